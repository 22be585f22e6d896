"""Simulation.fdtd(...).run() -> Result on the device (GPU): the front end runs
the same discretised problem as a direct plan (checked against the float64
oracle), and a trainable MaterialRegion makes result.monitor() payloads
differentiable with oracle-exact gradients."""

import numpy as np
import pytest
import torch

import witwin_maxwell_amd as mw
from witwin_maxwell_amd import configs
from witwin_maxwell_amd.fdtd.plan import compile_plan

from conftest import family_err, norm_err

pytestmark = pytest.mark.gpu

D = 5e-8


def test_simulation_forward_matches_oracle():
    from oracle.driver import OracleRun
    print("case: Simulation.fdtd vacuum dipole 16 cells, 3 frequencies, 180 steps", flush=True)
    scene, kw = configs.vacuum_dipole(cells=16, pml=6)
    scene.add_monitor(mw.PointMonitor("probe", (2 * configs.DX, -configs.DX, configs.DX), fields=("Ez", "Hx")))
    sim = mw.Simulation.fdtd(scene, frequencies=kw["frequencies"], run_time=mw.TimeConfig(time_steps=180),
                             spectral_sampler=mw.SpectralSampler(window="none"))
    res = sim.run()
    assert res.solver_stats["time_steps"] == 180 and res.solver_stats["full_field_dft"]
    plan = compile_plan(scene, frequencies=kw["frequencies"], time_steps=180, full_field_dft=True, window="none",
                        dt=res.solver_stats["dt"])
    ref = OracleRun(plan).run()
    rm = ref.monitor_results()["probe"]
    dm = res.monitor("probe")
    errs = family_err({c: dm["components"][c].cpu().numpy() for c in ("Ez",)},
                      {c: rm["components"][c] for c in ("Ez",)})
    assert max(errs.values()) <= 1e-5, errs
    # full-field DFT payload: (F, *shape) complex, E components
    ez = res.tensor("Ez", freq_index=1)
    assert ez.shape == plan.field_shape("Ez") and ez.is_complex()


def test_simulation_trainable_region_backward():
    import test_gpu_adjoint as ta
    print("case: Simulation.fdtd trainable design region, loss.backward vs oracle", flush=True)
    scene, density = ta._design_case("cpml_ade", seed=2)
    dens = density.detach().clone().requires_grad_(True)
    region = scene.material_regions[0]
    scene.material_regions[0] = mw.MaterialRegion(region.name, region.geometry, dens, eps_bounds=region.eps_bounds)
    sim = mw.Simulation.fdtd(scene, frequencies=(4e14, 5e14), run_time=mw.TimeConfig(time_steps=140),
                             spectral_sampler=mw.SpectralSampler(window="none"), adjoint_checkpoint_stride=23)
    res = sim.run()
    loss = ta._flux_loss(res.monitors)
    loss.backward()
    plan_kw = dict(frequencies=(4e14, 5e14), time_steps=140, dt=res.solver_stats["dt"], window="none",
                   full_field_dft=True)
    scene2, _ = ta._design_case("cpml_ade", seed=2)
    ref_loss, ref_grad = ta._oracle_density_gradient(scene2, density, plan_kw, ta._flux_loss)
    assert abs(float(loss) - ref_loss) <= 1e-4 * abs(ref_loss)
    assert norm_err(dens.grad.numpy(), ref_grad) <= 1e-3
