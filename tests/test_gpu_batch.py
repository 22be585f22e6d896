"""Scene batches (GPU): several scenes of uneven sizes (one long enough in z for
an interior box) advanced by the batched
launches (direct and graph-replayed) give the same fields and monitors as each
scene run alone, and those match the float64 oracle."""

import numpy as np
import pytest
import torch

import witwin_maxwell_amd as mw
from witwin_maxwell_amd import configs
from witwin_maxwell_amd.fdtd.plan import COMPONENTS, compile_plan

from conftest import family_err

pytestmark = pytest.mark.gpu


def _scenes():
    out = []
    s1, k1 = configs.vacuum_dipole(cells=14, pml=5)
    out.append((s1, dict(k1, time_steps=120, full_field_dft=False)))
    s2, k2 = configs.dielectric_boxes(cells=20, pml=6, n_boxes=5, box_range=(3, 8), seed=7)
    out.append((s2, dict(k2, time_steps=120)))
    dx = configs.DX
    s3 = mw.Scene(domain=mw.Domain(((-8 * dx, 8 * dx), (-6 * dx, 6 * dx), (-11 * dx, 11 * dx))),
                  grid=mw.GridSpec.uniform(dx), boundary=mw.BoundarySpec.pml(5))
    s3.add_structure(mw.Structure(mw.Box((0, 0, 2 * dx), (5 * dx, 4 * dx, 3 * dx)), mw.Material(eps_r=3.0)))
    s3.add_source(mw.PlaneWave((0, 0, 1), (0, 1, 0), source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4),
                               injection=mw.TFSF.slab(axis="z", bounds=(-8 * dx, 8 * dx))))
    s3.add_monitor(mw.FluxMonitor("t", axis="z", position=9 * dx, frequencies=(configs.F0,)))
    out.append((s3, {"frequencies": (configs.F0,), "courant": 0.95, "time_steps": 120}))
    # long in z: an interior box, with z-edge tiles or z shell slabs beside it
    s4, k4 = configs.dielectric_boxes(cells=16, pml=6, n_boxes=6, box_range=(3, 8), seed=5)
    s4.domain = mw.Domain(((-8 * dx, 8 * dx), (-9 * dx, 9 * dx), (-44 * dx, 44 * dx)))
    out.append((s4, dict(k4, time_steps=120)))
    return out


@pytest.mark.parametrize("graph_steps,xchunk,z_tiling", [(1, 0, 0), (16, 0, 1), (8, 12, 2)])
def test_batch_matches_single_scene_runs(graph_steps, xchunk, z_tiling):
    from oracle.driver import OracleRun
    from witwin_maxwell_amd.fdtd.batch import BatchRun
    from witwin_maxwell_amd.fdtd.runtime import DeviceRun
    print(f"case: batch of 4 scenes, graph_steps={graph_steps}, xchunk={xchunk}, z_tiling={z_tiling}", flush=True)
    scenes = _scenes()
    plans = [compile_plan(s, **kw) for s, kw in scenes]
    singles = [DeviceRun(p, "cuda").set_tiling(0, 0, z_tiling).run() for p in plans]
    batch = BatchRun([DeviceRun(p, "cuda").set_tiling(0, 0, z_tiling) for p in plans], xchunk=xchunk)
    batch.run(50, graph_steps=graph_steps)
    batch.run(graph_steps=graph_steps)
    torch.cuda.synchronize()
    for p, single, brun in zip(plans, singles, batch.runs):
        ref = OracleRun(p).run()
        scale_e = max(ref.peak[c] for c in COMPONENTS[:3])
        scale_h = max(ref.peak[c] for c in COMPONENTS[3:])
        for fam, scale in ((COMPONENTS[:3], scale_e), (COMPONENTS[3:], scale_h)):
            eb = family_err({c: brun.fields[c].double().cpu().numpy() for c in fam},
                            {c: ref.fields[c] for c in fam}, scale=scale)
            assert max(eb.values()) <= 1e-5, eb
            # the batch runs the same kernels on the same data as a lone scene
            es = family_err({c: brun.fields[c].double().cpu().numpy() for c in fam},
                            {c: single.fields[c].double().cpu().numpy() for c in fam}, scale=scale)
            assert max(es.values()) <= 1e-6, es
        mb, ms = brun.monitor_results(), single.monitor_results()
        for name in ms:
            for comp, v in ms[name]["components"].items():
                a = v["data"] if isinstance(v, dict) else v
                b = mb[name]["components"][comp]
                b = b["data"] if isinstance(b, dict) else b
                assert np.max(np.abs(np.asarray(a) - np.asarray(b))) <= 1e-6 * max(np.max(np.abs(a)), 1e-30)
    batch.close()


@pytest.mark.skip(reason="batched reverse kernels switched off: their first device run ended in a GPU memory fault "
                         "whose cause is not established (CPU AddressSanitizer emulation of the same path is clean)")
def test_batched_gradients_match_oracle():
    """Three design scenes of different sizes differentiated in one batch by the
    batched reverse kernels: each scene's density gradient equals its own float64
    oracle gradient."""
    import test_gpu_adjoint as ta
    from witwin_maxwell_amd.fdtd.batch import batched_differentiable_run
    print("case: batched forward+adjoint, 3 design scenes, stride 19", flush=True)
    cases = []
    for seed, cells in ((3, (14, 12, 16)), (4, (16, 13, 18)), (5, (14, 15, 16))):
        scene, density = ta._design_case("plain_cpml", seed=seed, cells=cells)
        cases.append((scene, density))
    plan_kw = dict(frequencies=(4e14, 5e14), time_steps=120, courant=0.9)
    dens = [d.detach().clone().requires_grad_(True) for _, d in cases]
    scenes = []
    for (scene, _), d in zip(cases, dens):
        r = scene.material_regions[0]
        scene.material_regions[0] = mw.MaterialRegion(r.name, r.geometry, d, eps_bounds=r.eps_bounds)
        scenes.append(scene)
    plans, mons, session = batched_differentiable_run(scenes, plan_kw, stride=19, reverse="kernels")
    loss = sum(ta._flux_loss(m) for m in mons)
    loss.backward()
    for (scene, density), d in zip(cases, dens):
        ref_loss, ref_grad = ta._oracle_density_gradient(scene, density, plan_kw, ta._flux_loss)
        assert np.abs(ref_grad).max() > 0
        from conftest import norm_err
        assert norm_err(d.grad.numpy(), ref_grad) <= 1e-3


def test_batched_forward_stream_reverse_gradients_match_oracle():
    """Four design scenes of different sizes: the forward advances all of them in
    the batched launches (graph-replayed), each scene's reverse sweep runs the
    single-scene entry points on one of two streams (so two scenes share each
    stream and sweeps of different scenes overlap); every scene's density
    gradient equals its own float64 oracle gradient."""
    import test_gpu_adjoint as ta
    from conftest import norm_err
    from witwin_maxwell_amd.fdtd.batch import batched_differentiable_run
    print("case: batched forward + per-scene stream reverse, 4 design scenes, stride 23, 2 streams", flush=True)
    cases = []
    for seed, cells in ((6, (14, 14, 16)), (7, (16, 12, 17)), (8, (12, 15, 18)), (9, (15, 13, 16))):
        scene, density = ta._design_case("plain_cpml", seed=seed, cells=cells)
        cases.append((scene, density))
    plan_kw = dict(frequencies=(4e14, 5e14), time_steps=110, courant=0.9)
    dens = [d.detach().clone().requires_grad_(True) for _, d in cases]
    scenes = []
    for (scene, _), d in zip(cases, dens):
        r = scene.material_regions[0]
        scene.material_regions[0] = mw.MaterialRegion(r.name, r.geometry, d, eps_bounds=r.eps_bounds)
        scenes.append(scene)
    plans, mons, session = batched_differentiable_run(scenes, plan_kw, stride=23, graph_steps=8, streams=2)
    loss = sum(ta._flux_loss(m) for m in mons)
    loss.backward()
    assert len(session.scene_done_s) == 4 and all(t > 0 for t in session.scene_done_s)
    for (scene, density), d in zip(cases, dens):
        ref_loss, ref_grad = ta._oracle_density_gradient(scene, density, plan_kw, ta._flux_loss)
        assert np.abs(ref_grad).max() > 0
        assert norm_err(d.grad.numpy(), ref_grad) <= 1e-3
