"""Device adjoint vs the float64 oracle adjoint (GPU).

The oracle's reverse step is pinned on the CPU (transpose identity and central
differences).  Here the HIP reverse sweep must reproduce it:

* adjoint fields after a reverse sweep and the design-window coefficient
  gradients dL/d(decay_E), dL/d(curl_E), for the same observer seeds, across
  CPML + ADE and PEC/PMC/periodic faces;
* the full autograd path (checkpointed forward, segment replay, reverse sweep,
  coefficient -> density chain) against the oracle chain.

Tolerances: adjoint fields 1e-5 of their peak (float32 vs float64 arithmetic of
a linear recurrence, as for the forward fields); gradients 1e-3.
"""

import numpy as np
import pytest
import torch

import witwin_maxwell_amd as mw
from witwin_maxwell_amd.fdtd.plan import COMPONENTS, compile_plan

from conftest import family_err, norm_err

pytestmark = pytest.mark.gpu

D = 5e-8


def _scene(boundary, cells, structures=(), sources=(), monitors=(), regions=()):
    dom = mw.Domain(tuple((0.0, n * D) for n in cells))
    scene = mw.Scene(domain=dom, grid=mw.GridSpec.uniform(D), boundary=boundary, device="cuda")
    for s in structures:
        scene.add_structure(s)
    for s in sources:
        scene.add_source(s)
    for m in monitors:
        scene.add_monitor(m)
    for r in regions:
        scene.add_material_region(r)
    return scene


def _design_case(bname, seed=0, cells=(14, 12, 16)):
    rng = np.random.default_rng(seed)
    density = torch.as_tensor(rng.uniform(0.1, 0.9, size=(4, 3, 3)), dtype=torch.float64)
    two_pi = 2 * np.pi
    structures = [mw.Structure(mw.Box((3 * D, 3 * D, 11 * D), (3 * D, 3 * D, 3 * D)), mw.Material(eps_r=2.2))]
    if bname == "plain_cpml":
        boundary = mw.BoundarySpec.pml(3)
    elif bname == "cpml_ade":
        boundary = mw.BoundarySpec.pml(3)
        structures.append(mw.Structure(mw.Box((10 * D, 8 * D, 4 * D), (3 * D, 3 * D, 3 * D)),
                                       mw.Material(eps_r=2.0, drude_poles=(mw.DrudePole(2e15 / two_pi, 1e14 / two_pi),),
                                                   lorentz_poles=(mw.LorentzPole(1.2, 3e15 / two_pi, 2e14 / two_pi),))))
    else:
        boundary = mw.BoundarySpec.faces(x="periodic", y="pec", z="pmc")
    src = mw.PointDipole((4.2 * D, 5.1 * D, 3.3 * D), "Ez", profile="ideal", source_time=mw.GaussianPulse(4e14, 1.5e14))
    mon = mw.PlaneMonitor("out", axis="z", position=12.0 * D, fields=("Ex", "Ey", "Hx", "Hy"), frequencies=(4e14, 5e14))
    pt = mw.PointMonitor("probe", (9.0 * D, 6.0 * D, 10.0 * D), fields=("Ez", "Hx"), frequencies=(4e14,))
    region = mw.MaterialRegion("design", mw.Box((7 * D, 6 * D, 8 * D), (4 * D, 3 * D, 3 * D)), density,
                               eps_bounds=(1.0, 5.0))
    scene = _scene(boundary, cells, structures, [src], [mon, pt], [region])
    return scene, density


@pytest.mark.parametrize("bname", ["cpml_ade", "periodic_pec_pmc"])
def test_reverse_sweep_matches_oracle(bname):
    from oracle.driver import OracleAdjoint, OracleRun
    from witwin_maxwell_amd.fdtd.adjoint import DeviceAdjoint, design_window
    from witwin_maxwell_amd.fdtd.runtime import DeviceRun
    print(f"case: reverse sweep vs oracle, {bname}", flush=True)
    scene, _ = _design_case(bname)
    steps = 120
    plan = compile_plan(scene, frequencies=(4e14, 5e14), time_steps=steps, courant=0.9)
    ref = OracleRun(plan)
    ref.trajectory = []
    ref.run(steps)
    rng = np.random.default_rng(11)
    g_planes = [(rng.standard_normal(a.shape), rng.standard_normal(a.shape)) for a in ref.plane_re]
    g_point = (rng.standard_normal(ref.point_re.shape), rng.standard_normal(ref.point_im.shape))
    radj = OracleAdjoint(ref)
    radj.run_reverse({"plane_re": [g[0] for g in g_planes], "plane_im": [g[1] for g in g_planes],
                      "point_re": g_point[0], "point_im": g_point[1]})

    window = design_window(plan)
    dev = DeviceRun(plan, "cuda")
    dev.run(steps)
    adj = DeviceAdjoint(dev, window, capacity=steps)
    adj.zero()
    t = lambda a: torch.as_tensor(a, dtype=torch.float32)
    adj.set_seeds(point=(t(g_point[0]), t(g_point[1])), planes=[(t(r), t(i)) for r, i in g_planes])
    from witwin_maxwell_amd.fdtd.adjoint import restore
    restore(dev, None, 0)
    adj.capture(0, steps)
    adj.reverse(0, steps)
    torch.cuda.synchronize()
    errs = family_err({c: adj.fields[c].double().cpu().numpy() for c in COMPONENTS[:3]},
                      {c: radj.ae[i] for i, c in enumerate(COMPONENTS[:3])})
    assert max(errs.values()) <= 1e-5, errs
    lo, hi = window
    box = tuple(slice(l, h) for l, h in zip(lo, hi))
    for ci in range(3):
        gd = adj.g_decay[ci].double().cpu().numpy()
        gc = adj.g_curl[ci].double().cpu().numpy()
        assert norm_err(gc, radj.g_curl[ci][box]) <= 1e-3, ("curl", ci)
        assert norm_err(gd, radj.g_decay[ci][box]) <= 1e-3, ("decay", ci)
    assert np.abs(radj.g_curl[0][box]).max() > 0


def _oracle_density_gradient(scene, density, plan_kw, loss_fn):
    """dL/d(density) from the float64 oracle: forward, torch loss on the
    assembled monitors, oracle reverse sweep, coefficient -> density chain."""
    from oracle.driver import OracleAdjoint, OracleRun
    from witwin_maxwell_amd.fdtd import materials as mat
    from witwin_maxwell_amd.fdtd.spectral import assemble_monitor_results
    from witwin_maxwell_amd.scene import prepare_scene
    plan = compile_plan(scene, **plan_kw)
    ref = OracleRun(plan)
    ref.trajectory = []
    ref.run(plan.time_steps)
    pr = torch.tensor(ref.point_re, requires_grad=True)
    pi = torch.tensor(ref.point_im, requires_grad=True)
    qr = [torch.tensor(a, requires_grad=True) for a in ref.plane_re]
    qi = [torch.tensor(a, requires_grad=True) for a in ref.plane_im]
    mons = assemble_monitor_results(plan.observers, pr, pi, qr, qi, plan.obs_window_sum)
    loss = loss_fn(mons)
    loss.backward()
    radj = OracleAdjoint(ref)
    g = lambda q: q.grad.numpy() if q.grad is not None else np.zeros(q.shape)
    radj.run_reverse({"point_re": g(pr), "point_im": g(pi), "plane_re": [g(q) for q in qr],
                      "plane_im": [g(q) for q in qi]})
    dens = density.detach().clone().requires_grad_(True)
    prepared = prepare_scene(scene)
    model = mat.NodeMaterialModel(prepared)
    eps_node = mat.node_permittivity(prepared, model, [dens])
    total = 0.0
    for ci, c in enumerate(("Ex", "Ey", "Ez")):
        decay, curl = mat.electric_coefficients(mat.average_node_to_component(eps_node, c),
                                                mat.average_node_to_component(model.sigma_e, c), plan.dt)
        total = total + (decay * torch.as_tensor(radj.g_decay[ci])).sum() + (curl * torch.as_tensor(radj.g_curl[ci])).sum()
    total.backward()
    return float(loss), dens.grad.numpy()


def _flux_loss(mons):
    e = mons["out"]
    return (e["Ex"].abs() ** 2).sum() + 0.5 * (e["Hy"].abs() ** 2).sum() * 1e5 + mons["probe"]["Ez"].real


@pytest.mark.parametrize("bname,stride", [("cpml_ade", 17), ("periodic_pec_pmc", None)])
def test_density_gradient_matches_oracle(bname, stride):
    from witwin_maxwell_amd.fdtd.adjoint import differentiable_run
    print(f"case: density gradient vs oracle, {bname}, stride {stride}", flush=True)
    scene, density = _design_case(bname, seed=1)
    plan_kw = dict(frequencies=(4e14, 5e14), time_steps=140, courant=0.9)
    ref_loss, ref_grad = _oracle_density_gradient(scene, density, plan_kw, _flux_loss)
    dens = density.detach().clone().requires_grad_(True)
    plan, mons, session = differentiable_run(scene, region_densities=[dens], device="cuda", stride=stride, **plan_kw)
    loss = _flux_loss(mons)
    loss.backward()
    assert abs(float(loss) - ref_loss) <= 1e-4 * abs(ref_loss)
    g = dens.grad.numpy()
    assert np.abs(ref_grad).max() > 0
    assert norm_err(g, ref_grad) <= 1e-3, (g.ravel()[:6], ref_grad.ravel()[:6])


def test_gpu_resident_filtered_projected_region_gradient():
    """A MaterialRegion whose density lives on the GPU, with non-unit bounds, a
    box filter and a tanh projection: the region compile (normalise, filter,
    project, sample, blend, coefficients) runs on the device in float64 and the
    density gradient of the checkpointed device adjoint equals the float64
    oracle chain's (the same chain on the CPU over the oracle adjoint, itself
    pinned by central differences in tests/test_adjoint_oracle.py)."""
    from witwin_maxwell_amd.fdtd.adjoint import differentiable_run
    print("case: GPU-resident filtered + projected region gradient", flush=True)
    scene, density = _design_case("plain_cpml", seed=2)
    r = scene.material_regions[0]

    def with_density(d):
        scene.material_regions[0] = mw.MaterialRegion(r.name, r.geometry, d, bounds=(0.05, 0.95),
                                                      eps_bounds=r.eps_bounds, filter_radius=1.5 * D,
                                                      projection_beta=4.0)
        return scene

    plan_kw = dict(frequencies=(4e14, 5e14), time_steps=140, courant=0.9)
    ref_loss, ref_grad = _oracle_density_gradient(with_density(density), density, plan_kw, _flux_loss)
    dens = density.detach().clone().to("cuda").requires_grad_(True)
    plan, mons, session = differentiable_run(with_density(dens), region_densities=[dens], device="cuda", stride=23,
                                             **plan_kw)
    loss = _flux_loss(mons)
    loss.backward()
    assert dens.grad is not None and dens.grad.is_cuda
    assert abs(float(loss) - ref_loss) <= 1e-4 * abs(ref_loss)
    g = dens.grad.double().cpu().numpy()
    assert np.abs(ref_grad).max() > 0
    assert norm_err(g, ref_grad) <= 1e-3, (g.ravel()[:6], ref_grad.ravel()[:6])
