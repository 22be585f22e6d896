"""Pins the float64 reverse step of the oracle (CPU only).

* Transpose identity: for one forward step F (linear in the state once sources
  are removed) and random states x, y, <F x, y> = <x, F^T y> to round-off.  The
  forward step is the physics-pinned oracle step, so a reverse step that passes
  is its exact transpose: every stencil sign, index, boundary mode, CPML psi
  term, conductivity decay and Debye/Drude/Lorentz pole is covered (the pattern
  of tests/fdtd/cuda/test_cuda_adjoint_native.py:17-69 in the reference).
* Gradient: the adjoint dL/d(eps_node) through the coefficient map equals central
  differences of the forward oracle (tests/gradients/test_fdtd_adjoint_rigorous.py
  :40-59 pattern), for a loss built from the monitor DFT sums.
"""

import numpy as np
import pytest
import torch

import witwin_maxwell_amd as mw
from witwin_maxwell_amd.fdtd.plan import compile_plan

from conftest import norm_err

D = 5e-8


def _scene(boundary, structures=(), sources=(), monitors=(), cells=(6, 7, 8)):
    dom = mw.Domain(tuple((0.0, n * D) for n in cells))
    scene = mw.Scene(domain=dom, grid=mw.GridSpec.uniform(D), boundary=boundary, device="cpu")
    for s in structures:
        scene.add_structure(s)
    for s in sources:
        scene.add_source(s)
    for m in monitors:
        scene.add_monitor(m)
    return scene


def _lossy_dispersive_structures():
    two_pi = 2 * np.pi
    return [
        mw.Structure(mw.Box((2 * D, 3 * D, 4 * D), (4 * D, 4 * D, 6 * D)), mw.Material(eps_r=3.0, sigma_e=2e5)),
        mw.Structure(mw.Box((4 * D, 4 * D, 3 * D), (3 * D, 5 * D, 4 * D)),
                     mw.Material(eps_r=2.0, drude_poles=(mw.DrudePole(2e15 / two_pi, 1e14 / two_pi),),
                                 lorentz_poles=(mw.LorentzPole(1.2, 3e15 / two_pi, 2e14 / two_pi),))),
        mw.Structure(mw.Box((1 * D, 2 * D, 6 * D), (3 * D, 3 * D, 3 * D)),
                     mw.Material(eps_r=1.5, debye_poles=(mw.DebyePole(0.8, 2e-15),))),
    ]


BOUNDARIES = {
    "cpml": lambda: mw.BoundarySpec.pml(2),
    "periodic_pec_pmc": lambda: mw.BoundarySpec.faces(x="periodic", y="pec", z="pmc"),
    "mixed_faces": lambda: mw.BoundarySpec.faces(default="none", num_layers=2, x=("pml", "pmc"),
                                                 y=("pec", "none"), z=("pmc", "pml")),
}


def _state_arrays(run):
    arrs = [run.fields[c] for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")] + list(run.psi)
    for c in ("Ex", "Ey", "Ez"):
        for p in run.ade[c]["poles"]:
            arrs += [p["pol"], p["cur"]]
    return arrs


def _zero_unreachable(plan, run):
    """Zero the state the forward dynamics never populate: psi outside the
    absorbing layers (b = 1, c = 0 there, so psi stays zero) and pole state where
    the pole's drive is zero.  The oracle's step skips that state, so the
    transpose identity is taken over the reachable states."""
    axis_of = "yzxzxyyzxzxy"
    for q, psi in enumerate(run.psi):
        ax = "xyz".index(axis_of[q])
        if not plan.uses_cpml:
            psi[...] = 0.0
            continue
        prof = plan.cpml["xyz"[ax]]
        b, c = (prof["b_e"], prof["c_e"]) if q < 6 else (prof["b_h"], prof["c_h"])
        n = psi.shape[ax]
        live = ~((np.asarray(b)[:n] == 1.0) & (np.asarray(c)[:n] == 0.0))
        shape = [1, 1, 1]
        shape[ax] = n
        psi *= live.reshape(shape)
    for c in ("Ex", "Ey", "Ez"):
        for p in run.ade[c]["poles"]:
            acts = p["drive"] != 0
            p["pol"] *= acts
            p["cur"] *= acts


def _adj_arrays(adj):
    arrs = list(adj.ae) + list(adj.ah) + list(adj.apsi)
    for c in ("Ex", "Ey", "Ez"):
        for p in adj.adj_poles[c]:
            arrs += [p["pol"], p["cur"]]
    return arrs


@pytest.mark.parametrize("bname", sorted(BOUNDARIES))
def test_reverse_step_is_exact_transpose(oracle_lib, bname):
    from oracle.driver import OracleAdjoint, OracleRun
    scene = _scene(BOUNDARIES[bname](), _lossy_dispersive_structures())
    plan = compile_plan(scene, time_steps=4, courant=0.9)
    assert plan.dispersive and any(np.any(plan.coef[c][0] != 1.0) for c in ("Ex", "Ey", "Ez"))
    rng = np.random.default_rng(3)
    run = OracleRun(plan)
    xs = _state_arrays(run)
    for a in xs:
        a[...] = rng.standard_normal(a.shape)
    _zero_unreachable(plan, run)
    x0 = [a.copy() for a in xs]
    run.step()
    fx = [a.copy() for a in xs]
    adj = OracleAdjoint(run)
    ys = _adj_arrays(adj)
    assert len(ys) == len(xs)
    for a in ys:
        a[...] = rng.standard_normal(a.shape)
    y0 = [a.copy() for a in ys]
    adj.reverse_step(0, gradients=False)
    lhs = sum(float(np.vdot(f, y)) for f, y in zip(fx, y0))
    rhs = sum(float(np.vdot(x, a)) for x, a in zip(x0, ys))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1.0), (lhs, rhs)


def _design_scene(density, boundary):
    src = mw.PointDipole((1.6 * D, 2.1 * D, 1.7 * D), "Ez", profile="ideal",
                         source_time=mw.GaussianPulse(4e14, 1e14))
    mon = mw.PlaneMonitor("out", axis="z", position=6.0 * D, fields=("Ex", "Ey", "Hx", "Hy"), frequencies=(4e14, 5e14))
    scene = _scene(boundary, [mw.Structure(mw.Box((4 * D, 3.5 * D, 4 * D), (2 * D, 2 * D, 2 * D)),
                                           mw.Material(eps_r=2.5))], [src], [mon], cells=(7, 7, 8))
    scene.add_material_region(mw.MaterialRegion("design", mw.Box((3.5 * D, 3.5 * D, 4.0 * D), (3 * D, 3 * D, 2 * D)),
                                                density, eps_bounds=(1.0, 6.0)))
    return scene


def _loss_from_run(run, wre, wim):
    re = run.plane_re
    im = run.plane_im
    return sum(float(np.sum(a * b)) for a, b in zip(re, wre)) + sum(float(np.sum(a * b)) for a, b in zip(im, wim))


@pytest.mark.parametrize("bname", ["cpml", "periodic_pec_pmc"])
def test_oracle_gradient_matches_central_differences(oracle_lib, bname):
    """dL/d(node eps) by the oracle adjoint vs central differences of the oracle
    forward, with L linear in the plane DFT sums (random weights)."""
    from oracle.driver import OracleAdjoint, OracleRun
    from witwin_maxwell_amd.fdtd import materials as mat
    rng = np.random.default_rng(7)
    density = torch.as_tensor(rng.uniform(0.2, 0.8, size=(3, 3, 2)), dtype=torch.float64)
    scene = _design_scene(density, BOUNDARIES[bname]())
    steps = 60
    plan = compile_plan(scene, frequencies=(4e14, 5e14), time_steps=steps, courant=0.9)
    run = OracleRun(plan)
    run.trajectory = []
    run.run(steps)
    wre = [rng.standard_normal(a.shape) for a in run.plane_re]
    wim = [rng.standard_normal(a.shape) for a in run.plane_im]
    adj = OracleAdjoint(run)
    adj.run_reverse({"plane_re": wre, "plane_im": wim})

    # chain dL/d(decay, curl) -> node eps through the coefficient formulas
    eps_node = torch.as_tensor(plan.node_eps, dtype=torch.float64).requires_grad_(True)
    total = 0.0
    for ci, comp in enumerate(("Ex", "Ey", "Ez")):
        eps = mat.average_node_to_component(eps_node, comp)
        decay, curl = mat.electric_coefficients(eps, mat.average_node_to_component(plan.material_model.sigma_e,
                                                                                       comp), plan.dt)
        total = total + (decay * torch.as_tensor(adj.g_decay[ci])).sum() + (curl * torch.as_tensor(adj.g_curl[ci])).sum()
    (g_node,) = torch.autograd.grad(total, eps_node)
    g_node = g_node.numpy()

    # central differences on a few nodes inside the design region and the box
    picks = [(4, 4, 5), (5, 4, 4), (4, 5, 5), (6, 4, 5), (3, 3, 4)]
    h = 1e-5
    for idx in picks:
        vals = []
        for sgn in (1, -1):
            eps = plan.node_eps.copy()
            eps[idx] += sgn * h
            p2 = compile_plan(scene, frequencies=(4e14, 5e14), time_steps=steps, courant=0.9, node_eps_override=eps)
            r2 = OracleRun(p2).run(steps)
            vals.append(_loss_from_run(r2, wre, wim))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g_node[idx]) <= 1e-6 * max(abs(fd), abs(g_node).max()), (idx, fd, g_node[idx])
    assert np.abs(g_node).max() > 0


def test_filtered_projected_density_gradient_matches_central_differences(oracle_lib):
    """dL/d(density) through normalisation to non-unit bounds, the box filter, the
    tanh projection, trilinear sampling and the region blend (the torch chain of
    fdtd/materials.py, compiler/materials.py:1310-1470 in the reference) on top of
    the oracle adjoint, vs central differences of the oracle forward in the raw
    density: pins the region compile's derivative, filter and projection included."""
    from oracle.driver import OracleAdjoint, OracleRun
    from witwin_maxwell_amd.fdtd import materials as mat
    from witwin_maxwell_amd.scene import prepare_scene
    rng = np.random.default_rng(9)
    raw = torch.as_tensor(rng.uniform(0.15, 0.85, size=(5, 4, 3)), dtype=torch.float64)

    def scene_with(dens):
        scene = _design_scene(dens, BOUNDARIES["cpml"]())
        r = scene.material_regions[0]
        scene.material_regions[0] = mw.MaterialRegion(r.name, r.geometry, dens, bounds=(0.1, 0.9),
                                                      eps_bounds=r.eps_bounds, filter_radius=1.2 * D,
                                                      projection_beta=3.0)
        return scene

    steps = 50
    kw = dict(frequencies=(4e14, 5e14), time_steps=steps, courant=0.9)
    scene = scene_with(raw)
    plan = compile_plan(scene, **kw)
    run = OracleRun(plan)
    run.trajectory = []
    run.run(steps)
    wre = [rng.standard_normal(a.shape) for a in run.plane_re]
    wim = [rng.standard_normal(a.shape) for a in run.plane_im]
    adj = OracleAdjoint(run)
    adj.run_reverse({"plane_re": wre, "plane_im": wim})
    dens = raw.clone().requires_grad_(True)
    prepared = prepare_scene(scene)
    model = mat.NodeMaterialModel(prepared)
    eps_node = mat.node_permittivity(prepared, model, [dens])
    total = 0.0
    for ci, comp in enumerate(("Ex", "Ey", "Ez")):
        decay, curl = mat.electric_coefficients(mat.average_node_to_component(eps_node, comp),
                                                mat.average_node_to_component(model.sigma_e, comp), plan.dt)
        total = total + (decay * torch.as_tensor(adj.g_decay[ci])).sum() + (curl * torch.as_tensor(adj.g_curl[ci])).sum()
    (g,) = torch.autograd.grad(total, dens)
    g = g.numpy()
    assert np.abs(g).max() > 0
    h = 1e-6
    for idx in [(0, 0, 0), (2, 1, 1), (4, 3, 2), (1, 2, 0), (3, 0, 2)]:
        vals = []
        for sgn in (1, -1):
            d2 = raw.clone()
            d2[idx] += sgn * h
            r2 = OracleRun(compile_plan(scene_with(d2), **kw)).run(steps)
            vals.append(_loss_from_run(r2, wre, wim))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-5 * np.abs(g).max(), (idx, fd, g[idx])
