"""The five benchmark configurations at their stated sizes on the HIP path (GPU).

* vacuum dipole (64^3 cells, CPML 10, 500 steps, 3-frequency full-field DFT,
  point monitor): fields mid-run and at the end, psi, DFT and monitor sums
  against the float64 oracle, the whole configuration;
* dielectric boxes (256^3 cells, 281^3 nodes): the full grid against the
  oracle for 450 steps, by which the pulse has crossed the lower flux plane and
  is crossing the upper one;
* dispersive spheres (384x384x256 cells, 409x409x281 nodes, Drude + 2 Lorentz
  poles, TFSF slab): the full grid against the oracle for the first steps of the
  injection (the resolved-plasma time step is ~1/300 period, so the stated
  4000-step run is beyond an oracle run); the pulse starts at once and an extra
  flux plane 6 cells past the TFSF face sees it;
* design-region gradient (225x225x145 nodes, 64x64x32 design voxels, 3000
  steps, checkpoints every 100): the adjoint gradient against central finite
  differences of device forward runs, along a random direction over every voxel
  and on 16 seeded voxels (h = 1e-3 in theta);
* scene batch (32 scenes 64^3 .. 192x192x128 cells): every scene of the batch
  against the same scene run alone, and the smallest and largest against the
  oracle.

Tolerance: 1e-5 of the reference peak for fields, psi, pole state, DFT and
monitor sums; 1e-3 for gradients.
"""

import numpy as np
import pytest
import torch

from witwin_maxwell_amd import configs
from witwin_maxwell_amd.fdtd.plan import COMPONENTS, compile_plan

from conftest import family_err
from test_gpu_parity import TOL, _check_fields, _check_monitors, _run_both

pytestmark = pytest.mark.gpu


def test_vacuum_dipole_full_configuration():
    print("case: vacuum dipole 64^3 cells CPML 10, 500 steps, full-field DFT x3", flush=True)
    scene, kw = configs.vacuum_dipole()
    plan, ref, dev = _run_both(scene, kw, 500, True)
    assert plan.shape == (85, 85, 85)
    _check_fields(ref, dev)
    _check_monitors(ref, dev)
    for f in range(len(plan.dft_frequencies)):
        d = {c + p: (dev.dft_re if p == "re" else dev.dft_im)[c][f].double().cpu().numpy()
             for c in ("Ex", "Ey", "Ez") for p in ("re", "im")}
        r = {c + p: ref.dft[c][0 if p == "re" else 1][f] for c in ("Ex", "Ey", "Ez") for p in ("re", "im")}
        errs = family_err(d, r)
        assert max(errs.values()) <= TOL, errs


def test_dielectric_boxes_full_grid():
    print("case: dielectric boxes 256^3 cells (281^3 nodes), 40 boxes, 450 steps", flush=True)
    scene, kw = configs.dielectric_boxes()
    plan, ref, dev = _run_both(scene, kw, 450, True)
    assert plan.shape == (281, 281, 281)
    _check_fields(ref, dev)
    _check_monitors(ref, dev)


def test_dispersive_spheres_full_grid():
    import witwin_maxwell_amd as mw
    print("case: dispersive spheres 384x384x256 cells (409x409x281 nodes), TFSF, 150 steps", flush=True)
    scene, kw = configs.dispersive_spheres(pulse_delay=2e-15)
    half_z = 0.5 * 256 * configs.DX
    freqs = kw["frequencies"]
    scene.add_monitor(mw.FluxMonitor("flux_face", axis="z", position=-half_z + 10 * configs.DX, frequencies=freqs))
    plan, ref, dev = _run_both(scene, kw, 150, True)
    assert plan.shape == (409, 409, 281) and plan.dispersive and plan.aux is not None
    _check_fields(ref, dev)
    d, r = {}, {}
    for c in ("Ex", "Ey", "Ez"):
        for i, (p, st) in enumerate(zip(ref.ade[c]["poles"], dev.ade_state[c])):
            d[(c, i)] = st["current"].double().cpu().numpy()
            r[(c, i)] = p["cur"]
    errs = family_err(d, r)
    assert max(errs.values()) <= TOL, errs
    mr, md = ref.monitor_results(), dev.monitor_results()
    face = {k: v for k, v in mr.items() if k == "flux_face"}
    assert face
    for name in face:
        for c, v in mr[name]["components"].items():
            a = np.asarray(v["data"])
            b = md[name]["components"][c]["data"]
            b = b.detach().cpu().numpy() if hasattr(b, "detach") else np.asarray(b)
            assert np.max(np.abs(a)) > 0, (name, c)
            assert np.max(np.abs(b - a)) <= TOL * np.max(np.abs(a)), (name, c)


def _objective(monitors):
    """|E|^2 of the central 40x40 of the objective plane at f0 (the design-region
    configuration's loss)."""
    e = monitors["objective"]
    loss = 0.0
    for c in ("Ex", "Ey"):
        d = e[c]
        a0, b0 = (d.shape[-2] - 40) // 2, (d.shape[-1] - 40) // 2
        loss = loss + (d[..., a0:a0 + 40, b0:b0 + 40].abs() ** 2).sum()
    return loss


def test_design_region_gradient_finite_differences():
    """Adjoint gradient of the full design-region configuration against central
    differences of device forward runs.  Each difference perturbs theta, recomputes
    the design window's E coefficients with the float64 formulas and writes them
    into the device run, whose forward run then gives the perturbed loss.

    * Directional: theta +- h d for a seeded random sign vector d over all
      131072 voxels, h = 1e-2: (L+ - L-) / 2h against grad . d, within 1e-3.
      The difference moves the loss by ~2e-3 of itself, far above the float32
      resolution of a forward run (~1e-7 of the loss), so it resolves the
      gradient to ~1e-4.
    * Per voxel: 16 seeded voxels among those whose gradient is within 1/10 of
      the largest, h = 1e-3.  One voxel moves the loss by only ~5e-7 of itself,
      a few float32 resolutions, so each difference carries ~10% noise: the
      check is the sign and 0.2 of the largest gradient."""
    import witwin_maxwell_amd as mw
    from witwin_maxwell_amd.fdtd import materials as mat
    from witwin_maxwell_amd.fdtd.adjoint import differentiable_run, window_coefficients
    from witwin_maxwell_amd.scene import prepare_scene
    print("case: design region 200x200x120 cells, 3000 steps, stride 100, directional FD + FD on 16 voxels", flush=True)
    scene, kw = configs.design_region()
    theta = kw.pop("theta").clone().requires_grad_(True)
    region = scene.material_regions[0]
    scene.material_regions[0] = mw.MaterialRegion(region.name, region.geometry, torch.sigmoid(theta),
                                                  eps_bounds=region.eps_bounds)
    plan, mons, session = differentiable_run(scene, stride=100, **kw)
    assert plan.shape == (225, 225, 145) and plan.time_steps == 3000
    loss = _objective(mons)
    loss.backward()
    grad = theta.grad.detach().numpy()
    assert np.isfinite(grad).all() and np.abs(grad).max() > 0
    run = session.run
    lo, hi = session.window
    box = tuple(slice(l, h) for l, h in zip(lo, hi))
    prepared = prepare_scene(scene)
    model = mat.NodeMaterialModel(prepared)

    def forward_loss(th):
        eps_node = mat.node_permittivity(prepared, model, [torch.sigmoid(th)])
        coefs = window_coefficients(plan, eps_node, lo, hi)
        for ci, c in enumerate(("Ex", "Ey", "Ez")):
            run.set_coefficients(c, "curl", box, coefs[ci][1].detach().numpy())
        run.reset()
        run.run()
        torch.cuda.synchronize()
        e = run.monitor_results()["objective"]
        total = 0.0
        for c in ("Ex", "Ey"):
            d = np.asarray(e[c])  # the plane-aligned component, as the loss above reads it
            a0, b0 = (d.shape[-2] - 40) // 2, (d.shape[-1] - 40) // 2
            total += float((np.abs(d[..., a0:a0 + 40, b0:b0 + 40]) ** 2).sum())
        return total

    th0 = theta.detach().clone()
    base_loss = forward_loss(th0)
    assert abs(base_loss - float(loss)) <= 1e-5 * abs(base_loss), (base_loss, float(loss))
    rng = np.random.default_rng(0)
    # directional derivative over every voxel
    d = torch.as_tensor(rng.choice([-1.0, 1.0], size=grad.shape), dtype=th0.dtype)
    h_dir = 1e-2
    fd_dir = (forward_loss(th0 + h_dir * d) - forward_loss(th0 - h_dir * d)) / (2 * h_dir)
    ad_dir = float((grad * d.numpy()).sum())
    err_dir = abs(fd_dir - ad_dir) / abs(ad_dir)
    print(f"directional: adjoint {ad_dir:.6e}, finite difference {fd_dir:.6e}, relative error {err_dir:.2e}",
          flush=True)
    # single voxels
    strong = np.flatnonzero(np.abs(grad).ravel() >= 0.1 * np.abs(grad).max())
    voxels = rng.choice(strong, size=16, replace=strong.size < 16)
    h = 1e-3
    fd, ad = [], []
    for v in voxels:
        idx = np.unravel_index(v, grad.shape)
        tp, tm = th0.clone(), th0.clone()
        tp[idx] += h
        tm[idx] -= h
        fd.append((forward_loss(tp) - forward_loss(tm)) / (2 * h))
        ad.append(grad[idx])
    fd, ad = np.array(fd), np.array(ad)
    err = np.max(np.abs(fd - ad)) / np.max(np.abs(ad))
    print("adjoint:", ad, "\nfinite differences:", fd, f"\nmax error / max |gradient| = {err:.2e}", flush=True)
    assert err_dir <= 1e-3, (fd_dir, ad_dir)
    assert np.all(np.sign(fd) == np.sign(ad)), (fd, ad)
    assert err <= 0.2, err


def test_scene_batch_at_stated_sizes():
    """All 32 scenes of the batch configuration (their stated sizes) advanced by
    the batched launches for 100 steps: each equals the same scene run alone, and
    the smallest and the largest scene equal their float64 oracle runs."""
    from oracle.driver import OracleRun
    from witwin_maxwell_amd.fdtd.batch import BatchRun
    from witwin_maxwell_amd.fdtd.runtime import DeviceRun
    print("case: 32-scene batch at stated sizes, 100 steps", flush=True)
    scs, kw = configs.batched_scenes(steps=100)
    plans = [compile_plan(sc, region_densities=[r.density.detach() for r in sc.material_regions], **kw)
             for sc, _ in scs]
    sizes = [int(np.prod(p.shape)) for p in plans]
    batch = BatchRun([DeviceRun(p, "cuda") for p in plans])
    batch.run(graph_steps=16)
    torch.cuda.synchronize()
    check = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    for i, (p, brun) in enumerate(zip(plans, batch.runs)):
        single = DeviceRun(p, "cuda").run()
        torch.cuda.synchronize()
        for fam in (COMPONENTS[:3], COMPONENTS[3:]):
            scale = max(float(single.fields[c].abs().max()) for c in fam)
            es = family_err({c: brun.fields[c].double().cpu().numpy() for c in fam},
                            {c: single.fields[c].double().cpu().numpy() for c in fam}, scale=scale)
            assert max(es.values()) <= 1e-6, (i, es)
        if i in check:
            ref = OracleRun(p).run()
            _check_fields(ref, brun)
            _check_monitors(ref, brun)
        del single
    batch.close()
