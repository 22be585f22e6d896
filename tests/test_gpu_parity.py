"""HIP runtime vs the float64 oracle on the same seeded plans (normalised max error).

Tolerance: fields, psi, ADE state and DFT/monitor sums <= 1e-5 of the reference
field's peak (FP32 device arithmetic vs FP64 oracle).
"""

import numpy as np
import pytest
import torch

import witwin_maxwell_amd as mw
from witwin_maxwell_amd import configs
from witwin_maxwell_amd.fdtd.cpml import PSI_NAMES
from witwin_maxwell_amd.fdtd.plan import COMPONENTS, compile_plan

from conftest import family_err, norm_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


FUSED = [True, False]


def _run_both(scene, kw, steps, fused=None, tiling=None, on_mid=None):
    """Run oracle and device side by side; fields are also compared at mid-run,
    while the pulse is still inside the grid.  ``tiling`` = (xchunk, shell_chunk,
    z_tiling) of the single-pass step; ``on_mid(plan, ref, dev)`` runs at mid-run."""
    from oracle.driver import OracleRun
    from witwin_maxwell_amd.fdtd.runtime import DeviceRun
    kw = dict(kw)
    kw.pop("theta", None)
    kw["time_steps"] = steps
    plan = compile_plan(scene, **kw)
    ref = OracleRun(plan)
    dev = DeviceRun(plan, "cuda", fused=fused)
    if tiling is not None:
        dev.set_tiling(*tiling)
    mid = steps // 3
    ref.run(mid)
    dev.run(mid)
    torch.cuda.synchronize()
    _check_fields(ref, dev)
    _check_psi(plan, ref, dev)
    if on_mid is not None:
        on_mid(plan, ref, dev)
    ref.run(steps - mid)
    dev.run()
    torch.cuda.synchronize()
    return plan, ref, dev


def _check_fields(ref, dev, tol=TOL):
    """Family error normalised by the family's peak over the run so far: a
    component that has decayed (pulse gone, PML-absorbed) is judged at the scale
    float32 could represent it, max(|x|) over time times ~1e-7 per step."""
    errs = {}
    for fam in (COMPONENTS[:3], COMPONENTS[3:]):
        scale = max(ref.peak[c] for c in fam)
        errs.update(family_err({c: dev.fields[c].double().cpu().numpy() for c in fam},
                               {c: ref.fields[c] for c in fam}, scale=scale))
    assert max(errs.values()) <= tol, errs
    return errs


def _check_psi(plan, ref, dev, tol=TOL):
    """psi compared mid-run, while the pulse is inside the absorbing layers
    (after it has been absorbed psi decays towards round-off)."""
    if not plan.uses_cpml:
        return
    for fam in (range(0, 6), range(6, 12)):
        errs = family_err({PSI_NAMES[q]: dev.psi_dense(PSI_NAMES[q]).double().cpu().numpy() for q in fam},
                          {PSI_NAMES[q]: ref.psi[q] for q in fam})
        assert max(errs.values()) <= tol, errs


def _check_monitors(ref, dev, tol=TOL):
    mr, md = ref.monitor_results(), dev.monitor_results()
    for name, entry in mr.items():
        for fam in ("E", "H"):
            comps = [c for c in entry["fields"] if c[0] == fam]
            if not comps:
                continue
            def data(e, c):
                return e["components"][c]["data"] if isinstance(e["components"][c], dict) else e["components"][c]
            refs = {c: data(entry, c) for c in comps}
            # a monitor the pulse never reached compares round-off, not physics
            assert max(float(np.max(np.abs(r))) for r in refs.values()) > 1e-20, (name, "monitor saw no field")
            errs = family_err({c: data(md[name], c) for c in comps}, refs)
            assert max(errs.values()) <= tol, (name, errs)
        if "flux" in entry:
            assert norm_err(md[name]["flux"], entry["flux"]) <= tol, (name, md[name]["flux"], entry["flux"])


@pytest.mark.parametrize("fused", FUSED)
def test_vacuum_dipole_cpml(fused):
    print(f"case: vacuum dipole 20 cells CPML 8, 300 steps, fused={fused}", flush=True)
    scene, kw = configs.vacuum_dipole(cells=20, pml=8)
    plan, ref, dev = _run_both(scene, kw, 300, fused)
    _check_fields(ref, dev)
    _check_monitors(ref, dev)
    for f in range(len(plan.dft_frequencies)):
        d = {c + p: (dev.dft_re if p == "re" else dev.dft_im)[c][f].double().cpu().numpy()
             for c in ("Ex", "Ey", "Ez") for p in ("re", "im")}
        r = {c + p: ref.dft[c][0 if p == "re" else 1][f] for c in ("Ex", "Ey", "Ez") for p in ("re", "im")}
        errs = family_err(d, r)
        assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize("fused", FUSED)
def test_dielectric_boxes_soft_plane_wave(fused):
    print(f"case: dielectric boxes 32 cells, soft plane wave, 250 steps, fused={fused}", flush=True)
    scene, kw = configs.dielectric_boxes(cells=32, pml=8, n_boxes=6, box_range=(4, 12), seed=3)
    plan, ref, dev = _run_both(scene, kw, 250, fused)
    assert plan.h_terms and plan.e_terms
    _check_fields(ref, dev)
    _check_monitors(ref, dev)


@pytest.mark.parametrize("fused", FUSED)
def test_dispersive_drude_lorentz(fused):
    print(f"case: Drude + 2 Lorentz spheres 28x28x24, 400 steps, fused={fused}", flush=True)
    scene, kw = configs.dispersive_spheres(cells=(28, 28, 24), pml=8, fill=0.3, radius=(3, 6), seed=1,
                                           injection="soft", pulse_delay=3e-15)
    plan, ref, dev = _run_both(scene, kw, 400, fused)
    assert plan.dispersive
    _check_fields(ref, dev)
    d, r = {}, {}
    for c in ("Ex", "Ey", "Ez"):
        for i, (p, st) in enumerate(zip(ref.ade[c]["poles"], dev.ade_state[c])):
            d[(c, i)] = st["current"].double().cpu().numpy()
            r[(c, i)] = p["cur"]
    errs = family_err(d, r)
    assert max(errs.values()) <= TOL, errs
    _check_monitors(ref, dev)


@pytest.mark.parametrize("fused", FUSED)
def test_debye_drude_lorentz_mixed_media(fused):
    """A Debye medium (relaxation pole, polarisation state) beside a Drude +
    Lorentz medium: all three pole kinds of the fused ADE advance/subtraction
    (reference kernels/dispersive.cu:18-83) on the device."""
    print(f"case: Debye box + Drude/Lorentz box, dipole, 300 steps, fused={fused}", flush=True)
    dx = configs.DX
    L = 10 * dx
    two_pi = 2 * np.pi
    scene = mw.Scene(domain=mw.Domain(((-L, L),) * 3), grid=mw.GridSpec.uniform(dx),
                     boundary=mw.BoundarySpec.pml(6))
    scene.add_structure(mw.Structure(mw.Box((3 * dx, 0, 0), (5 * dx, 9 * dx, 9 * dx)),
                                     mw.Material(eps_r=2.5, debye_poles=(mw.DebyePole(3.0, 2e-15),))))
    scene.add_structure(mw.Structure(mw.Box((-4 * dx, 1 * dx, 0), (4 * dx, 6 * dx, 7 * dx)),
                                     mw.Material(eps_r=1.8,
                                                 drude_poles=(mw.DrudePole(1.5e15 / two_pi, 1e14 / two_pi),),
                                                 lorentz_poles=(mw.LorentzPole(1.1, 3e15 / two_pi, 2e14 / two_pi),))))
    scene.add_source(mw.PointDipole((-0.4 * dx, -4.2 * dx, 0.3 * dx), (0.3, 0.4, 1.0), profile="ideal",
                                    source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4)))
    scene.add_monitor(mw.PointMonitor("p", (4 * dx, 1 * dx, -1 * dx), fields=("Ex", "Ez", "Hy")))
    states = []

    def ade_state(plan, ref, dev):
        d, r = {}, {}
        for c in ("Ex", "Ey", "Ez"):
            for i, (p, st) in enumerate(zip(ref.ade[c]["poles"], dev.ade_state[c])):
                for kind, key in (("cur", "current"), ("pol", "polarization")):
                    if st[key] is not None:
                        d[(c, i, kind)] = st[key].double().cpu().numpy()
                        r[(c, i, kind)] = p[kind].copy()
        states.append((d, r))

    plan, ref, dev = _run_both(scene, {"frequencies": (configs.F0,), "courant": 0.9}, 300, fused, on_mid=ade_state)
    ade_state(plan, ref, dev)
    kinds = {p["kind"] for c in ("Ex", "Ey", "Ez") for p in plan.ade[c]["poles"]}
    assert kinds == {"debye", "drude", "lorentz"}
    _check_fields(ref, dev)
    # pole state judged, like the fields, against its peak over the run (mid-run,
    # while the pulse drives the media, and at the end): the Debye polarisation
    # relaxes within ~15 steps and has decayed by the end
    for kind in ("cur", "pol"):
        keys = [k for k in states[0][1] if k[2] == kind]
        scale = max(float(np.max(np.abs(r[k]))) for _, r in states for k in keys)
        for d, r in states:
            errs = family_err({k: d[k] for k in keys}, {k: r[k] for k in keys}, scale=scale)
            assert max(errs.values()) <= TOL, (kind, errs)
    _check_monitors(ref, dev)


@pytest.mark.parametrize("fused", FUSED)
@pytest.mark.parametrize("kind", ["pec", "pmc", "periodic"])
def test_boundary_modes(kind, fused):
    print(f"case: boundary {kind}, 18^3 cells, dipole, 200 steps, fused={fused}", flush=True)
    dx = configs.DX
    L = 9 * dx
    scene = mw.Scene(domain=mw.Domain(((-L, L),) * 3), grid=mw.GridSpec.uniform(dx),
                     boundary=mw.BoundarySpec.faces(default=kind, z="pml", num_layers=6) if kind != "periodic"
                     else mw.BoundarySpec.faces(default="periodic", z="pml", num_layers=6))
    scene.add_source(mw.PointDipole((0.3 * dx, 0.2 * dx, 0.1 * dx), (1.0, 0.5, 0.25), profile="ideal",
                                    source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4)))
    scene.add_monitor(mw.PointMonitor("p", (2 * dx, -dx, 0.0), fields=("Ex", "Hz")))
    plan, ref, dev = _run_both(scene, {"frequencies": (configs.F0,), "courant": 0.99}, 200, fused)
    _check_fields(ref, dev)
    _check_monitors(ref, dev)


@pytest.mark.parametrize("tiling", [(3, 2, 2), (0, 4, 1), (0, 4, 3)], ids=["aligned", "zedge-asked", "zends-asked"])
@pytest.mark.parametrize("faces", ["x_periodic_y_pmc", "z_periodic", "xy_periodic"])
def test_periodic_faces_single_pass(faces, tiling):
    """Periodic faces in the single-pass step (the shell's E pass reads H across
    the seam), next to PMC/PEC/CPML faces, on grids long enough in z for interior
    tiles and z-edge tiles; a dielectric and the source straddle the seam."""
    print(f"case: single-pass periodic faces {faces}, tiling {tiling}", flush=True)
    dx = configs.DX
    if faces == "x_periodic_y_pmc":
        bnd = mw.BoundarySpec.faces(x="periodic", y="pmc", z="pml", num_layers=6)
    elif faces == "z_periodic":
        bnd = mw.BoundarySpec.faces(default="pml", z="periodic", num_layers=6)
    else:
        bnd = mw.BoundarySpec.faces(x="periodic", y="periodic", z="pec", num_layers=6)
    scene = mw.Scene(domain=mw.Domain(((-10 * dx, 10 * dx), (-9 * dx, 9 * dx), (-40 * dx, 40 * dx))),
                     grid=mw.GridSpec.uniform(dx), boundary=bnd)
    scene.add_structure(mw.Structure(mw.Box((9 * dx, 0, 3 * dx), (6 * dx, 5 * dx, 8 * dx)), mw.Material(eps_r=3.0)))
    scene.add_source(mw.PointDipole((9.4 * dx, 0.3 * dx, 39.2 * dx), (0.5, 0.3, 1.0), profile="ideal",
                                    source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4)))
    scene.add_monitor(mw.PointMonitor("p", (-9 * dx, 2 * dx, -38 * dx), fields=("Ex", "Hz")))
    plan, ref, dev = _run_both(scene, {"frequencies": (configs.F0,), "courant": 0.95}, 240, True, tiling)
    assert 2 in plan.bc_lo and plan.shape[2] >= 63 + 4
    _check_fields(ref, dev)
    _check_monitors(ref, dev)


@pytest.mark.parametrize("fused", FUSED)
def test_conductive_and_gaussian_dipole(fused):
    print(f"case: conductive box + gaussian-profile dipole, 150 steps, fused={fused}", flush=True)
    dx = configs.DX
    L = 10 * dx
    scene = mw.Scene(domain=mw.Domain(((-L, L),) * 3), grid=mw.GridSpec.uniform(dx),
                     boundary=mw.BoundarySpec.pml(6))
    scene.add_structure(mw.Structure(mw.Box((2 * dx, 0, 0), (6 * dx, 8 * dx, 8 * dx)),
                                     mw.Material(eps_r=4.0, sigma_e=2e4)))
    scene.add_source(mw.PointDipole((-3 * dx, 0, 0), "Ey", width=1.5 * dx,
                                    source_time=mw.GaussianPulse(configs.F0, configs.F0 / 3)))
    scene.add_monitor(mw.PlaneMonitor("pl", axis="x", position=4.3 * dx, fields=("Ey", "Ez", "Hy", "Hz"),
                                      compute_flux=True))
    plan, ref, dev = _run_both(scene, {"frequencies": (configs.F0, 1.1 * configs.F0), "courant": 0.95}, 150, fused)
    _check_fields(ref, dev)
    _check_monitors(ref, dev)


TILINGS = [None, (3, 1, 1), (8, 4, 2)]


@pytest.mark.parametrize("tiling", TILINGS, ids=["default", "x3-s1-zedge", "x8-s4-aligned"])
@pytest.mark.parametrize("case", ["vacuum_dft", "boxes_soft_source", "dispersive"])
def test_lean_interior_kernel(case, tiling):
    """Grids whose interior spans more than one 63-column tile inside the PML, so
    the single-pass step runs its interior kernel next to the PML-shell kernels
    (vacuum: homogeneous coefficients + full-field DFT; boxes: per-sample E curl
    with the soft source planes crossing interior tiles; dispersive: fused ADE),
    under the default tiling, short chunks with z-edge tiles (z CPML and faces in
    the interior-row layout) and long chunks with PML-aligned z tiles."""
    print(f"case: lean interior kernel, {case}, tiling {tiling}", flush=True)
    if case == "vacuum_dft":
        scene, kw = configs.vacuum_dipole(cells=20, pml=8)
        dx = configs.DX
        scene.domain = mw.Domain(((-12 * dx, 12 * dx), (-14 * dx, 14 * dx), (-50 * dx, 50 * dx)))
        steps = 160
    elif case == "boxes_soft_source":
        scene, kw = configs.dielectric_boxes(cells=24, pml=8, n_boxes=8, box_range=(4, 12), seed=5)
        dx = configs.DX
        scene.domain = mw.Domain(((-12 * dx, 12 * dx), (-14 * dx, 14 * dx), (-52 * dx, 52 * dx)))
        steps = 200
    else:
        scene, kw = configs.dispersive_spheres(cells=(24, 26, 100), pml=8, fill=0.3, radius=(3, 7), seed=2,
                                               injection="soft", pulse_delay=3e-15)
        steps = 1400  # the resolved-plasma dt is ~1/110 period: the pulse needs ~1000 steps to reach the monitor
    plan, ref, dev = _run_both(scene, kw, steps, True, tiling)
    assert plan.shape[2] >= 63 + 2 * 8 + 4
    _check_fields(ref, dev)
    _check_monitors(ref, dev)


def _z_ends_fit(plan):
    """Do both z ends of this grid, with the interior tiles aligned to the z PML
    (63 owned columns per tile, one halo column before), fit half a wavefront with
    a halo lane each?  (The condition under which z_tiling 3 takes effect.)"""
    nz = plan.shape[2]
    lay = {n: plan.psi_layout[n] for n in ("psi_hx_z", "psi_ex_z")}
    klo = max(1, lay["psi_hx_z"][1], lay["psi_ex_z"][1])
    khi = min(nz - 1, lay["psi_hx_z"][2] if lay["psi_hx_z"][3] > 0 else nz - 1,
              lay["psi_ex_z"][2] if lay["psi_ex_z"][3] > 0 else nz)
    first = klo + 1
    off = first % 63 - 63 if first % 63 else 0
    z0 = off + 63 * ((first - off) // 63)
    n = 0
    while first + 63 * (n + 1) <= khi:
        n += 1
    z1 = z0 + 63 * n
    return n > 0 and (z0 + 1) + (nz - z1 + 1) <= 32


@pytest.mark.parametrize("z_tiling", [3, 4], ids=["one-stream", "side-stream"])
@pytest.mark.parametrize("case", ["vacuum_dft", "boxes_soft_source", "dispersive"])
def test_z_end_tiles(case, z_tiling):
    """z_tiling 3: interior tiles aligned to the z PML, and both z ends (CPML
    layers and faces) of two rows advanced by one wavefront, on grids whose z ends
    fit that layout; homogeneous + full-field DFT, per-sample curl with soft
    source planes, fused ADE.  z_tiling 4 runs the z-end launch on a second
    stream beside the interior launch, joined before shell E."""
    print(f"case: z-end tiles, {case}, z_tiling {z_tiling}", flush=True)
    dx = configs.DX
    if case == "vacuum_dft":
        scene, kw = configs.vacuum_dipole(cells=20, pml=8)
        scene.domain = mw.Domain(((-12 * dx, 12 * dx), (-14 * dx, 14 * dx), (-34 * dx, 34 * dx)))
        steps = 160
    elif case == "boxes_soft_source":
        scene, kw = configs.dielectric_boxes(cells=24, pml=8, n_boxes=8, box_range=(4, 12), seed=5)
        scene.domain = mw.Domain(((-12 * dx, 12 * dx), (-14 * dx, 14 * dx), (-34 * dx, 34 * dx)))
        steps = 200
    else:
        scene, kw = configs.dispersive_spheres(cells=(24, 26, 68), pml=8, fill=0.3, radius=(3, 7), seed=2,
                                               injection="soft", pulse_delay=3e-15)
        steps = 1400
    plan, ref, dev = _run_both(scene, kw, steps, True, (4, 2, z_tiling))
    assert _z_ends_fit(plan), plan.shape
    _check_fields(ref, dev)
    _check_psi(plan, ref, dev)
    _check_monitors(ref, dev)


@pytest.mark.parametrize("mode,fused", [("slab", True), ("slab", False), ("box", True), ("box", False)])
def test_tfsf_plane_wave(mode, fused):
    """TFSF faces driven by the auxiliary line: slab (dispersive spheres, CPML
    on every face) and box (-y incidence into a dielectric scene)."""
    print(f"case: TFSF {mode}, fused={fused}", flush=True)
    if mode == "slab":
        scene, kw = configs.dispersive_spheres(cells=(20, 22, 70), pml=6, fill=0.25, radius=(3, 6), seed=4,
                                               injection="tfsf", pulse_delay=3e-15)
        steps = 1400  # ~0.07 cells/step at the resolved-plasma dt: the pulse crosses the monitor by ~1100
    else:
        dx = configs.DX
        scene = mw.Scene(domain=mw.Domain(((-10 * dx, 10 * dx), (-16 * dx, 16 * dx), (-9 * dx, 9 * dx))),
                         grid=mw.GridSpec.uniform(dx), boundary=mw.BoundarySpec.pml(6), device="cuda")
        scene.add_structure(mw.Structure(mw.Box((0, 0, 0), (6 * dx, 5 * dx, 4 * dx)), mw.Material(eps_r=4.0)))
        scene.add_source(mw.PlaneWave((0, -1, 0), (0, 0, 1), source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4),
                                      injection=mw.TFSF(((-7 * dx, 7 * dx), (-11 * dx, 11 * dx), (-6 * dx, 6 * dx)))))
        scene.add_monitor(mw.FluxMonitor("sf_flux", axis="y", position=-13 * dx, frequencies=(configs.F0,)))
        kw = {"frequencies": (configs.F0,), "courant": 0.95}
        steps = 400
    plan, ref, dev = _run_both(scene, kw, steps, fused)
    assert plan.aux is not None and (plan.line_e_terms and plan.line_h_terms)
    _check_fields(ref, dev)
    _check_monitors(ref, dev)


@pytest.mark.parametrize("tiling", [None, (3, 1, 1), (3, 1, 3), (3, 1, 4)], ids=["aligned", "zedge", "zends",
                                                                            "zends-side-stream"])
def test_staged_planes_equal_per_step_sums(tiling):
    """z-normal plane monitors staged and folded a block at a time (in calls that
    end inside blocks and at their ends) give the per-step sums to the bit: the
    fold adds a block's rows in step order with one fused multiply-add each, the
    rounding chain of adding each step as it is taken.  Planes in the interior,
    the z CPML and at a z end of the lean box; both against the oracle too."""
    from oracle.driver import OracleRun
    from witwin_maxwell_amd.fdtd.runtime import DeviceRun
    print(f"case: staged vs per-step z planes, tiling {tiling}", flush=True)
    dx = configs.DX
    scene, kw = configs.dielectric_boxes(cells=20, pml=6, n_boxes=5, box_range=(3, 8), seed=3)
    scene.domain = mw.Domain(((-10 * dx, 10 * dx), (-11 * dx, 11 * dx), (-40 * dx, 40 * dx)))
    scene.monitors.clear()
    for zc in (-38.3 * dx, 0.2 * dx, 33.7 * dx):
        scene.add_source(mw.PointDipole((0.3 * dx, -0.2 * dx, zc), (0.4, 0.6, 1.0), profile="ideal",
                                        source_time=mw.CW(configs.F0)))
    for name, z in (("z_pml", -38.5 * dx), ("z_mid", 0.0), ("z_end", 34 * dx)):
        scene.add_monitor(mw.PlaneMonitor(name, axis="z", position=z, fields=("Ex", "Hy"), frequencies=kw["frequencies"]))
    kw = dict(kw, time_steps=150, window="none")
    kw.pop("theta", None)
    plan = compile_plan(scene, **kw)
    runs = {}
    for depth, splits in ((0, ()), (16, (7, 40, 64, 99))):
        dev = DeviceRun(plan, "cuda", plane_stage_depth=depth)
        if tiling is not None:
            dev.set_tiling(*tiling)
        assert (dev.plane_stage is not None) == (depth > 0)
        for end in splits:
            dev.run(end - dev.step_index)
        dev.run()
        torch.cuda.synchronize()
        runs[depth] = dev
    for a, b in zip(runs[0].plane_re + runs[0].plane_im, runs[16].plane_re + runs[16].plane_im):
        assert torch.equal(a, b)
    ref = OracleRun(plan).run()
    _check_monitors(ref, runs[16])
