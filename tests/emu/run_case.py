"""One small single-pass case on the CPU emulation of the HIP library, compared
with the float64 oracle.  Run by tests/test_emulated_kernels.py in a child process
(AddressSanitizer must be preloaded into the interpreter that loads the library).

usage: python run_case.py LIB [CASE ...]   (cases: boxes staged dispersive periodic adjoint; default all)"""
import contextlib
import ctypes
import inspect
import os
import sys
import textwrap

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402
import torch  # noqa: E402

from witwin_maxwell_amd.fdtd import native as nat  # noqa: E402

lib = nat.load_library(sys.argv[1])
nat.load_library = lambda path=None: lib
nat.ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
nat.current_stream = lambda: ctypes.c_void_p(None)
torch.cuda.device = lambda *a, **k: contextlib.nullcontext()
torch.cuda.synchronize = lambda *a, **k: None

from witwin_maxwell_amd.fdtd import runtime  # noqa: E402

# DeviceRun refuses non-HIP devices; the emulated library works on host memory
_src = textwrap.dedent(inspect.getsource(runtime.DeviceRun.__init__))
_src = _src.replace('if self.device.type != "cuda":', "if False:")
_ns = {}
exec(compile(_src, "emulated DeviceRun.__init__", "exec"), runtime.__dict__, _ns)
runtime.DeviceRun.__init__ = _ns["__init__"]

import witwin_maxwell_amd as mw  # noqa: E402
from oracle.driver import OracleRun  # noqa: E402
from witwin_maxwell_amd import configs  # noqa: E402
from witwin_maxwell_amd.fdtd.cpml import PSI_NAMES  # noqa: E402
from witwin_maxwell_amd.fdtd.plan import COMPONENTS, compile_plan  # noqa: E402

from conftest import family_err  # noqa: E402


def _monitor_errs(ref, dev, need_field=False):
    """Family errors of every plane/point monitor payload against the oracle's
    (need_field: every monitor must have seen field, so the comparison is not of
    zeros)."""
    def data(e, c):
        return e["components"][c]["data"] if isinstance(e["components"][c], dict) else e["components"][c]
    errs = {}
    mr, md = ref.monitor_results(), dev.monitor_results()
    for name, entry in mr.items():
        for fam in ("E", "H"):
            comps = [c for c in entry["fields"] if c[0] == fam]
            if comps:
                if need_field:
                    peak = max(float(np.max(np.abs(data(entry, c)))) for c in comps)
                    assert peak > 1e-20, (name, fam, "monitor saw no field")
                got = family_err({c: np.asarray(data(md[name], c)) for c in comps}, {c: data(entry, c) for c in comps})
                errs.update({f"{name}/{c}": v for c, v in got.items()})
    return errs


lib.emu_kernel_launches.argtypes = [ctypes.c_char_p]
lib.emu_kernel_launches.restype = ctypes.c_long


def launches(part):
    return lib.emu_kernel_launches(part.encode())


def run(name, scene, kw, steps, tiling, splits=(), stage_depth=32, need_field=False, fused=True, expect=None):
    """Run `steps` steps in calls ending at `splits` (then the rest) with z planes
    staged `stage_depth` steps at a time, and compare fields, psi and monitors;
    `expect`: a kernel name that the run must launch."""
    before = launches(expect) if expect else 0
    scene.device = "cpu"
    kw = dict(kw)
    kw.pop("theta", None)
    kw["time_steps"] = steps
    plan = compile_plan(scene, **kw)
    ref = OracleRun(plan).run(steps)
    dev = runtime.DeviceRun(plan, "cpu", fused=fused, plane_stage_depth=stage_depth).set_tiling(*tiling)
    for end in splits:
        dev.run(end - dev.step_index)
    dev.run()
    errs = _monitor_errs(ref, dev, need_field)
    if plan.observers is not None and any(a == "z" for _, a, _ in plan.observers.planes):
        assert dev.plane_stage is not None, "z planes not staged"
    for fam in (COMPONENTS[:3], COMPONENTS[3:]):
        scale = max(ref.peak[c] for c in fam)
        errs.update(family_err({c: dev.fields[c].double().numpy() for c in fam}, {c: ref.fields[c] for c in fam},
                               scale=scale))
    if plan.uses_cpml:
        for fam in (range(0, 6), range(6, 12)):
            errs.update(family_err({PSI_NAMES[q]: dev.psi_dense(PSI_NAMES[q]).double().numpy() for q in fam},
                                   {PSI_NAMES[q]: ref.psi[q] for q in fam}))
    worst = max(errs.values())
    print(f"{name}: {plan.shape} tiling {tiling} max err {worst:.2e}", flush=True)
    assert worst <= 1e-5, errs
    if expect:
        assert launches(expect) > before, f"{name}: no {expect} launch"


dx = configs.DX


def case_boxes():
    # a grid with an interior box, PML shell boxes, soft-source terms and per-sample
    # curl coefficients, under both z tilings; z-normal planes in the interior (the
    # flux monitors), in the z PML slab and at the last interior column are staged
    # (blocks of 3 steps, calls ending inside and at the end of a block), an x-normal
    # plane is sampled after each step
    s, kw = configs.dielectric_boxes(cells=12, pml=5, n_boxes=4, box_range=(3, 6), seed=5)
    s.domain = mw.Domain(((-6 * dx, 6 * dx), (-7 * dx, 7 * dx), (-37 * dx, 37 * dx)))
    s.add_monitor(mw.PlaneMonitor("z_pml", axis="z", position=-34.5 * dx, fields=("Ez", "Hx"),
                                  frequencies=kw["frequencies"]))
    s.add_monitor(mw.PlaneMonitor("z_edge", axis="z", position=31 * dx, fields=("Hz",),
                                  frequencies=kw["frequencies"]))
    s.add_monitor(mw.PlaneMonitor("x_mid", axis="x", position=0.0, fields=("Ey",), frequencies=kw["frequencies"]))
    run("boxes, aligned z tiling", s, kw, 8, (3, 2, 2), splits=(2, 6), stage_depth=3)
    run("boxes, z-edge tiles", s, kw, 8, (0, 4, 1), splits=(3,), stage_depth=4, expect="zedge_step_kernel")
    run("boxes, z-end tiles", s, kw, 8, (3, 2, 3), splits=(5,), stage_depth=3, expect="zends_step_kernel")
    run("boxes, z-end tiles, second stream", s, kw, 8, (3, 2, 4), splits=(5,), stage_depth=3,
        expect="zends_step_kernel")


def case_staged():
    # staged z planes under field from the first step: CW dipoles (full amplitude at
    # t = 0) beside planes in the z PML slab, the interior and the high z end (shell
    # or z-edge tile); blocks of 4 steps, calls ending inside blocks and at their ends,
    # so a fold that adds a step twice or drops one shows at once
    st = mw.Scene(domain=mw.Domain(((-6 * dx, 6 * dx), (-7 * dx, 7 * dx), (-37 * dx, 37 * dx))),
                  grid=mw.GridSpec.uniform(dx), boundary=mw.BoundarySpec.pml(5), device="cpu")
    for zc in (-34.3 * dx, 0.2 * dx, 30.8 * dx):
        st.add_source(mw.PointDipole((0.3 * dx, -0.2 * dx, zc), (0.4, 0.6, 1.0), profile="ideal",
                                     source_time=mw.CW(configs.F0)))
    freqs = (configs.F0, 1.3 * configs.F0)
    st.add_monitor(mw.PlaneMonitor("z_pml", axis="z", position=-34.5 * dx, fields=("Ez", "Hx"), frequencies=freqs))
    st.add_monitor(mw.PlaneMonitor("z_mid", axis="z", position=0.0, fields=("Ey", "Hy"), frequencies=freqs))
    st.add_monitor(mw.PlaneMonitor("z_end", axis="z", position=31 * dx, fields=("Ex", "Hz"), frequencies=freqs))
    st.add_monitor(mw.PlaneMonitor("x_mid", axis="x", position=0.0, fields=("Ez",), frequencies=freqs))
    # a second monitor on one of those planes: both read the same staged samples
    st.add_monitor(mw.PlaneMonitor("z_mid_again", axis="z", position=0.0, fields=("Ey",), frequencies=freqs[:1]))
    skw = {"frequencies": freqs, "courant": 0.9, "window": "none"}
    run("staged planes, aligned z tiling", st, skw, 11, (3, 2, 2), splits=(2, 6, 9), stage_depth=4, need_field=True)
    run("staged planes, z-edge tiles", st, skw, 11, (0, 4, 1), splits=(3, 5), stage_depth=4, need_field=True)
    run("staged planes, z-end tiles", st, skw, 11, (0, 4, 3), splits=(3, 5), stage_depth=4, need_field=True,
        expect="zends_step_kernel")
    run("staged planes, two-kernel step", st, skw, 11, (0, 0, 0), splits=(4,), stage_depth=3, need_field=True,
        fused=False)


def case_dispersive():
    # every pole kind (Debye, Drude, two Lorentz) with the per-sample activity mask,
    # in the interior box, the z-edge tiles and the shell, under both z tilings
    two_pi = 2 * np.pi
    d = mw.Scene(domain=mw.Domain(((-6 * dx, 6 * dx), (-7 * dx, 7 * dx), (-37 * dx, 37 * dx))),
                 grid=mw.GridSpec.uniform(dx), boundary=mw.BoundarySpec.pml(5), device="cpu")
    d.add_structure(mw.Structure(mw.Box((0, 0, -20 * dx), (8 * dx, 9 * dx, 20 * dx)),
                                 mw.Material(eps_r=2.0, debye_poles=(mw.DebyePole(2.0, 3e-15),))))
    d.add_structure(mw.Structure(mw.Sphere((1 * dx, -1 * dx, 12 * dx), 5 * dx),
                                 mw.Material(eps_r=2.0, drude_poles=(mw.DrudePole(1.2e15 / two_pi, 1e14 / two_pi),),
                                             lorentz_poles=(mw.LorentzPole(1.5, 4e15 / two_pi, 2e14 / two_pi),
                                                            mw.LorentzPole(0.8, 6e15 / two_pi, 2e14 / two_pi)))))
    d.add_source(mw.PointDipole((0.2 * dx, 0.3 * dx, -0.4 * dx), (0.5, 0.3, 1.0), profile="ideal",
                                source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4)))
    dkw = {"frequencies": (configs.F0,), "courant": 0.9}
    run("dispersive, aligned z tiling", d, dkw, 12, (3, 2, 2))
    run("dispersive, z-edge tiles", d, dkw, 12, (0, 4, 1))
    run("dispersive, z-end tiles", d, dkw, 12, (0, 4, 3), expect="zends_step_kernel")


def case_periodic():
    # periodic x faces (H read across the seam by the shell's E pass), PMC y faces,
    # CPML z, with an interior box and a dielectric crossing the seam
    pz = mw.Scene(domain=mw.Domain(((-8 * dx, 8 * dx), (-7 * dx, 7 * dx), (-37 * dx, 37 * dx))),
                  grid=mw.GridSpec.uniform(dx),
                  boundary=mw.BoundarySpec.faces(x="periodic", y="pmc", z="pml", num_layers=5), device="cpu")
    pz.add_structure(mw.Structure(mw.Box((7 * dx, 0, 3 * dx), (6 * dx, 5 * dx, 8 * dx)), mw.Material(eps_r=3.0)))
    pz.add_source(mw.PointDipole((6.6 * dx, 0.3 * dx, -0.4 * dx), (0.5, 0.3, 1.0), profile="ideal",
                                 source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4)))
    run("periodic x, PMC y, aligned z tiling", pz, {"frequencies": (configs.F0,), "courant": 0.9}, 10, (3, 2, 2))
    run("periodic x, PMC y, z-edge tiles", pz, {"frequencies": (configs.F0,), "courant": 0.9}, 10, (0, 4, 1))
    run("periodic x, PMC y, z-end tiles", pz, {"frequencies": (configs.F0,), "courant": 0.9}, 10, (0, 4, 3),
        expect="zends_step_kernel")
    # periodic z faces: the z ends stay shell cells even when z-edge tiles are asked for
    pk = mw.Scene(domain=mw.Domain(((-6 * dx, 6 * dx), (-7 * dx, 7 * dx), (-37 * dx, 37 * dx))),
                  grid=mw.GridSpec.uniform(dx), boundary=mw.BoundarySpec.faces(default="pml", z="periodic",
                                                                                num_layers=5), device="cpu")
    pk.add_source(mw.PointDipole((0.2 * dx, 0.3 * dx, 36.6 * dx), (0.5, 0.3, 1.0), profile="ideal",
                                 source_time=mw.GaussianPulse(configs.F0, configs.F0 / 4)))
    run("periodic z, z-edge tiles asked", pk, {"frequencies": (configs.F0,), "courant": 0.9}, 10, (0, 4, 1))
    run("periodic z, z-end tiles asked", pk, {"frequencies": (configs.F0,), "courant": 0.9}, 10, (0, 4, 3))


def case_adjoint():
    # the reverse sweep (interior and boundary-box reverse kernels, seeds, window
    # capture and coefficient gradients) against the oracle adjoint: CPML with
    # Drude + Lorentz poles, and periodic x / PEC y / PMC z faces
    import test_gpu_adjoint as ta
    from oracle.driver import OracleAdjoint
    from witwin_maxwell_amd.fdtd.adjoint import DeviceAdjoint, design_window, restore
    from conftest import norm_err
    for bname in ("cpml_ade", "periodic_pec_pmc"):
        scene, _ = ta._design_case(bname)
        scene.device = "cpu"
        steps = 40
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
        dev = runtime.DeviceRun(plan, "cpu")
        dev.run(steps)
        adj = DeviceAdjoint(dev, window, capacity=steps)
        adj.zero()
        t = lambda a: torch.as_tensor(a, dtype=torch.float32)
        adj.set_seeds(point=(t(g_point[0]), t(g_point[1])), planes=[(t(r), t(i)) for r, i in g_planes])
        restore(dev, None, 0)
        adj.capture(0, steps)
        adj.reverse(0, steps)
        errs = family_err({c: adj.fields[c].double().numpy() for c in COMPONENTS[:3]},
                          {c: radj.ae[i] for i, c in enumerate(COMPONENTS[:3])})
        lo, hi = window
        box = tuple(slice(l, h) for l, h in zip(lo, hi))
        for ci in range(3):
            errs[f"g_curl{ci}"] = norm_err(adj.g_curl[ci].double().numpy(), radj.g_curl[ci][box]) * 1e-2
            errs[f"g_decay{ci}"] = norm_err(adj.g_decay[ci].double().numpy(), radj.g_decay[ci][box]) * 1e-2
        worst = max(errs.values())
        print(f"adjoint {bname}: {plan.shape} max err {worst:.2e}", flush=True)
        assert worst <= 1e-5, errs


CASES = {"boxes": case_boxes, "staged": case_staged, "dispersive": case_dispersive, "periodic": case_periodic,
         "adjoint": case_adjoint}

for name in (sys.argv[2:] or list(CASES)):
    CASES[name]()
print("OK")
