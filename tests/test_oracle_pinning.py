"""Pins the float64 oracle against closed forms and invariants (CPU only).

The reference holds no golden vectors for this path, so every oracle routine is
checked against mathematics that a plausible transcription mistake (wrong sign,
index, spacing, coefficient or profile) would violate:

* PEC-cavity eigenmodes evolve exactly as the leapfrog recurrence of the 3-D Yee
  dispersion relation sin^2(w dt/2)/(c dt)^2 = sum_a sin^2(k_a d_a/2)/d_a^2
  (update stencils, signs, spacings, PEC faces; one mode per E component).
* The discrete Yee energy E.eps.E + H^{n-1/2}.mu.H^{n+1/2} is conserved to
  round-off in a PEC box with an inhomogeneous dielectric (curl coefficients).
* A uniform field in a periodic box oscillates/relaxes at the analytic
  Drude plasma, Lorentz longitudinal and Debye relaxation rates (ADE).
* CPML: psi stays zero outside the layers, the profile grades monotonically and a
  pulse leaves a PML-terminated box with small reflection versus a larger box.
* DFT accumulation equals a direct numpy sum of the recorded samples.
"""

import math

import numpy as np
import pytest

import witwin_maxwell_amd as mw
from witwin_maxwell_amd.constants import EPSILON_0, MU_0
from witwin_maxwell_amd.fdtd.plan import compile_plan

C = 1.0 / math.sqrt(MU_0 * EPSILON_0)


def _box_scene(n, d, boundary, structures=(), sources=(), monitors=()):
    L = n * d
    scene = mw.Scene(domain=mw.Domain(((0, L),) * 3), grid=mw.GridSpec.uniform(d), boundary=boundary, device="cpu")
    for s in structures:
        scene.add_structure(s)
    for s in sources:
        scene.add_source(s)
    for m in monitors:
        scene.add_monitor(m)
    return scene


def _oracle(plan):
    from oracle.driver import OracleRun
    return OracleRun(plan)


@pytest.mark.parametrize("comp", ["Ex", "Ey", "Ez"])
def test_pec_cavity_mode_follows_yee_dispersion(oracle_lib, comp):
    n, d = (10, 12, 14), 1e-7
    dims = {"x": n[0] * d, "y": n[1] * d, "z": n[2] * d}
    scene = mw.Scene(domain=mw.Domain(((0, dims["x"]), (0, dims["y"]), (0, dims["z"]))),
                     grid=mw.GridSpec.uniform(d), boundary=mw.BoundarySpec.pec(), device="cpu")
    plan = compile_plan(scene, frequencies=(1e14,), time_steps=400, courant=0.9)
    run = _oracle(plan)
    ax = "xyz".index(comp[1])
    others = [a for a in range(3) if a != ax]
    prep = plan.prepared
    coords = prep.component_coords(comp)
    ks, mode = [], np.ones(plan.field_shape(comp))
    for a in others:
        L = (prep.x_nodes64, prep.y_nodes64, prep.z_nodes64)[a][-1]
        m = 1 if a == others[0] else 2
        k = m * math.pi / L
        ks.append((k, d))
        shape = [1, 1, 1]
        shape[a] = -1
        mode = mode * np.sin(k * coords[a]).reshape(shape)
    run.fields[comp][...] = mode
    s2 = (C * plan.dt / 2) ** 2 * sum((2 * math.sin(k * dd / 2) / dd) ** 2 for k, dd in ks)
    theta = 2 * math.asin(math.sqrt(s2))
    for step in range(1, 401):
        run.step()
        if step % 50 == 0:
            expect = math.cos(step * theta) - math.tan(theta / 2) * math.sin(step * theta)
            got = run.fields[comp]
            err = np.max(np.abs(got - expect * mode))
            assert err < 1e-9, (comp, step, err)
    # the other E components stay identically zero for this mode
    for c in ("Ex", "Ey", "Ez"):
        if c != comp:
            assert np.max(np.abs(run.fields[c])) < 1e-12


def test_energy_conserved_in_pec_box_with_dielectric(oracle_lib):
    rng = np.random.default_rng(4)
    d = 1e-7
    boxes = [mw.Structure(mw.Box(tuple(rng.uniform(0.3e-6, 0.9e-6, 3)), tuple(rng.uniform(2e-7, 6e-7, 3))),
                          mw.Material(eps_r=float(rng.uniform(1.5, 9.0)))) for _ in range(4)]
    scene = _box_scene(12, d, mw.BoundarySpec.pec(), structures=boxes)
    plan = compile_plan(scene, frequencies=(1e14,), time_steps=200, courant=0.95)
    run = _oracle(plan)
    for c in ("Ex", "Ey", "Ez"):
        run.fields[c][...] = rng.standard_normal(run.fields[c].shape)
    run.step()  # PEC clears the tangential E on the walls
    eps = {c: plan.eps_edges[c] for c in ("Ex", "Ey", "Ez")}
    mu = MU_0

    def step_energy():
        # leapfrog invariant: eps E^n.E^{n+1} + mu |H^{n+1/2}|^2
        e_prev = {c: run.fields[c].copy() for c in ("Ex", "Ey", "Ez")}
        run.step()
        we = sum(np.sum(eps[c] * e_prev[c] * run.fields[c]) for c in ("Ex", "Ey", "Ez"))
        wh = sum(np.sum(mu * run.fields[c] ** 2) for c in ("Hx", "Hy", "Hz"))
        return we + wh

    w0 = step_energy()
    for _ in range(150):
        w1 = step_energy()
    assert abs(w1 - w0) / abs(w0) < 1e-10, (w0, w1)
    # and the energy actually moved between E and H (the test is not vacuous)
    assert sum(np.sum(run.fields[c] ** 2) for c in ("Hx", "Hy", "Hz")) > 0


@pytest.mark.parametrize("kind", ["drude", "lorentz", "debye"])
def test_uniform_field_ade_rates(oracle_lib, kind):
    d = 1e-7
    two_pi = 2 * math.pi
    wp, gamma = 2.0e15, 1.0e13
    w0, dl = 3.0e15, 1.5
    tau = 2e-15
    eps_inf = 2.0
    if kind == "drude":
        mat = mw.Material(eps_r=eps_inf, drude_poles=(mw.DrudePole(wp / two_pi, gamma / two_pi),))
    elif kind == "lorentz":
        mat = mw.Material(eps_r=eps_inf, lorentz_poles=(mw.LorentzPole(dl, w0 / two_pi, gamma / two_pi),))
    else:
        mat = mw.Material(eps_r=eps_inf, debye_poles=(mw.DebyePole(dl, tau),))
    scene = _box_scene(4, d, mw.BoundarySpec.periodic(),
                       structures=[mw.Structure(mw.Box((0, 0, 0), (1, 1, 1)), mat)])
    steps = 3000
    plan = compile_plan(scene, frequencies=(1e14,), time_steps=steps, dt=2e-18)
    run = _oracle(plan)
    run.fields["Ez"][...] = 1.0
    trace = []
    for _ in range(steps):
        run.step()
        trace.append(run.fields["Ez"][1, 1, 1])
    trace = np.asarray(trace)
    assert np.allclose(run.fields["Ez"], run.fields["Ez"][1, 1, 1])  # stays uniform
    t = (np.arange(steps) + 1) * plan.dt
    if kind == "debye":
        # eps_inf E + P = D constant; P relaxes to dl*E with tau_eff = tau*eps_inf/(eps_inf+dl)
        e_inf = eps_inf / (eps_inf + dl)
        tau_eff = tau * eps_inf / (eps_inf + dl)
        expect = e_inf + (1 - e_inf) * np.exp(-t / tau_eff)
        assert np.max(np.abs(trace - expect)) < 2e-3
        return
    # oscillation frequency from zero crossings, damping from the envelope
    w_expect = wp / math.sqrt(eps_inf) if kind == "drude" else w0 * math.sqrt((eps_inf + dl) / eps_inf)
    if kind == "lorentz":
        # D = eps_inf*E + P is conserved, so E oscillates about eps_inf/(eps_inf+dl)
        centred = trace - eps_inf / (eps_inf + dl)
    else:
        centred = trace
    sgn = np.sign(centred)
    crossings = np.nonzero(np.diff(sgn) != 0)[0]
    period = 2 * np.mean(np.diff(crossings)) * plan.dt
    assert abs(2 * math.pi / period - w_expect) / w_expect < 5e-3, (2 * math.pi / period, w_expect)
    peaks = np.abs(centred)
    early = peaks[: steps // 4].max()
    late = peaks[3 * steps // 4:].max()
    decay = math.log(early / late) / (0.5 * steps * plan.dt)
    assert 0.3 * gamma / 2 < decay < 3.0 * gamma / 2 + 1e12


def test_cpml_profile_and_psi_confinement(oracle_lib):
    d = 1e-7
    scene = _box_scene(16, d, mw.BoundarySpec.pml(6), sources=[
        mw.PointDipole((0.8e-6, 0.8e-6, 0.8e-6), "Ez", profile="ideal",
                       source_time=mw.GaussianPulse(1.2e14, 4e13))])
    plan = compile_plan(scene, frequencies=(1.2e14,), time_steps=120, courant=0.95)
    prof = plan.cpml["x"]
    n = plan.shape[0]
    sigma_like = 1.0 - prof["b_e"]
    # graded: zero interior, monotone increasing towards both outer faces
    assert np.all(sigma_like[6:n - 6] == 0)
    assert np.all(np.diff(sigma_like[:7]) <= 0) and np.all(np.diff(sigma_like[n - 7:]) >= 0)
    assert np.all(prof["c_e"] <= 0)
    run = _oracle(plan).run()
    from witwin_maxwell_amd.fdtd.cpml import PSI_NAMES
    for q, name in enumerate(PSI_NAMES):
        axis, low, hs, hl = plan.psi_layout[name]
        psi = np.moveaxis(run.psi[q], axis, 0)
        inner = psi[low:hs]
        assert np.max(np.abs(inner)) == 0.0, name
    assert max(np.max(np.abs(p)) for p in run.psi) > 0


def test_cpml_reflection_small(oracle_lib):
    d = 1e-7
    src = mw.PointDipole((0.0, 0.0, 0.0), "Ez", profile="ideal", source_time=mw.GaussianPulse(2e14, 8e13))
    probe = (3 * d, 0.0, 0.0)

    def run(half_cells, steps):
        L = half_cells * d
        scene = mw.Scene(domain=mw.Domain(((-L, L),) * 3), grid=mw.GridSpec.uniform(d),
                         boundary=mw.BoundarySpec.pml(10), device="cpu")
        scene.add_source(src)
        plan = compile_plan(scene, frequencies=(2e14,), time_steps=steps, courant=0.95)
        r = _oracle(plan)
        idx = tuple(int(np.argmin(np.abs(c - p))) for c, p in zip(plan.prepared.component_coords("Ez"), probe))
        out = []
        for _ in range(steps):
            r.step()
            out.append(r.fields["Ez"][idx])
        return np.asarray(out)

    steps = 260
    small = run(12, steps)
    big = run(40, steps)  # reflections from the far boundary arrive after the window
    err = np.max(np.abs(small - big)) / np.max(np.abs(big))
    assert err < 2e-2, err


def test_dft_accumulation_matches_direct_sum(oracle_lib):
    d = 1e-7
    f0 = 1.5e14
    scene = _box_scene(12, d, mw.BoundarySpec.pml(4), sources=[
        mw.PointDipole((0.6e-6, 0.6e-6, 0.6e-6), "Ey", profile="ideal", source_time=mw.GaussianPulse(f0, f0 / 3))],
        monitors=[mw.PointMonitor("p", (0.7e-6, 0.6e-6, 0.6e-6), fields=("Ey", "Hz"))])
    freqs = (0.8 * f0, f0, 1.2 * f0)
    plan = compile_plan(scene, frequencies=freqs, time_steps=150, courant=0.9, full_field_dft=True)
    run = _oracle(plan)
    (cy, iy), (ch, ih) = plan.observers.points
    ey, hz = [], []
    for _ in range(plan.time_steps):
        run.step()
        ey.append(run.fields["Ey"][iy])
        hz.append(run.fields["Hz"][ih])
    n = np.arange(plan.time_steps)
    res = run.monitor_results()["p"]
    for fi, f in enumerate(freqs):
        w = 2 * np.pi * f
        e_direct = np.sum(np.asarray(ey) * np.exp(1j * w * n * plan.dt)) * 2 / plan.time_steps
        h_direct = np.sum(np.asarray(hz) * np.exp(1j * w * (n - 0.5) * plan.dt)) * 2 / plan.time_steps
        assert abs(res["components"]["Ey"][fi] - e_direct) <= 1e-12 * abs(e_direct)
        assert abs(res["components"]["Hz"][fi] - h_direct) <= 1e-12 * abs(h_direct)
        full = run.dft["Ey"][0][fi][iy] + 1j * run.dft["Ey"][1][fi][iy]
        assert abs(full * 2 / plan.time_steps - e_direct) <= 1e-12 * abs(e_direct)


def test_soft_plane_wave_dispersion_and_one_way_injection(oracle_lib):
    """Soft plane wave in a transversely periodic vacuum column: the field is
    uniform across the plane, propagates at the phase speed fixed by the 1-D Yee
    dispersion relation, and (the E and H surface terms cancelling backwards)
    leaves almost nothing behind the launch plane."""
    d = 5e-8
    f0 = 2.0e14
    nz_cells, nxy = 160, 4
    scene = mw.Scene(domain=mw.Domain(((0, nxy * d), (0, nxy * d), (0, nz_cells * d))),
                     grid=mw.GridSpec.uniform(d),
                     boundary=mw.BoundarySpec.faces(default="periodic", z="pml", num_layers=12), device="cpu")
    # -z propagation: the pulse exits through the low z face (see the reflection
    # test below for why the high face is avoided here)
    scene.add_source(mw.PlaneWave((0, 0, -1), (1, 0, 0), source_time=mw.GaussianPulse(f0, f0 / 5)))
    steps = 1400
    plan = compile_plan(scene, frequencies=(f0,), time_steps=steps, courant=0.9)
    run = _oracle(plan)
    z = plan.prepared.z_nodes64
    kz_src = plan.e_terms[0]["offset"][2]
    za, zb = kz_src - 30, kz_src - 70
    back = kz_src + 6
    ta, tb, tback = [], [], []
    for _ in range(steps):
        run.step()
        ex = run.fields["Ex"]
        ta.append(ex[1, 2, za])
        tb.append(ex[1, 2, zb])
        tback.append(ex[1, 2, back])
        assert np.allclose(ex[:, :, za], ex[1, 2, za], rtol=0, atol=1e-12 * (abs(ex[1, 2, za]) + 1e-30) + 1e-30)
    n = np.arange(steps)
    w = 2 * np.pi * f0
    pa = np.sum(np.asarray(ta) * np.exp(1j * w * n * plan.dt))
    pb = np.sum(np.asarray(tb) * np.exp(1j * w * n * plan.dt))
    # sum_n E_n e^{+i w n dt} of cos(w t - k s) is proportional to e^{+i k s}, s
    # the distance travelled, so pb/pa = e^{i k |zb - za|}
    dphi = np.angle(pb / pa)
    s = math.sin(w * plan.dt / 2) * d / (C * plan.dt)
    k_num = 2 * math.asin(s) / d
    expect = (k_num * abs(z[zb] - z[za]) + np.pi) % (2 * np.pi) - np.pi
    assert abs(((dphi - expect) + np.pi) % (2 * np.pi) - np.pi) < 2e-3, (dphi, expect)
    assert np.max(np.abs(tback)) < 2e-2 * np.max(np.abs(ta))


@pytest.mark.parametrize("grading,high_bound", [(None, 0.1), ("staggered", 1e-3)])
def test_cpml_face_reflection(oracle_lib, grading, high_bound):
    """Normal-incidence reflection of a plane-wave pulse from the low and high z
    faces (12 layers, 30 cells per wavelength).  The low face is graded at the
    true Yee positions; the reference grades high-face H half a cell deep
    (fdtd/boundary/cpml.py:408-420), which reflects a few percent."""
    d = 5e-8
    f0 = 2.0e14
    out = {}
    for direction in (1, -1):
        scene = mw.Scene(domain=mw.Domain(((0, 4 * d), (0, 4 * d), (0, 160 * d))), grid=mw.GridSpec.uniform(d),
                         boundary=mw.BoundarySpec.faces(default="periodic", z="pml", num_layers=12), device="cpu")
        scene.add_source(mw.PlaneWave((0, 0, direction), (1, 0, 0), source_time=mw.GaussianPulse(f0, f0 / 5)))
        steps = 1400
        cfg = {"h_grading": grading} if grading else None
        plan = compile_plan(scene, frequencies=(f0,), time_steps=steps, courant=0.9, cpml_config=cfg)
        run = _oracle(plan)
        probe = plan.e_terms[0]["offset"][2] + 70 * direction
        trace = []
        for _ in range(steps):
            run.step()
            trace.append(run.fields["Ex"][1, 2, probe])
        trace = np.abs(np.asarray(trace))
        peak = int(np.argmax(trace))
        sigma_steps = 1.0 / (2 * np.pi * f0 / 5) / plan.dt
        out[direction] = trace[int(peak + 5 * sigma_steps):].max() / trace.max()
    assert out[-1] < 1e-3, out
    assert out[1] < high_bound, out


def test_sparse_pole_update_equals_dense(oracle_lib):
    """The oracle advances and applies each pole over the samples where its
    drive is non-zero; over many steps that gives exactly the dense routines'
    numbers (a pole with zero drive keeps a zero state and adds nothing)."""
    import ctypes
    lib = oracle_lib.load("f64")
    rng = np.random.default_rng(11)
    n = 4000
    P = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int64)
    ptr = lambda a: a.ctypes.data_as(P)
    for kind in (0, 1, 2):
        drive = rng.standard_normal(n) * (rng.random(n) < 0.3)
        idx = np.flatnonzero(drive).astype(np.int64)
        inv_eps = 1.0 + rng.random(n)
        dense = {k: np.zeros(n) for k in ("pol", "cur", "e")}
        sparse = {k: np.zeros(n) for k in ("pol", "cur", "e")}
        for step in range(20):
            e_in = rng.standard_normal(n)
            for st in (dense, sparse):
                st["e"][:] = e_in
            lib.or_ade_advance(kind, n, ptr(dense["e"]), ptr(drive), 0.9, 0.05, 1e-2, ptr(dense["pol"]),
                               ptr(dense["cur"]))
            lib.or_ade_advance_idx(kind, idx.size, idx.ctypes.data_as(IP), ptr(sparse["e"]), ptr(drive), 0.9, 0.05,
                                   1e-2, ptr(sparse["pol"]), ptr(sparse["cur"]))
            lib.or_ade_apply(n, ptr(dense["cur"]), ptr(inv_eps), 1e-2, ptr(dense["e"]))
            lib.or_ade_apply_idx(idx.size, idx.ctypes.data_as(IP), ptr(sparse["cur"]), ptr(inv_eps), 1e-2,
                                 ptr(sparse["e"]))
            for k in ("pol", "cur", "e"):
                assert np.array_equal(dense[k], sparse[k]), (kind, step, k)
        assert np.any(dense["cur"] != 0)
