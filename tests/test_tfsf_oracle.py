"""Pins the float64 TFSF restatement (CPU only).

For a normally incident plane wave in vacuum, the 3-D Yee update of a field that
is uniform across the transverse axes reduces exactly to the 1-D auxiliary
line's update (same dt, same spacing, same coefficients).  So with exact face
corrections the total-field region holds (up to one global sign) the line's own
samples at the mapped indices, to round-off, and the scattered-field region stays
at round-off.  A wrong face sign, offset, coefficient, sample index (incl. the
mirrored indices of a negative direction) or step placement breaks one of these.
"""

import numpy as np
import pytest

import witwin_maxwell_amd as mw
from witwin_maxwell_amd.fdtd.plan import compile_plan

D = 5e-8


def _run(direction, polarization, mode):
    cells = (10, 11, 40)
    dom = mw.Domain(tuple((0.0, n * D) for n in cells))
    if mode == "box":
        boundary = mw.BoundarySpec.pml(6)
    else:
        boundary = mw.BoundarySpec.faces(x="periodic", y="periodic", z="pml", num_layers=6)
    scene = mw.Scene(domain=dom, grid=mw.GridSpec.uniform(D), boundary=boundary, device="cpu")
    if mode == "box":
        inj = mw.TFSF(((2.6 * D, 7.4 * D), (2.6 * D, 8.4 * D), (8.2 * D, 31.6 * D)))
    else:
        inj = mw.TFSF.slab(axis="z", bounds=(8.2 * D, 31.6 * D))
    scene.add_source(mw.PlaneWave(direction, polarization, source_time=mw.GaussianPulse(3e14, 1e14),
                                  injection=inj))
    plan = compile_plan(scene, time_steps=260, courant=0.95)
    from oracle.driver import OracleRun
    run = OracleRun(plan)
    return plan, run


@pytest.mark.parametrize("direction,polarization", [((0, 0, 1), (1, 0, 0)), ((0, 0, -1), (0, 1, 0))])
def test_box_tfsf_total_field_is_the_line_and_scattered_field_is_empty(oracle_lib, direction, polarization):
    plan, run = _run(direction, polarization, "box")
    lo, hi = plan.aux["lower"], plan.aux["upper"]
    nz = plan.shape[2]
    comp = "Ex" if polarization[0] else "Ey"
    worst_in, worst_out, peak = 0.0, 0.0, 0.0
    for n in range(plan.time_steps):
        run.step()
        if n % 20 != 19:
            continue
        f = run.fields[comp]
        # E samples inside the TF box (component-specific index ranges)
        si = slice(lo[0], hi[0] + (0 if comp == "Ex" else 1))
        sj = slice(lo[1], hi[1] + (1 if comp == "Ex" else 0))
        ks = np.arange(lo[2], hi[2] + 1)
        line_idx = ks if direction[2] > 0 else nz - 1 - ks
        line = run.aux_e[line_idx]
        inside = f[si, sj, lo[2]:hi[2] + 1]
        peak = max(peak, float(np.abs(line).max()))
        # a uniform transverse field equal to -line or +line (one global sign)
        d_neg = np.abs(inside + line[None, None, :]).max()
        d_pos = np.abs(inside - line[None, None, :]).max()
        worst_in = max(worst_in, min(d_neg, d_pos))
        mask = np.ones(f.shape, dtype=bool)
        mask[si, sj, lo[2]:hi[2] + 1] = False
        worst_out = max(worst_out, float(np.abs(f[mask]).max()))
    assert peak > 0.1, peak
    assert worst_in <= 1e-9 * peak, (worst_in, peak)
    assert worst_out <= 1e-9 * peak, (worst_out, peak)


def test_slab_tfsf_scattered_side_is_empty(oracle_lib):
    """Slab mode with periodic transverse faces: the slab spans the whole cross
    section, so both half-spaces outside it stay empty.  (With CPML transverse
    faces the slab stops at the layers, as in the reference
    (tfsf_state.py:150-158), and its edges radiate.)"""
    plan, run = _run((0, 0, 1), (1, 0, 0), "slab")
    lo, hi = plan.aux["lower"], plan.aux["upper"]
    peak = worst = 0.0
    for n in range(plan.time_steps):
        run.step()
        f = run.fields["Ex"]
        peak = max(peak, float(np.abs(f[:, :, lo[2]:hi[2] + 1]).max()))
        worst = max(worst, float(np.abs(f[:, :, :lo[2]]).max()), float(np.abs(f[:, :, hi[2] + 1:]).max()))
    assert peak > 0.1
    assert worst <= 1e-9 * peak, (worst, peak)
