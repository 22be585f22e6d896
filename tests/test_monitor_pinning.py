"""Monitor payloads and flux quadrature pinned by closed forms (CPU only).

The product assembles monitor payloads in witwin_maxwell_amd.fdtd.spectral and the
oracle in oracle/monitors.py, two independent restatements of the reference's
observer and quadrature code (fdtd/observers.py:25-37, :126-241, :296-405;
fdtd/quadrature.py:78-174).  Both are fed the accumulated DFT sums of a known
field pattern on a flux plane and must return the analytic flux:

* tangential E = E0 (1 + a u + b v), tangential H = H0 (constant) on a plane whose
  physical part spans the box [u0, u1] x [v0, v1]: the time-averaged flux is
  1/2 E0 H0 (A + a int u dA + b int v dA), exact for the midpoint rule on cell
  centres after the node->centre averaging (the field is linear);
* the flux flips sign with normal_direction="-";
* a point monitor returns the sample of the nearest Yee node;
* a plane between two sample planes interpolates linearly.
A wrong half factor, cell width, averaging axis, crop, sign or scale fails one.
"""

import math

import numpy as np
import pytest

import witwin_maxwell_amd as mw
from witwin_maxwell_amd.fdtd.plan import compile_plan
from witwin_maxwell_amd.fdtd.spectral import assemble_monitor_results

D = 1e-7
_AX = {"x": 0, "y": 1, "z": 2}


def _plan(monitors, nonuniform=False):
    n = (12, 10, 14)
    if nonuniform:
        rng = np.random.default_rng(4)
        coords = [np.concatenate([[0.0], np.cumsum(rng.uniform(0.7, 1.3, size=k) * D)]) for k in n]
        grid = mw.GridSpec.custom(*coords)
        dom = mw.Domain(tuple((0.0, float(c[-1])) for c in coords))
    else:
        grid = mw.GridSpec.uniform(D)
        dom = mw.Domain(tuple((0.0, k * D) for k in n))
    scene = mw.Scene(domain=dom, grid=grid, boundary=mw.BoundarySpec.pml(3), device="cpu")
    scene.add_source(mw.PointDipole((5 * D, 5 * D, 5 * D), "Ez", profile="ideal",
                                    source_time=mw.GaussianPulse(3e14, 1e14)))
    for m in monitors:
        scene.add_monitor(m)
    return compile_plan(scene, frequencies=(3e14,), time_steps=8, courant=0.9)


def _pattern(plan, comp, plane_axis, coeffs, amp):
    """Samples of amp * (1 + sum_a c_a * coord_a) on the plane of ``comp`` at
    every sample, as a (plane a, plane b) array."""
    coords = plan.prepared.component_coords(comp)
    across = [c for i, c in enumerate(coords) if i != _AX[plane_axis]]
    names = [a for a in "xyz" if a != plane_axis]
    val = np.ones((across[0].size, across[1].size))
    for k, name in enumerate(names):
        c = coeffs.get(name, 0.0)
        shape = [1, 1]
        shape[k] = -1
        val = val + c * across[k].reshape(shape)
    return amp * val


def _analytic_flux(plan, plane_axis, coeffs, e0, h0, sign):
    r = plan.prepared.physical_domain_range
    names = [a for a in "xyz" if a != plane_axis]
    lo = {a: r[2 * _AX[a]] for a in "xyz"}
    hi = {a: r[2 * _AX[a] + 1] for a in "xyz"}
    area = (hi[names[0]] - lo[names[0]]) * (hi[names[1]] - lo[names[1]])
    mean = 1.0 + sum(coeffs.get(a, 0.0) * 0.5 * (lo[a] + hi[a]) for a in names)
    return 0.5 * e0 * h0 * area * mean * sign


# (plane axis, E component carrying the field, H component, sign of S.n for E, H > 0)
CASES = [("z", "Ex", "Hy", +1.0), ("z", "Ey", "Hx", -1.0), ("x", "Ey", "Hz", +1.0), ("y", "Ez", "Hx", +1.0)]


@pytest.mark.parametrize("nonuniform", [False, True])
@pytest.mark.parametrize("direction", ["+", "-"])
@pytest.mark.parametrize("axis,ecomp,hcomp,orient", CASES)
def test_flux_of_linear_tangential_field(axis, ecomp, hcomp, orient, direction, nonuniform):
    from oracle.monitors import OracleObservers
    probe = _plan([], nonuniform)
    nodes = probe.prepared.nodes(axis)
    pos = float(nodes[nodes.size // 2])
    mon = mw.FluxMonitor("f", axis=axis, position=pos, frequencies=(3e14,), normal_direction=direction)
    plan = _plan([mon], nonuniform)
    coeffs = {a: 0.3 / (D * 10) * (k + 1) for k, a in enumerate("xyz")}
    e0, h0 = 2.0, 0.7
    half_sum = 0.5 * plan.obs_window_sum  # payload scale is 2 / sum(w)
    # product: fill its plane sums with the pattern (real part only)
    layout = plan.observers
    shapes = {c: plan.field_shape(c) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    plane_re, plane_im = [], []
    for comp, ax, index in layout.planes:
        shp = layout.plane_shape(shapes, (comp, ax, index))
        if comp == ecomp:
            val = _pattern(plan, comp, ax, coeffs, e0)
        elif comp == hcomp:
            val = np.full(shp, h0)
        else:
            val = np.zeros(shp)
        plane_re.append(val[None] * half_sum[:, None, None])
        plane_im.append(np.zeros((1,) + shp))
    nf = len(layout.frequencies)
    res = assemble_monitor_results(layout, np.zeros((nf, 0)), np.zeros((nf, 0)), plane_re, plane_im,
                                   plan.obs_window_sum)
    sign = orient * (1.0 if direction == "+" else -1.0)
    expected = _analytic_flux(plan, axis, coeffs, e0, h0, sign)
    assert res["f"]["flux"] == pytest.approx(expected, rel=1e-12)
    # oracle: its own sample planes and assembly
    obs = OracleObservers(plan)
    wsum = obs.wsum
    for (comp, ax, index), acc in obs.planes.items():
        if comp == ecomp:
            acc[0] = _pattern(plan, comp, ax, coeffs, e0)[None] * (0.5 * wsum)[:, None, None]
        elif comp == hcomp:
            acc[0] = h0 * 0.5 * wsum[:, None, None] * np.ones_like(acc[0])
    assert obs.results()["f"]["flux"] == pytest.approx(expected, rel=1e-12)


def test_point_and_interpolated_plane_payloads():
    from oracle.monitors import OracleObservers
    probe = _plan([])
    hx = probe.prepared.component_coords("Hx")
    zpos = 0.25 * hx[2][6] + 0.75 * hx[2][7]
    point = (4.2 * D, 6.8 * D, 3.1 * D)
    mons = [mw.PointMonitor("p", point, fields=("Ez", "Hx"), frequencies=(3e14,)),
            mw.PlaneMonitor("q", axis="z", position=zpos, fields=("Hx",), frequencies=(3e14,))]
    plan = _plan(mons)
    obs = OracleObservers(plan)
    fields = {c: np.random.default_rng(7).standard_normal(plan.field_shape(c))
              for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    for n in range(plan.time_steps):
        obs.record(n, fields)
    res = obs.results()
    wc = obs.w["E"][0][:, 0]
    ws = obs.w["E"][1][:, 0]
    scale = 2.0 / obs.wsum[0]
    ez = plan.prepared.component_coords("Ez")
    idx = tuple(int(np.argmin(np.abs(ez[a] - point[a]))) for a in range(3))
    expect = fields["Ez"][idx] * (wc.sum() + 1j * ws.sum()) * scale
    assert res["p"]["components"]["Ez"] == pytest.approx(expect, rel=1e-12)
    hc = obs.w["H"][0][:, 0].sum() + 1j * obs.w["H"][1][:, 0].sum()
    plane = 0.25 * fields["Hx"][:, :, 6] + 0.75 * fields["Hx"][:, :, 7]
    got = res["q"]["components"]["Hx"]["data"]
    assert np.allclose(got, plane * hc * scale, rtol=1e-9, atol=0)
    # the product's layout picks the same samples with the same weights
    rec = [r for r in plan.observers.records if r["monitor"].name == "q"][0]
    ws_ = sorted((plan.observers.planes[p][2], w) for p, w, _ in rec["components"]["Hx"]["samples"])
    assert [k for k, _ in ws_] == [6, 7] and np.allclose([w for _, w in ws_], [0.25, 0.75])
    # and the H weights lag the E weights by half a step
    f = 3e14
    n = np.arange(plan.time_steps)
    assert np.allclose(obs.w["H"][0][:, 0], np.cos(2 * math.pi * f * (n - 0.5) * plan.dt))
