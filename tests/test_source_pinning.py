"""Source normalisation pinned by closed forms (CPU only).

* Point dipole (reference fdtd/excitation/injection.py:36-71, :127-202, :264-272):
  the patch added to E each step is -dt * p * profile / eps, with a profile of unit
  integral.  So the patch's moment sum_g patch_g eps_g dV_g equals -dt * p exactly,
  whatever the permittivity under it, and the trilinear "ideal" stencil reproduces
  the dipole position as its first moment; the Gaussian profile is centred on the
  dipole with variance width^2 per axis (to the quadrature error of a resolved
  Gaussian).  A dropped 1/eps, a wrong control volume, swapped interpolation
  weights or an off-by-one offset fails one of these.
* Soft plane-wave power scale (reference injection.py:732-765, :884-891): the
  surface amplitude is 1/sqrt(A cos(theta) / (2 eta0)) / sqrt(cos(k dz / 2)) over
  the full computational aperture A, so a unit-amplitude source carries 1 W.
"""

import math

import numpy as np
import pytest

import witwin_maxwell_amd as mw
from witwin_maxwell_amd.constants import ETA_0
from witwin_maxwell_amd.fdtd.plan import compile_plan

D = 5e-8


def _dipole_plan(profile, polarization, position, eps_box=None, width=None, d=D):
    n = 24
    scene = mw.Scene(domain=mw.Domain(((0.0, n * d),) * 3), grid=mw.GridSpec.uniform(d),
                     boundary=mw.BoundarySpec.pml(4), device="cpu")
    if eps_box is not None:
        scene.add_structure(mw.Structure(mw.Box(position, (4 * d, 3 * d, 5 * d)), mw.Material(eps_r=eps_box)))
    kw = dict(profile=profile, source_time=mw.GaussianPulse(4e14, 1e14))
    if width is not None:
        kw["width"] = width
    scene.add_source(mw.PointDipole(position, polarization, **kw))
    return compile_plan(scene, frequencies=(4e14,), time_steps=4, courant=0.9)


def _moments(plan, term):
    from witwin_maxwell_amd.fdtd.excitation import _control_widths
    comp, off, patch = term["field"], term["offset"], term["patch"]
    coords = plan.prepared.component_coords(comp)
    sl = [slice(off[a], off[a] + patch.shape[a]) for a in range(3)]
    w = [_control_widths(coords[a])[sl[a]] for a in range(3)]
    vol = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
    eps = plan.eps_edges[comp][tuple(sl)]
    mass = patch * eps * vol
    pos = [coords[a][sl[a]] for a in range(3)]
    first = [float(np.sum(mass * np.expand_dims(pos[a], [b for b in range(3) if b != a]))) for a in range(3)]
    return mass, pos, first


@pytest.mark.parametrize("eps_box", [None, 3.5])
@pytest.mark.parametrize("polarization", ["Ez", (0.6, 0.0, 0.8), (0.3, -0.5, 0.81)])
def test_ideal_dipole_patch_moment_and_position(polarization, eps_box):
    # the stencil interpolates only where a cell is wider than float32 epsilon in
    # metres (the reference's guard, injection.py:141-143), so use 0.25 um cells
    d = 2.5e-7
    pos = (11.3 * d, 12.7 * d, 10.45 * d)
    plan = _dipole_plan("ideal", polarization, pos, eps_box, d=d)
    pol = mw.sources.polarization_vector(polarization)
    seen = set()
    for term in plan.point_terms:
        a = "xyz".index(term["field"][1])
        seen.add(a)
        mass, _, first = _moments(plan, term)
        total = float(mass.sum())
        assert total == pytest.approx(-plan.dt * pol[a], rel=1e-12)
        for b in range(3):
            assert first[b] / total == pytest.approx(pos[b], rel=1e-12)
    assert seen == {a for a in range(3) if abs(pol[a]) > 1e-12}


def test_ideal_dipole_on_sub_epsilon_cells_uses_the_lower_sample():
    """Cells narrower than float32 epsilon (in metres) make the reference's guard
    put the whole moment on the lower sample of each axis (injection.py:141-143)."""
    pos = (11.3 * D, 12.7 * D, 10.45 * D)
    plan = _dipole_plan("ideal", "Ez", pos)
    (term,) = plan.point_terms
    assert term["patch"].shape == (1, 1, 1)
    coords = plan.prepared.component_coords("Ez")
    for a in range(3):
        c = coords[a][term["offset"][a]]
        assert c <= pos[a] < coords[a][term["offset"][a] + 1]
    mass, _, _ = _moments(plan, term)
    assert float(mass.sum()) == pytest.approx(-plan.dt, rel=1e-12)


@pytest.mark.parametrize("eps_box", [None, 6.0])
def test_gaussian_dipole_patch_moment_centre_and_width(eps_box):
    pos = (12.0 * D, 11.5 * D, 12.25 * D)
    width = 1.6 * D
    plan = _dipole_plan("gaussian", "Ey", pos, eps_box, width=width)
    (term,) = plan.point_terms
    mass, coords, first = _moments(plan, term)
    total = float(mass.sum())
    assert total == pytest.approx(-plan.dt, rel=1e-12)
    for b in range(3):
        assert first[b] / total == pytest.approx(pos[b], abs=1e-3 * D)
        other = [c for c in range(3) if c != b]
        var = float(np.sum(mass * np.expand_dims((coords[b] - pos[b]) ** 2, other))) / total
        assert math.sqrt(var) == pytest.approx(width, rel=0.05)


def test_soft_plane_wave_unit_power_scale():
    """E amplitude of the H-side terms equals 1/sqrt(A/(2 eta0)) / sqrt(cos(k dz/2))
    at normal incidence: the patch divided by the curl coefficient over dz."""
    from witwin_maxwell_amd.fdtd.excitation import solve_numerical_wavenumber
    n = 20
    scene = mw.Scene(domain=mw.Domain(((-n * D, n * D), (-0.5 * n * D, 0.5 * n * D), (-n * D, n * D))),
                     grid=mw.GridSpec.uniform(D), boundary=mw.BoundarySpec.pml(6), device="cpu")
    f0 = 3e14
    scene.add_source(mw.PlaneWave((0, 0, 1), (1, 0, 0), source_time=mw.GaussianPulse(f0, f0 / 4)))
    plan = compile_plan(scene, frequencies=(f0,), time_steps=2, courant=0.95)
    r = plan.prepared.domain_range
    area = (r[1] - r[0]) * (r[3] - r[2])
    k = solve_numerical_wavenumber(2 * math.pi * f0, plan.dt, (0.0, 0.0, 1.0), {"x": D, "y": D, "z": D})
    expected = 1.0 / math.sqrt(area / (2.0 * ETA_0)) / math.sqrt(math.cos(0.5 * k * D))
    # H-side term of Hy sampling Ex: patch = sign * E0 * curl_H / dz
    (hy,) = [t for t in plan.h_terms if t["field"] == "Hy"]
    curl_h = plan.coef["Hy"][1]
    ratio = np.abs(hy["patch"]) / (np.broadcast_to(curl_h, plan.field_shape("Hy"))[
        tuple(slice(o, o + s) for o, s in zip(hy["offset"], hy["patch"].shape))] / D)
    assert np.allclose(ratio, expected, rtol=1e-12)
    # 1/2 |E0|^2 A / eta0 = 1 / cos(k dz / 2): unit power once the Yee factor is removed
    assert 0.5 * expected ** 2 * area / ETA_0 * math.cos(0.5 * k * D) == pytest.approx(1.0, rel=1e-12)
