"""Scene -> FDTD plan: every array and schedule one run needs, in float64 on the host.

The plan is what both the native runtime (after a float32 cast) and the CPU
checkers consume, so the two always run the same discretised problem.  The
compile steps follow the reference's ``initialize_solver`` (dt and spacing
reciprocals, fdtd/runtime/initialization.py:72-140), ``FDTD.init_field`` (field
shapes, fdtd/runtime/stepping.py:1624-1654) and ``solve`` preparation
(stepping.py:2153-2253).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ..constants import C_0
from ..scene import PreparedScene, prepare_scene
from ..sources import TFSF, PlaneWave, PointDipole, cw_signal, gaussian_pulse_signal, ricker_wavelet_signal
from . import cpml as _cpml
from . import excitation as _exc
from . import materials as _mat
from . import spectral as _spec

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
BOUNDARY_CODE = {"none": 0, "pml": 1, "periodic": 2, "pec": 4, "pmc": 5}


def signal_value(st, t):
    """s(t) of a compiled source time (sources.py:497-521)."""
    if st["kind"] == "cw":
        return cw_signal(t, frequency=st["frequency"], amplitude=st["amplitude"], phase=st["phase"])
    if st["kind"] == "ricker_wavelet":
        return ricker_wavelet_signal(t, frequency=st["frequency"], amplitude=st["amplitude"], delay=st["delay"])
    return gaussian_pulse_signal(t, frequency=st["frequency"], fwidth=st["fwidth"], amplitude=st["amplitude"],
                                 phase=st["phase"], delay=st["delay"])


def auto_dt(dx, dy, dz, frequency, steps_per_cycle=30):
    """min(period/30, Courant limit) (fdtd/solver.py:145-159)."""
    dt1 = (1.0 / frequency) / steps_per_cycle
    dt2 = 1.0 / (C_0 * math.sqrt(1.0 / dx ** 2 + 1.0 / dy ** 2 + 1.0 / dz ** 2))
    return min(dt1, dt2)


@dataclass
class FDTDPlan:
    prepared: PreparedScene
    shape: tuple
    dt: float
    time_steps: int
    inv_e: list
    inv_h: list
    bc_lo: tuple
    bc_hi: tuple
    coef: dict                   # comp -> (decay, curl), float64 arrays of the component shape
    eps_edges: dict              # E comp -> absolute permittivity on the edge
    uses_cpml: bool
    cpml: dict                   # axis -> profiles
    psi_layout: dict             # psi name -> (axis, low, high_start, high_len)
    ade: dict                    # E comp -> {"inv_eps", "poles"}
    point_terms: list            # additive patches with per-step signal
    e_terms: list                # time-shifted E surface terms (t + dt/2)
    h_terms: list                # time-shifted H surface terms (t)
    signals: np.ndarray          # (n_point_terms, steps) per-step host signal
    observers: _spec.ObserverLayout | None
    obs_weights: tuple | None    # (E_cos, E_sin, H_cos, H_sin) (steps, F)
    obs_window_sum: np.ndarray | None
    dft_frequencies: tuple
    dft_weights: tuple | None    # (E_cos, E_sin) (steps, F)
    dft_window_sum: np.ndarray | None
    node_eps: np.ndarray
    material_model: object = None
    line_e_terms: list = field(default_factory=list)   # TFSF E faces sampling the auxiliary line's H
    line_h_terms: list = field(default_factory=list)   # TFSF H faces sampling the auxiliary line's E
    aux: dict | None = None                            # TFSF auxiliary 1-D line
    meta: dict = field(default_factory=dict)

    def field_shape(self, comp):
        return _mat.component_shape(comp, self.shape)

    @property
    def dispersive(self):
        return any(self.ade[c]["poles"] for c in ("Ex", "Ey", "Ez"))


def _step_dt(scene, prepared, sts, default_f, courant):
    """The time step: Courant number times the 3-D stability limit of the smallest
    cells, or the reference's automatic step for the highest frequency in play."""
    min_d = [float(prepared.primal(a).min()) for a in "xyz"]
    if courant is not None:
        return float(courant) / (C_0 * math.sqrt(sum(1.0 / d ** 2 for d in min_d)))
    f_char = max([st["characteristic_frequency"] for st in sts] or [default_f])
    f_char = max(f_char, _mat.material_characteristic_frequency(scene))
    return float(auto_dt(*min_d, f_char))


def plan_dt(scene, *, frequencies=(), courant=None, default_frequency=None, **_unused) -> float:
    """The time step ``compile_plan`` would choose for these arguments, without
    compiling the plan (takes and ignores compile_plan's other keywords)."""
    prepared = prepare_scene(scene)
    default_f = float(default_frequency or (frequencies[0] if frequencies else 1e9))
    sts = [_exc.primary_source_time(s, default_f) for s in scene.sources]
    return _step_dt(scene, prepared, sts, default_f, courant)


def compile_plan(scene, *, frequencies=(), time_steps=None, dt=None, courant=None, full_field_dft=False,
                 window="hanning", cpml_config=None, default_frequency=None, region_densities=None,
                 node_eps_override=None) -> FDTDPlan:
    prepared = prepare_scene(scene)
    shape = prepared.shape
    if min(shape) < 2:
        raise ValueError("every axis needs at least two Yee nodes")

    sources = list(scene.sources)
    for s in sources:
        if not isinstance(s, (PointDipole, PlaneWave)):
            raise NotImplementedError(f"source type {type(s).__name__} is outside this FDTD runtime")
    default_f = float(default_frequency or (frequencies[0] if frequencies else 1e9))
    sts = [_exc.primary_source_time(s, default_f) for s in sources]
    pulsed = any(st["kind"] != "cw" for st in sts[:1])

    min_d = [float(prepared.primal(a).min()) for a in "xyz"]
    dt = float(dt) if dt is not None else _step_dt(scene, prepared, sts, default_f, courant)
    if time_steps is None:
        raise ValueError("time_steps must be given")
    steps = int(time_steps)

    inv_e = [1.0 / prepared.dual(a) for a in "xyz"]
    inv_h = [1.0 / prepared.primal(a) for a in "xyz"]
    bc_lo = tuple(BOUNDARY_CODE[prepared.boundary.face_kind(a, "low")] for a in "xyz")
    bc_hi = tuple(BOUNDARY_CODE[prepared.boundary.face_kind(a, "high")] for a in "xyz")

    for region in scene.material_regions:
        if tuple(region.mu_bounds) != (1.0, 1.0):
            raise NotImplementedError("MaterialRegion.mu_bounds other than (1, 1) is outside this FDTD runtime")
    model = _mat.NodeMaterialModel(prepared)
    if node_eps_override is not None:
        eps_node = np.asarray(node_eps_override, dtype=np.float64)
    else:
        eps_node_t = _mat.node_permittivity(prepared, model, region_densities)
        eps_node = eps_node_t.detach().cpu().numpy()
    coef, eps_edges = _mat.update_coefficients(prepared, model, eps_node, dt)
    ade = _mat.dispersive_templates(model, eps_edges, dt)

    uses_cpml = prepared.boundary.uses_kind("pml") and prepared.pml_thickness > 0
    profiles = _cpml.build_cpml_profiles(prepared, dt, cpml_config) if uses_cpml else None
    layout = {}
    if uses_cpml:
        for name in _cpml.PSI_NAMES:
            comp, _ = _cpml.PSI_SPECS[name]
            layout[name] = _cpml.psi_layout(prepared, name, _mat.component_shape(comp, shape))

    point_terms, e_terms, h_terms = [], [], []
    line_e, line_h, aux = [], [], None
    for src, st in zip(sources, sts):
        if isinstance(src, PointDipole):
            point_terms.extend(_exc.point_dipole_terms(prepared, src, eps_edges, dt, st))
        elif isinstance(src, PlaneWave):
            if isinstance(src.injection, TFSF):
                if aux is not None:
                    raise ValueError("FDTD currently supports at most one TFSF source per scene.")
                from .tfsf import tfsf_setup
                values = np.array([signal_value(st, n * dt) for n in range(steps)])
                line_e, line_h, aux = tfsf_setup(prepared, src, coef, dt, st, values, eps_node, model)
            else:
                et, ht = _exc.soft_plane_wave_terms(prepared, src, coef, dt, st)
                e_terms.extend(et)
                h_terms.extend(ht)
    signals = np.zeros((len(point_terms), steps))
    for ti, term in enumerate(point_terms):
        for n in range(steps):
            signals[ti, n] = signal_value(term["st"], n * dt)

    win = "none" if pulsed else str(window)
    observers = None
    obs_weights = None
    obs_sum = None
    monitors = list(scene.monitors)
    if monitors:
        observers = _spec.ObserverLayout(prepared, monitors, tuple(frequencies) or (sts[0]["frequency"] if sts else default_f,))
        starts = [_spec.spectral_start_step(prepared, f, dt, win, pulsed) for f in observers.frequencies]
        ec, es, hc, hs, wsum, _ = _spec.weight_tables(observers.frequencies, dt, steps, starts, win)
        obs_weights = (ec, es, hc, hs)
        obs_sum = wsum
    dft_freqs = tuple(float(f) for f in frequencies) if full_field_dft else ()
    dft_weights = None
    dft_sum = None
    if dft_freqs:
        starts = [_spec.spectral_start_step(prepared, f, dt, win, pulsed) for f in dft_freqs]
        ec, es, _, _, wsum, _ = _spec.weight_tables(dft_freqs, dt, steps, starts, win)
        dft_weights = (ec, es)
        dft_sum = wsum

    return FDTDPlan(prepared=prepared, shape=shape, dt=dt, time_steps=steps, inv_e=inv_e, inv_h=inv_h,
                    bc_lo=bc_lo, bc_hi=bc_hi, coef=coef, eps_edges=eps_edges, uses_cpml=uses_cpml,
                    cpml=profiles, psi_layout=layout, ade=ade, point_terms=point_terms, e_terms=e_terms,
                    h_terms=h_terms, signals=signals, observers=observers, obs_weights=obs_weights,
                    obs_window_sum=obs_sum, dft_frequencies=dft_freqs, dft_weights=dft_weights,
                    dft_window_sum=dft_sum, node_eps=eps_node, material_model=model,
                    line_e_terms=line_e, line_h_terms=line_h, aux=aux,
                    meta={"window": win, "pulsed": pulsed})
