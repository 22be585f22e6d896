"""``Simulation.fdtd(scene, ...).run() -> Result`` for the FDTD hot path.

Mirrors the reference's front end for this path (witwin/maxwell/simulation.py:122-200
``TimeConfig``/``SpectralSampler``/``FDTDConfig``, :434-549 ``Simulation.fdtd``,
:1042-1054 ``_run_fdtd``, :1522-1591 step resolution and solve, :1612-1700 result
assembly; result.py:1593-1745 ``Result.field``/``tensor``/``monitor``): same names,
argument meanings and errors, running on the MI355X runtime.  A scene whose
``MaterialRegion`` densities require grad runs through the checkpointed adjoint,
so ``result.monitor(...)`` payloads are differentiable.
"""

from __future__ import annotations

from collections.abc import Iterable
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from .constants import C_0

_WINDOWS = ("none", "hanning", "hann")


def _normalize_frequencies(*, frequencies=None, frequency=None) -> tuple:
    """simulation.py:49-63."""
    if frequency is not None and frequencies is not None:
        raise ValueError("Pass either frequency or frequencies, not both.")
    if frequencies is None:
        if frequency is None:
            raise ValueError("A target frequency is required.")
        return (float(frequency),)
    if isinstance(frequencies, Iterable) and not isinstance(frequencies, (str, bytes)):
        values = tuple(float(v) for v in frequencies)
    else:
        values = (float(frequencies),)
    if not values:
        raise ValueError("frequencies must not be empty.")
    return values


@dataclass
class TimeConfig:
    time_steps: int | None = None
    steady_cycles: int = 20
    transient_cycles: int = 15

    @classmethod
    def auto(cls, steady_cycles: int = 20, transient_cycles: int = 15) -> "TimeConfig":
        return cls(time_steps=None, steady_cycles=steady_cycles, transient_cycles=transient_cycles)


@dataclass
class SpectralSampler:
    window: str = "hanning"
    normalize_source: bool = False

    def __post_init__(self):
        w = str(getattr(self.window, "value", self.window)).lower()
        if w not in _WINDOWS:
            raise ValueError(f"Unsupported spectral window {self.window!r}.")
        self.window = "hanning" if w == "hann" else w
        self.normalize_source = bool(self.normalize_source)


@dataclass
class FDTDConfig:
    run_time: TimeConfig = field(default_factory=TimeConfig.auto)
    spectral_sampler: SpectralSampler = field(default_factory=SpectralSampler)
    enable_plot: bool = False
    full_field_dft: bool = False
    absorber: str = "cpml"
    cpml_config: dict = field(default_factory=dict)
    adjoint_checkpoint_stride: int | None = None
    shutoff: float = 0.0
    shutoff_check_interval: int = 100
    cuda_graph: bool = True
    parallel: Any = None
    initial_condition: Any = None

    def __post_init__(self):
        if not isinstance(self.run_time, TimeConfig):
            raise TypeError("run_time must be a TimeConfig instance.")
        if not isinstance(self.spectral_sampler, SpectralSampler):
            raise TypeError("spectral_sampler must be a SpectralSampler instance.")
        self.absorber = str(getattr(self.absorber, "value", self.absorber)).lower()
        self.cpml_config = dict(self.cpml_config or {})
        self.shutoff = float(self.shutoff)
        if self.shutoff < 0:
            raise ValueError("shutoff must be >= 0.")
        self.shutoff_check_interval = int(self.shutoff_check_interval)
        if self.shutoff_check_interval <= 0:
            raise ValueError("shutoff_check_interval must be > 0.")


def calculate_required_steps(frequency, dt, c=C_0, num_cycles=20, transient_cycles=15, domain_size=None,
                             source_time=None):
    """fdtd/solver.py:108-131."""
    period = 1.0 / frequency
    transient = max(transient_cycles, 15) * period
    if domain_size is not None:
        transient = max(transient, 2 * domain_size / c * 5)
    if source_time is not None:
        transient = max(transient, float(source_time.get("settling_time", 0.0)))
    return int((transient + num_cycles * period) / dt)


def _to_tensor(v, device):
    if isinstance(v, torch.Tensor):
        return v
    a = np.asarray(v)
    if np.iscomplexobj(a):
        return torch.as_tensor(a.astype(np.complex64), device=device)
    return torch.as_tensor(a, device=device)


def _tensorize(obj, device):
    """Monitor payloads as torch tensors (the reference's payload type); tensors
    that carry an autograd graph pass through untouched."""
    if isinstance(obj, dict):
        return {k: _tensorize(v, device) for k, v in obj.items()}
    if isinstance(obj, np.ndarray):
        return _to_tensor(obj, device)
    if isinstance(obj, tuple) and obj and all(isinstance(x, np.ndarray) for x in obj):
        return tuple(_to_tensor(x, device) for x in obj)
    return obj


class Result:
    """FDTD result (result.py:1593-1745 subset): monitor payloads, frequency-domain
    (or last-step) E fields, solver statistics."""

    def __init__(self, *, method, scene, prepared_scene, frequency, frequencies, fields, monitors, solver_stats,
                 materials=None):
        self._materials = dict(materials or {})
        self.method = method
        self.scene = scene
        self.prepared_scene = prepared_scene
        self.frequency = frequency
        self.frequencies = tuple(frequencies)
        self._fields = {k.upper(): v for k, v in fields.items()}
        self._monitors = monitors
        self._solver_stats = solver_stats

    @property
    def fields(self) -> dict:
        return dict(self._fields)

    @property
    def monitors(self) -> dict:
        return dict(self._monitors)

    @property
    def solver_stats(self) -> dict:
        return dict(self._solver_stats)

    def _freq_index(self, freqs, frequency, freq_index):
        if frequency is not None and freq_index is not None:
            raise ValueError("Pass either frequency or freq_index, not both.")
        if freq_index is not None:
            if not -len(freqs) <= int(freq_index) < len(freqs):
                raise IndexError("freq_index out of range")
            return int(freq_index)
        if frequency is not None:
            diffs = [abs(float(f) - float(frequency)) for f in freqs]
            i = int(np.argmin(diffs))
            if diffs[i] > 1e-6 * max(abs(float(frequency)), 1.0):
                raise ValueError(f"frequency {frequency} is not among {freqs}")
            return i
        return None

    def tensor(self, name: str, *, frequency=None, freq_index=None):
        key = name.upper()
        if key not in self._fields:
            raise KeyError(f"Field {name!r} is not available in this result.")
        t = self._fields[key]
        i = self._freq_index(self.frequencies, frequency, freq_index)
        if i is not None and len(self.frequencies) > 1 and t.ndim > 0 and t.shape[0] == len(self.frequencies):
            t = t[i]
        return t

    def field(self, name: str = "E", *, frequency=None, freq_index=None):
        key = name.upper()
        if key in ("E", "H"):
            return {c: self.tensor(c, frequency=frequency, freq_index=freq_index)
                    for c in (key + "X", key + "Y", key + "Z") if c in self._fields}
        return self.tensor(key, frequency=frequency, freq_index=freq_index)

    def raw_monitor(self, name: str, *, frequency=None, freq_index=None):
        if name not in self._monitors:
            raise KeyError(f"Monitor {name!r} is not available in this result.")
        payload = self._monitors[name]
        freqs = tuple(payload.get("frequencies", (payload.get("frequency"),)))
        i = self._freq_index(freqs, frequency, freq_index)
        if i is None or len(freqs) <= 1:
            return dict(payload)
        return _select_frequency(payload, freqs, i)

    def monitor(self, name: str, *, frequency=None, freq_index=None, resolve_modal: bool = True):
        return self.raw_monitor(name, frequency=frequency, freq_index=freq_index)

    def material(self, name: str = "eps_r"):
        """Node-sampled material of the compiled scene: ``eps_r`` (design regions
        blended in), ``mu_r`` or ``sigma_e`` (result.py:1088-1092)."""
        key = {"eps": "eps_r", "permittivity": "eps_r", "mu": "mu_r", "permeability": "mu_r"}.get(name, name)
        if key not in self._materials:
            raise KeyError(f"Material {name!r} is not available in this result.")
        return self._materials[key]

    @property
    def materials(self) -> "ResultMaterials":
        return ResultMaterials(self)

    def at(self, *, frequency=None, freq_index=None) -> "ResultSelection":
        """A view of this result at one frequency (result.py:1110-1130)."""
        return ResultSelection(self, frequency=frequency, freq_index=freq_index)

    select = at

    @property
    def E(self):
        return self.field("E")

    @property
    def H(self):
        return self.field("H")


class ResultMaterials:
    """``result.materials.eps`` / ``.mu`` (result.py:1025-1043)."""

    def __init__(self, result):
        self._result = result

    @property
    def eps(self):
        return self._result.material("eps_r")

    @property
    def mu(self):
        return self._result.material("mu_r")

    permittivity = eps
    permeability = mu


class ResultSelection:
    """A result seen at one frequency (result.py:1046-1130): every accessor passes
    the selected frequency on."""

    def __init__(self, result, frequency=None, freq_index=None):
        if frequency is not None and freq_index is not None:
            raise ValueError("Pass either frequency or freq_index, not both.")
        self.result = result
        self.frequency = frequency
        self.freq_index = freq_index

    @property
    def E(self):
        return self.field("E")

    @property
    def H(self):
        return self.field("H")

    @property
    def materials(self):
        return self.result.materials

    def field(self, name: str = "E"):
        return self.result.field(name, frequency=self.frequency, freq_index=self.freq_index)

    def tensor(self, name: str):
        return self.result.tensor(name, frequency=self.frequency, freq_index=self.freq_index)

    def material(self, name: str = "eps_r"):
        return self.result.material(name)

    def raw_monitor(self, name: str):
        return self.result.raw_monitor(name, frequency=self.frequency, freq_index=self.freq_index)

    def monitor(self, name: str, *, resolve_modal: bool | None = None):
        return self.result.monitor(name, frequency=self.frequency, freq_index=self.freq_index)

    def at(self, *, frequency=None, freq_index=None):
        return ResultSelection(self.result, frequency=frequency if frequency is not None else self.frequency,
                               freq_index=freq_index if freq_index is not None else self.freq_index)

    select = at


def _select_frequency(payload, freqs, i):
    out = {}
    n = len(freqs)
    for k, v in payload.items():
        if k == "components":
            out[k] = {c: ({**d, "data": d["data"][i]} if isinstance(d, dict) else d[i]) for c, d in v.items()}
        elif k == "frequencies":
            out[k] = (freqs[i],)
        elif k == "frequency":
            out[k] = freqs[i]
        elif hasattr(v, "shape") and getattr(v, "ndim", 0) > 0 and v.shape[0] == n and k not in ("coords",):
            out[k] = v[i]
        else:
            out[k] = v
    return out


def _resolve_scene_input(scene_like):
    """A Scene, or a SceneModule materialised by its ``to_scene()``
    (simulation.py:234-239)."""
    from .scene import Scene, SceneModule
    if isinstance(scene_like, Scene):
        return scene_like, None
    if isinstance(scene_like, SceneModule):
        return scene_like.to_scene(), scene_like
    raise TypeError("scene must be a maxwell.Scene or maxwell.SceneModule.")


class Simulation:
    """simulation.py:434-549 (FDTD method).  ``scene`` may be a ``SceneModule``:
    it is materialised again by every ``run()`` / ``prepare()`` (simulation.py:
    927-933), so trainable parameters feeding its regions are read fresh."""

    def __init__(self, scene, method, frequencies, config, excitations=None):
        self.scene_input = scene
        self.scene, self.scene_module = _resolve_scene_input(scene)
        self.method = method
        self.frequencies = tuple(frequencies)
        self.config = config
        self.excitations = tuple(excitations or ())

    def _refresh_scene(self):
        if self.scene_module is not None:
            self.scene = self.scene_module.to_scene()

    @property
    def frequency(self):
        return self.frequencies[0]

    @classmethod
    def fdtd(cls, scene, *, frequencies=None, frequency=None, run_time: TimeConfig | None = None, absorber="cpml",
             spectral_sampler: SpectralSampler | None = None, cpml_config=None, enable_plot=False,
             full_field_dft=False, shutoff=0.0, shutoff_check_interval=100, cuda_graph=True, excitations=None,
             parallel=None, initial_condition=None, adjoint_checkpoint_stride=None) -> "Simulation":
        cfg = FDTDConfig(run_time=run_time or TimeConfig.auto(), absorber=absorber,
                         spectral_sampler=spectral_sampler or SpectralSampler(),
                         cpml_config={} if cpml_config is None else cpml_config, enable_plot=enable_plot,
                         full_field_dft=full_field_dft, shutoff=shutoff, shutoff_check_interval=shutoff_check_interval,
                         cuda_graph=cuda_graph, parallel=parallel, initial_condition=initial_condition,
                         adjoint_checkpoint_stride=adjoint_checkpoint_stride)
        return cls(scene=scene, method="fdtd", frequencies=_normalize_frequencies(frequencies=frequencies,
                                                                                  frequency=frequency),
                   config=cfg, excitations=excitations)

    # ------------------------------------------------------------ checks
    @property
    def has_trainable_parameters(self) -> bool:
        return any(isinstance(r.density, torch.Tensor) and r.density.requires_grad
                   for r in self.scene.material_regions)

    def _validate(self):
        c = self.config
        if c.absorber != "cpml":
            raise NotImplementedError(f"absorber {c.absorber!r} is outside this FDTD runtime (CPML only)")
        if c.parallel is not None:
            raise NotImplementedError("multi-GPU FDTD is outside this single-GPU runtime")
        if c.initial_condition is not None:
            raise NotImplementedError("electrostatic pre-bias initial conditions are outside this FDTD runtime")
        if self.excitations:
            raise NotImplementedError("port excitations are outside this FDTD runtime")
        if c.shutoff > 0:
            raise NotImplementedError("auto-shutoff is outside this FDTD runtime")
        if c.spectral_sampler.normalize_source and len(self.scene.sources) != 1:
            raise ValueError("normalize_source=True requires exactly one source.")
        if c.enable_plot:
            raise NotImplementedError("live plotting is outside this FDTD runtime")

    def _requested_frequencies(self):
        """simulation.py:1501-1520."""
        ordered = list(self.frequencies)
        seen = set(ordered)
        for m in self.scene.monitors:
            for f in (getattr(m, "frequencies", None) or ()):
                if float(f) not in seen:
                    ordered.append(float(f))
                    seen.add(float(f))
        return tuple(ordered)

    def _plan_kwargs(self):
        use_full = self.config.full_field_dft or len(self.frequencies) > 1
        return dict(frequencies=self.frequencies, full_field_dft=use_full,
                    window=self.config.spectral_sampler.window, cpml_config=self.config.cpml_config or None,
                    default_frequency=self.frequencies[0])

    def _resolve_time_steps(self, dt):
        rt = self.config.run_time
        if rt.time_steps is not None:
            return int(rt.time_steps)
        b = self.scene.domain.bounds
        domain_size = max(float(b[i][1] - b[i][0]) for i in range(3))
        from .fdtd.excitation import primary_source_time
        st = primary_source_time(self.scene.sources[0], self.frequencies[0]) if self.scene.sources else None
        return calculate_required_steps(min(self._requested_frequencies()), dt, C_0,
                                        rt.steady_cycles, rt.transient_cycles, domain_size, st)

    # ------------------------------------------------------------ run
    def _compile(self):
        """The run's time step, step count and plan keywords (dt is fixed by the
        scene and frequencies, so the plan is compiled once, with its steps)."""
        from .fdtd.plan import plan_dt
        kw = self._plan_kwargs()
        dt = plan_dt(self.scene, **kw)
        return dt, self._resolve_time_steps(dt), kw

    def run(self) -> Result:
        from .fdtd.plan import compile_plan
        self._refresh_scene()
        self._validate()
        device = getattr(self.scene, "device", "cuda")
        dt, steps, kw = self._compile()
        if self.has_trainable_parameters:
            from .fdtd.adjoint import differentiable_run
            plan, monitors, session = differentiable_run(self.scene, device=device, dt=dt, time_steps=steps,
                                                         stride=self.config.adjoint_checkpoint_stride, **kw)
            return self._result(plan, session.run, monitors, steps)
        from .fdtd.runtime import DeviceRun
        plan = compile_plan(self.scene, time_steps=steps, dt=dt, **kw)
        run = DeviceRun(plan, device)
        run.run()
        return self._result(plan, run, run.monitor_results(), steps)

    def prepare(self) -> "PreparedSimulation":
        """Compile the scene and upload it (simulation.py:590-640): the returned
        ``PreparedSimulation`` runs once, from the start or from a resume
        checkpoint, or advances to a step and returns a checkpoint."""
        from .fdtd.plan import compile_plan
        from .fdtd.runtime import DeviceRun
        self._refresh_scene()
        self._validate()
        device = getattr(self.scene, "device", "cuda")
        dt, steps, kw = self._compile()
        if self.has_trainable_parameters:
            return PreparedSimulation(self, None, steps)
        plan = compile_plan(self.scene, time_steps=steps, dt=dt, **kw)
        return PreparedSimulation(self, DeviceRun(plan, device), steps)

    def _result(self, plan, run, monitors, steps) -> Result:
        device = getattr(self.scene, "device", "cuda")
        fields = {}
        if plan.dft_frequencies:
            from .fdtd.spectral import frequency_solution
            sol = frequency_solution(plan.dft_frequencies,
                                     {c: run.dft_re[c].double().cpu().numpy() for c in ("Ex", "Ey", "Ez")},
                                     {c: run.dft_im[c].double().cpu().numpy() for c in ("Ex", "Ey", "Ez")},
                                     plan.dft_window_sum)
            for c, v in sol.items():
                fields[c] = _to_tensor(v if len(plan.dft_frequencies) > 1 else v[0], device)
        else:
            for c in ("Ex", "Ey", "Ez"):
                fields[c] = run.fields[c].detach().clone()
        if self.config.spectral_sampler.normalize_source:
            _normalize_by_source(plan, self.scene.sources[0], monitors, fields, self.frequencies[0])
        monitors = _tensorize(monitors, device)
        stats = {"time_steps": steps, "dt": plan.dt, "shape": plan.shape, "fused": run.fused,
                 "full_field_dft": bool(plan.dft_frequencies)}
        model = plan.material_model
        materials = {"eps_r": torch.as_tensor(plan.node_eps, device=device),
                     "mu_r": torch.as_tensor(model.mu_r, device=device),
                     "sigma_e": torch.as_tensor(model.sigma_e, device=device)}
        return Result(method="fdtd", scene=self.scene, prepared_scene=plan.prepared, frequency=self.frequencies[0],
                      frequencies=self.frequencies, fields=fields, monitors=monitors, solver_stats=stats,
                      materials=materials)


class PreparedSimulation:
    """A compiled, uploaded FDTD run (simulation.py:1830-1916) that executes once:
    ``run()`` from step 0 or from a resume checkpoint, or ``run_until(step)``,
    which advances exactly ``step`` steps and returns a detached checkpoint that a
    fresh ``prepare()`` can resume (the finished result equals the uninterrupted
    run's)."""

    def __init__(self, simulation: Simulation, run, total_steps: int):
        self.simulation = simulation
        self.solver = run
        self.total_steps = int(total_steps)
        self._consumed = False

    def _claim(self) -> None:
        if self._consumed:
            raise RuntimeError("A PreparedSimulation can execute only once; call Simulation.prepare() "
                               "again for a fresh solver.")
        self._consumed = True

    def _validate_resume_support(self) -> None:
        if self.simulation.has_trainable_parameters:
            raise NotImplementedError("FDTD resume requires a detached forward run; the adjoint replays from "
                                      "its own checkpoints.")

    def run(self, *, resume_from=None) -> Result:
        if resume_from is not None:
            from .fdtd.resume import preflight_resume_checkpoint, restore_resume_checkpoint
            self._validate_resume_support()
            preflight_resume_checkpoint(self.solver, resume_from, total_steps=self.total_steps)
            self._claim()
            restore_resume_checkpoint(self.solver, resume_from, total_steps=self.total_steps)
        else:
            self._claim()
            if self.simulation.has_trainable_parameters:
                return self.simulation.run()
        run = self.solver
        run.run(self.total_steps - run.step_index)
        return self.simulation._result(run.plan, run, run.monitor_results(), self.total_steps)

    def run_until(self, step: int):
        """Advance exactly ``step`` FDTD steps and return a detached resume state."""
        from .fdtd.resume import capture_resume_checkpoint
        self._validate_resume_support()
        if isinstance(step, bool) or not isinstance(step, int):
            raise TypeError("step must be an integer.")
        if not 0 <= step < self.total_steps:
            raise ValueError(f"step must satisfy 0 <= step < {self.total_steps} for this simulation.")
        self._claim()
        self.solver.run(step)
        return capture_resume_checkpoint(self.solver, step=step, total_steps=self.total_steps)


def source_spectrum(plan, source, frequencies, weights, window_sum, default_frequency):
    """Spectrum of the source's waveform envelope at ``frequencies``: the running
    DFT of s(n dt) with amplitude 1 and phase 0, taken with the monitors' own
    weights and scaled by 2 / sum(w) like the monitor sums
    (fdtd/runtime/spectral.py:584-621, sources.py:524-546)."""
    from .fdtd.excitation import primary_source_time
    from .fdtd.plan import signal_value
    st = dict(primary_source_time(source, default_frequency))
    st["amplitude"] = 1.0
    st["phase"] = 0.0
    sig = np.array([signal_value(st, n * plan.dt) for n in range(plan.time_steps)])
    wc, ws = weights
    scale = np.where(window_sum > 0, 2.0 / np.where(window_sum > 0, window_sum, 1.0), 0.0)
    spec = (sig @ wc + 1j * (sig @ ws)) * scale
    # near-zero values are floored at 1e-10 of the peak (observers.py:985-1011)
    peak = float(np.max(np.abs(spec))) if spec.size else 0.0
    floor = 1e-10 * peak if peak > 0 else 1e-30
    spec = np.where(np.abs(spec) > floor, spec, floor + 0j)
    index = {float(f): i for i, f in enumerate(frequencies)}
    return spec, index


def _divide(value, spec):
    if isinstance(value, torch.Tensor):
        s = torch.as_tensor(spec, dtype=torch.complex128, device=value.device)
        return value / s.reshape((-1,) + (1,) * (value.ndim - 1)) if s.ndim and value.ndim else value / s
    v = np.asarray(value)
    return v / (spec.reshape((-1,) + (1,) * (v.ndim - 1)) if np.ndim(spec) and v.ndim else spec)


def _normalize_by_source(plan, source, monitors, fields, default_frequency):
    """Divide monitor spectra by the source spectrum and fluxes by its squared
    magnitude (observers.py:1044-1084; full-field DFT postprocess.py:60-78)."""
    if plan.observers is not None and monitors:
        spec, index = source_spectrum(plan, source, plan.observers.frequencies, plan.obs_weights[:2],
                                      plan.obs_window_sum, default_frequency)
        for name, entry in monitors.items():
            freqs = entry.get("frequencies", (entry.get("frequency"),))
            s = spec[[index[float(f)] for f in freqs]]
            s = s[0] if len(freqs) == 1 else s
            comps = entry.get("components", {})
            for comp in list(comps):
                v = comps[comp]
                if isinstance(v, dict):
                    v["data"] = _divide(v["data"], s)
                else:
                    comps[comp] = _divide(v, s)
            for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
                if comp in entry:
                    entry[comp] = _divide(entry[comp], s)
            if "data" in entry and entry.get("component") in comps:
                v = comps[entry["component"]]
                entry["data"] = v["data"] if isinstance(v, dict) else v
            for key in ("flux", "power"):
                if key in entry:
                    entry[key] = entry[key] / (np.abs(s) ** 2 if not isinstance(entry[key], torch.Tensor)
                                               else torch.as_tensor(np.abs(s) ** 2, device=entry[key].device))
    if plan.dft_frequencies and fields:
        spec, _ = source_spectrum(plan, source, plan.dft_frequencies, plan.dft_weights, plan.dft_window_sum,
                                  default_frequency)
        s = spec if len(plan.dft_frequencies) > 1 else spec[0]
        for c in list(fields):
            fields[c] = _divide(fields[c], s)


def run(simulation: Simulation) -> Result:
    """simulation.py:1953."""
    return simulation.run()
