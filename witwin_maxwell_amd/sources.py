"""Sources and source-time waveforms.

``GaussianPulse``/``CW`` follow witwin/maxwell/sources.py:200-250 and the
formulas of witwin/maxwell/waveforms.py:30-53 (sigma_t = 1/(2*pi*fwidth), default
delay 6*sigma_t, characteristic frequency f0 + 3*fwidth).  ``PointDipole`` and
``PlaneWave`` follow sources.py:549+ and ``TFSF`` the slab form used by
fdtd/excitation/tfsf_*.py.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

POINT_DIPOLE_REFERENCE_WIDTH = 0.02  # metres, sources.py:21

KIND_CODE = {"cw": 0, "gaussian_pulse": 1, "ricker_wavelet": 2}


def gaussian_sigma_t(fwidth: float) -> float:
    return 1.0 / (2.0 * math.pi * fwidth)


def gaussian_pulse_signal(t, *, frequency, fwidth, amplitude, phase, delay):
    """A exp(-tau^2/(2 sigma^2)) cos(2 pi f tau + phi), waveforms.py:40-53."""
    sigma_t = gaussian_sigma_t(fwidth)
    tau = float(t) - float(delay)
    return amplitude * math.exp(-0.5 * (tau / sigma_t) ** 2) * math.cos(2.0 * math.pi * frequency * tau + phase)


def cw_signal(t, *, frequency, amplitude, phase):
    return amplitude * math.cos(2.0 * math.pi * frequency * float(t) + phase)


@dataclass(frozen=True)
class GaussianPulse:
    frequency: float
    fwidth: float
    amplitude: float = 1.0
    phase: float = 0.0
    delay: float | None = None
    kind: str = "gaussian_pulse"

    def __post_init__(self):
        f = float(self.frequency)
        fw = float(self.fwidth)
        if f <= 0 or fw <= 0:
            raise ValueError("GaussianPulse frequency and fwidth must be > 0")
        object.__setattr__(self, "frequency", f)
        object.__setattr__(self, "fwidth", fw)
        object.__setattr__(self, "amplitude", float(self.amplitude))
        object.__setattr__(self, "phase", float(self.phase))
        object.__setattr__(self, "delay", 6.0 * gaussian_sigma_t(fw) if self.delay is None else float(self.delay))
        object.__setattr__(self, "kind", "gaussian_pulse")

    @property
    def sigma_t(self):
        return gaussian_sigma_t(self.fwidth)

    def evaluate(self, t: float) -> float:
        return gaussian_pulse_signal(t, frequency=self.frequency, fwidth=self.fwidth, amplitude=self.amplitude,
                                     phase=self.phase, delay=self.delay)

    @property
    def characteristic_frequency(self) -> float:
        return max(self.frequency, self.frequency + 3.0 * self.fwidth)


@dataclass(frozen=True)
class CW:
    frequency: float
    amplitude: float = 1.0
    phase: float = 0.0
    kind: str = "cw"

    def __post_init__(self):
        if float(self.frequency) <= 0:
            raise ValueError("CW frequency must be > 0")
        object.__setattr__(self, "frequency", float(self.frequency))
        object.__setattr__(self, "amplitude", float(self.amplitude))
        object.__setattr__(self, "phase", float(self.phase))

    @property
    def delay(self):
        return 0.0

    @property
    def fwidth(self):
        return 0.0

    def evaluate(self, t: float) -> float:
        return cw_signal(t, frequency=self.frequency, amplitude=self.amplitude, phase=self.phase)

    @property
    def characteristic_frequency(self) -> float:
        return self.frequency


def ricker_wavelet_signal(t, *, frequency, amplitude, delay):
    """A (1 - 2 a^2) exp(-a^2), a = pi f (t - delay) (waveforms.py:56-61)."""
    a = math.pi * frequency * (float(t) - float(delay))
    return amplitude * (1.0 - 2.0 * a * a) * math.exp(-a * a)


@dataclass(frozen=True)
class RickerWavelet:
    """Mexican-hat wavelet; default delay 4/(pi f), characteristic frequency 3f
    (sources.py:254-283)."""
    frequency: float
    amplitude: float = 1.0
    delay: float | None = None
    kind: str = "ricker_wavelet"

    def __post_init__(self):
        f = float(self.frequency)
        if f <= 0:
            raise ValueError("RickerWavelet frequency must be > 0")
        object.__setattr__(self, "frequency", f)
        object.__setattr__(self, "amplitude", float(self.amplitude))
        object.__setattr__(self, "delay", 4.0 / (math.pi * f) if self.delay is None else float(self.delay))

    @property
    def phase(self):
        return 0.0

    @property
    def fwidth(self):
        return 0.0

    def evaluate(self, t: float) -> float:
        return ricker_wavelet_signal(t, frequency=self.frequency, amplitude=self.amplitude, delay=self.delay)

    @property
    def characteristic_frequency(self) -> float:
        return 3.0 * self.frequency


def source_time_dict(st) -> dict:
    """Compiled source-time record, the fields the device waveform needs."""
    return {"kind": st.kind, "kind_code": KIND_CODE[st.kind], "frequency": float(st.frequency),
            "fwidth": float(getattr(st, "fwidth", 0.0)), "amplitude": float(st.amplitude),
            "phase": float(getattr(st, "phase", 0.0)), "delay": float(st.delay),
            "characteristic_frequency": float(st.characteristic_frequency),
            "settling_time": float(settling_time(st))}


def settling_time(st) -> float:
    """Time after which the waveform has settled (sources.py:205, :249, :284):
    CW 0, Gaussian delay + 6 sigma_t, Ricker delay + 4/f."""
    if st.kind == "gaussian_pulse":
        return float(st.delay) + 6.0 * gaussian_sigma_t(st.fwidth)
    if st.kind == "ricker_wavelet":
        return float(st.delay) + 4.0 / max(float(st.frequency), 1e-30)
    return 0.0


def polarization_vector(polarization):
    if isinstance(polarization, str):
        table = {"ex": (1.0, 0.0, 0.0), "ey": (0.0, 1.0, 0.0), "ez": (0.0, 0.0, 1.0),
                 "x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0), "z": (0.0, 0.0, 1.0)}
        key = polarization.strip().lower()
        if key not in table:
            raise ValueError(f"Unknown polarization {polarization!r}")
        return table[key]
    vec = tuple(float(v) for v in polarization)
    if len(vec) != 3:
        raise ValueError("polarization must have three components")
    return vec


@dataclass(frozen=True)
class PointDipole:
    position: tuple
    polarization: tuple
    width: float = POINT_DIPOLE_REFERENCE_WIDTH
    profile: str = "gaussian"
    source_time: object = None
    name: str | None = None
    kind: str = "point_dipole"

    def __init__(self, position, polarization="Ez", width=POINT_DIPOLE_REFERENCE_WIDTH, profile="gaussian",
                 source_time=None, name=None):
        pos = tuple(float(v) for v in position)
        if len(pos) != 3:
            raise ValueError("position must have three components")
        prof = str(profile).lower()
        if prof not in ("gaussian", "ideal"):
            raise ValueError("PointDipole profile must be 'gaussian' or 'ideal'")
        if float(width) <= 0:
            raise ValueError("width must be > 0.")
        object.__setattr__(self, "position", pos)
        object.__setattr__(self, "polarization", polarization_vector(polarization))
        object.__setattr__(self, "width", float(width))
        object.__setattr__(self, "profile", prof)
        object.__setattr__(self, "source_time", source_time)
        object.__setattr__(self, "name", name)
        object.__setattr__(self, "kind", "point_dipole")


class TFSF:
    """Total-field/scattered-field injection region (sources.py:388-429).

    ``TFSF(bounds)`` is a box ((x0, x1), (y0, y1), (z0, z1)); ``TFSF.slab(axis=...,
    bounds=(s0, s1))`` keeps only the two faces normal to ``axis`` and spans the
    whole non-PML transverse extent."""

    kind = "tfsf"

    def __init__(self, bounds):
        if len(bounds) != 3:
            raise ValueError("bounds must contain x, y, z ranges.")
        norm = []
        for ab in bounds:
            if len(ab) != 2:
                raise ValueError("Each TFSF axis bound must contain exactly two values.")
            a, b = float(ab[0]), float(ab[1])
            if b <= a:
                raise ValueError("Each TFSF axis bound must have end > start.")
            norm.append((a, b))
        self.bounds = tuple(norm)
        self.mode = "box"
        self.axis = None
        self.axis_bounds = None

    @classmethod
    def slab(cls, *, axis, bounds):
        axis = str(axis).lower()
        if axis not in ("x", "y", "z"):
            raise ValueError("TFSF slab axis must be 'x', 'y', or 'z'.")
        if len(bounds) != 2:
            raise ValueError("TFSF slab bounds must contain exactly two values.")
        a, b = float(bounds[0]), float(bounds[1])
        if b <= a:
            raise ValueError("TFSF slab bounds must have end > start.")
        src = cls.__new__(cls)
        src.bounds = None
        src.mode = "slab"
        src.axis = axis
        src.axis_bounds = (a, b)
        return src


def _normalize(v):
    m = math.sqrt(sum(c * c for c in v))
    if m <= 1e-12:
        raise ValueError("vector must be non-zero")
    return tuple(c / m for c in v)


@dataclass(frozen=True)
class PlaneWave:
    direction: tuple
    polarization: tuple
    source_time: object = None
    injection: object = "soft"
    injection_axis: str | None = None
    name: str | None = None
    kind: str = "plane_wave"

    def __init__(self, direction=(0.0, 0.0, 1.0), polarization=(1.0, 0.0, 0.0), source_time=None,
                 injection="soft", injection_axis=None, name=None):
        d = _normalize(tuple(float(v) for v in direction))
        p = _normalize(tuple(float(v) for v in polarization))
        if abs(sum(a * b for a, b in zip(d, p))) > 1e-6:
            raise ValueError("PlaneWave polarization must be transverse to direction")
        if not isinstance(injection, TFSF) and str(injection).lower() != "soft":
            raise ValueError("PlaneWave injection must be 'soft' or TFSF(...).")
        object.__setattr__(self, "direction", d)
        object.__setattr__(self, "polarization", p)
        object.__setattr__(self, "source_time", source_time)
        object.__setattr__(self, "injection", injection if isinstance(injection, TFSF) else "soft")
        object.__setattr__(self, "injection_axis", None if injection_axis is None else str(injection_axis).lower())
        object.__setattr__(self, "name", name)
        object.__setattr__(self, "kind", "plane_wave")


__all__ = ["GaussianPulse", "CW", "PointDipole", "PlaneWave", "TFSF", "POINT_DIPOLE_REFERENCE_WIDTH",
           "gaussian_pulse_signal", "cw_signal", "source_time_dict", "polarization_vector"]
