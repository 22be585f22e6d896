"""Frequency-domain monitors (witwin/maxwell/monitors.py:155-255, point/plane/flux)."""

from __future__ import annotations

from dataclasses import dataclass

_FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
_FLUX_FIELDS = {"x": ("Ey", "Ez", "Hy", "Hz"), "y": ("Ex", "Ez", "Hx", "Hz"), "z": ("Ex", "Ey", "Hx", "Hy")}


def normalize_component(component: str) -> str:
    c = str(component).strip()
    if len(c) != 2 or c[0].upper() not in "EH" or c[1].lower() not in "xyz":
        raise ValueError(f"Unsupported field component: {component!r}")
    return c[0].upper() + c[1].lower()


def normalize_axis(axis: str) -> str:
    a = str(axis).strip().lower()
    if a not in ("x", "y", "z"):
        raise ValueError("axis must be 'x', 'y' or 'z'")
    return a


def _fields(fields):
    if isinstance(fields, str):
        fields = (fields,)
    out = tuple(normalize_component(f) for f in fields)
    if not out:
        raise ValueError("at least one field is required")
    return out


def _freqs(frequencies):
    if frequencies is None:
        return None
    if isinstance(frequencies, (int, float)):
        return (float(frequencies),)
    vals = tuple(float(f) for f in frequencies)
    if not vals or any(f <= 0 for f in vals):
        raise ValueError("monitor frequencies must be positive")
    return vals


def required_flux_fields(axis):
    return _FLUX_FIELDS[normalize_axis(axis)]


@dataclass(frozen=True)
class PointMonitor:
    name: str
    position: tuple
    fields: tuple
    frequencies: tuple | None = None
    kind: str = "point"

    def __init__(self, name, position, fields=("Ez",), frequencies=None):
        pos = tuple(float(v) for v in position)
        if len(pos) != 3:
            raise ValueError("position must contain exactly three values.")
        object.__setattr__(self, "name", str(name))
        object.__setattr__(self, "position", pos)
        object.__setattr__(self, "fields", _fields(fields))
        object.__setattr__(self, "frequencies", _freqs(frequencies))
        object.__setattr__(self, "kind", "point")


@dataclass(frozen=True)
class PlaneMonitor:
    name: str
    axis: str
    position: float
    fields: tuple
    frequencies: tuple | None = None
    compute_flux: bool = False
    normal_direction: str = "+"
    kind: str = "plane"

    def __init__(self, name, axis="z", position=0.0, fields=("Ez",), frequencies=None, compute_flux=False,
                 normal_direction="+"):
        ax = normalize_axis(axis)
        fl = _fields(fields)
        if compute_flux:
            missing = [f for f in _FLUX_FIELDS[ax] if f not in fl]
            if missing:
                raise ValueError(f"flux on a {ax}-normal plane needs fields {_FLUX_FIELDS[ax]}")
        if normal_direction not in ("+", "-"):
            raise ValueError("normal_direction must be '+' or '-'")
        object.__setattr__(self, "name", str(name))
        object.__setattr__(self, "axis", ax)
        object.__setattr__(self, "position", float(position))
        object.__setattr__(self, "fields", fl)
        object.__setattr__(self, "frequencies", _freqs(frequencies))
        object.__setattr__(self, "compute_flux", bool(compute_flux))
        object.__setattr__(self, "normal_direction", normal_direction)
        object.__setattr__(self, "kind", "plane")


class FluxMonitor(PlaneMonitor):
    def __init__(self, name, axis="z", position=0.0, frequencies=None, normal_direction="+"):
        super().__init__(name=name, axis=axis, position=position, fields=required_flux_fields(axis),
                         frequencies=frequencies, compute_flux=True, normal_direction=normal_direction)


__all__ = ["PointMonitor", "PlaneMonitor", "FluxMonitor", "normalize_component", "normalize_axis",
           "required_flux_fields"]
