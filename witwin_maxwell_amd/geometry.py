"""Geometry primitives and ``Structure``.

The reference re-exports ``Box``/``Sphere``/``Structure`` from the absent
``witwin.core`` package (witwin/maxwell/__init__.py; rasterised in
compiler/materials.py:475-505 through ``geometry.to_mask``).  Its occupancy and
subpixel smoothing cannot be reproduced without that package, so this runtime
uses a documented hard-voxel rasteriser: a Yee node belongs to a geometry when
its coordinate lies inside the closed shape.  Later structures override earlier
ones, as in the reference's structure ordering.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np


def _vec3(value, name):
    try:
        seq = [float(v) for v in value]
    except TypeError as exc:  # torch tensors iterate fine; scalars do not
        raise TypeError(f"{name} must be a length-3 sequence") from exc
    if len(seq) != 3:
        raise ValueError(f"{name} must have three components")
    return tuple(seq)


@dataclass(frozen=True)
class Box:
    position: tuple = (0.0, 0.0, 0.0)
    size: tuple = (1.0, 1.0, 1.0)

    def __init__(self, position=(0.0, 0.0, 0.0), size=(1.0, 1.0, 1.0)):
        object.__setattr__(self, "position", _vec3(position, "position"))
        s = _vec3(size, "size")
        if any(v < 0 for v in s):
            raise ValueError("Box size must be non-negative")
        object.__setattr__(self, "size", s)

    @property
    def bounds(self):
        return tuple((p - 0.5 * s, p + 0.5 * s) for p, s in zip(self.position, self.size))

    def contains(self, x, y, z):
        (x0, x1), (y0, y1), (z0, z1) = self.bounds
        return ((x >= x0) & (x <= x1)) & ((y >= y0) & (y <= y1)) & ((z >= z0) & (z <= z1))

    def axis_masks(self, x, y, z):
        """Separable per-axis membership (a box mask is their outer product)."""
        (x0, x1), (y0, y1), (z0, z1) = self.bounds
        return (x >= x0) & (x <= x1), (y >= y0) & (y <= y1), (z >= z0) & (z <= z1)


@dataclass(frozen=True)
class Sphere:
    position: tuple = (0.0, 0.0, 0.0)
    radius: float = 1.0

    def __init__(self, position=(0.0, 0.0, 0.0), radius=1.0):
        object.__setattr__(self, "position", _vec3(position, "position"))
        r = float(radius)
        if r < 0:
            raise ValueError("Sphere radius must be non-negative")
        object.__setattr__(self, "radius", r)

    @property
    def bounds(self):
        return tuple((p - self.radius, p + self.radius) for p in self.position)

    def contains(self, x, y, z):
        cx, cy, cz = self.position
        return (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= self.radius ** 2


@dataclass(frozen=True)
class Structure:
    geometry: Any
    material: Any
    name: str | None = None

    def __init__(self, geometry=None, material=None, name=None):
        if geometry is None or material is None:
            raise ValueError("Structure needs a geometry and a material")
        object.__setattr__(self, "geometry", geometry)
        object.__setattr__(self, "material", material)
        object.__setattr__(self, "name", name)


def rasterize_window(geometry, x, y, z):
    """Node occupancy of ``geometry`` restricted to its bounding window on the
    tensor-product grid (x, y, z): ``(slices, mask)`` with ``mask`` the boolean
    occupancy of the window ``slices``, or ``None`` when the window is empty.
    A small body inside a large grid costs O(window)."""
    (bx0, bx1), (by0, by1), (bz0, bz1) = geometry.bounds
    i0, i1 = np.searchsorted(x, bx0, "left"), np.searchsorted(x, bx1, "right")
    j0, j1 = np.searchsorted(y, by0, "left"), np.searchsorted(y, by1, "right")
    k0, k1 = np.searchsorted(z, bz0, "left"), np.searchsorted(z, bz1, "right")
    if i1 <= i0 or j1 <= j0 or k1 <= k0:
        return None
    xs = x[i0:i1][:, None, None]
    ys = y[j0:j1][None, :, None]
    zs = z[k0:k1][None, None, :]
    mask = np.broadcast_to(geometry.contains(xs, ys, zs), (i1 - i0, j1 - j0, k1 - k0))
    return (slice(i0, i1), slice(j0, j1), slice(k0, k1)), mask


def rasterize_nodes(geometry, x, y, z):
    """Boolean node-occupancy of ``geometry`` on the whole tensor-product grid."""
    out = np.zeros((x.size, y.size, z.size), dtype=bool)
    hit = rasterize_window(geometry, x, y, z)
    if hit is not None:
        out[hit[0]] = hit[1]
    return out


__all__ = ["Box", "Sphere", "Structure", "rasterize_nodes", "rasterize_window"]
