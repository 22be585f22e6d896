"""Scene description and the solver-ready grid.

Public types mirror witwin/maxwell/scene.py: ``Domain`` (:70-110), ``GridSpec``
(:174-290), ``BoundarySpec`` (:293-565), ``MaterialRegion`` (:567-610),
``SceneModule`` (:613-615) and ``Scene`` (:748-1212).  ``PreparedScene`` builds the
float64 Yee node/half/primal/dual arrays with the PML extension exactly as
``_build_axis_grid64`` / ``PreparedScene.__init__`` do (scene.py:1216-1395).

Only the subset the FDTD hot path needs is provided: uniform, anisotropic and
custom grids; none/pml/periodic/pec/pmc faces; boxes/spheres with isotropic and
dispersive media; density-parameterised design regions.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

_AXES = ("x", "y", "z")
_AXIS_INDEX = {"x": 0, "y": 1, "z": 2}
_BOUNDARY_KINDS = ("none", "pml", "periodic", "pec", "pmc")


@dataclass(frozen=True)
class Domain:
    """Axis-aligned physical simulation box (scene.py:70-110)."""

    bounds: tuple

    def __init__(self, bounds):
        if len(bounds) != 3:
            raise ValueError("bounds must contain x, y, z ranges.")
        normalized = []
        for axis_bounds in bounds:
            if len(axis_bounds) != 2:
                raise ValueError("Each axis bound must contain exactly two values.")
            start, end = float(axis_bounds[0]), float(axis_bounds[1])
            if end <= start:
                raise ValueError("Each axis bound must have end > start.")
            normalized.append((start, end))
        object.__setattr__(self, "bounds", tuple(normalized))

    @classmethod
    def from_domain_range(cls, domain_range):
        if len(domain_range) != 6:
            raise ValueError("domain_range must contain six values.")
        return cls(((domain_range[0], domain_range[1]), (domain_range[2], domain_range[3]),
                    (domain_range[4], domain_range[5])))

    @property
    def domain_range(self):
        b = self.bounds
        return (b[0][0], b[0][1], b[1][0], b[1][1], b[2][0], b[2][1])

    @property
    def size(self):
        return tuple(hi - lo for lo, hi in self.bounds)

    @property
    def center(self):
        return tuple(0.5 * (lo + hi) for lo, hi in self.bounds)


@dataclass(frozen=True)
class GridSpec:
    """Grid spacing specification (scene.py:174-290); auto-meshing is not provided."""

    dx: float | None
    dy: float | None
    dz: float | None
    x_coords: np.ndarray | None = None
    y_coords: np.ndarray | None = None
    z_coords: np.ndarray | None = None

    @classmethod
    def uniform(cls, dl: float) -> "GridSpec":
        step = float(dl)
        if step <= 0:
            raise ValueError("GridSpec.uniform spacing must be > 0.")
        return cls(step, step, step)

    @classmethod
    def anisotropic(cls, dx: float, dy: float, dz: float) -> "GridSpec":
        return cls(float(dx), float(dy), float(dz))

    @classmethod
    def custom(cls, x_coords, y_coords, z_coords) -> "GridSpec":
        def _check(c, name):
            arr = np.asarray(c, dtype=np.float64).reshape(-1)
            if arr.size < 2 or np.any(np.diff(arr) <= 0):
                raise ValueError(f"GridSpec.custom {name}_coords must be strictly increasing with >= 2 values.")
            return arr
        return cls(None, None, None, x_coords=_check(x_coords, "x"), y_coords=_check(y_coords, "y"),
                   z_coords=_check(z_coords, "z"))

    @property
    def is_custom(self) -> bool:
        return self.x_coords is not None

    @property
    def is_auto(self) -> bool:
        return False

    @property
    def spacing(self):
        if self.is_custom:
            raise ValueError("GridSpec.custom has no scalar spacing.")
        return (self.dx, self.dy, self.dz)

    def axis_coords(self, axis: str):
        return (self.x_coords, self.y_coords, self.z_coords)[_AXIS_INDEX[axis]]


def _normalize_kind(kind) -> str:
    k = str(kind).strip().lower()
    if k not in _BOUNDARY_KINDS:
        raise ValueError(f"Unsupported boundary kind {kind!r}; expected one of {_BOUNDARY_KINDS}.")
    return k


@dataclass(frozen=True)
class BoundarySpec:
    """Per-face boundary kinds (scene.py:293-565).

    ``pml(num_layers, strength)`` gives CPML on all six faces; ``faces(...)`` sets
    faces individually.  Bloch and Mur faces are not part of this runtime.
    """

    kind: str = "none"
    num_layers: int = 0
    strength: float = 0.0
    x_low: str | None = None
    x_high: str | None = None
    y_low: str | None = None
    y_high: str | None = None
    z_low: str | None = None
    z_high: str | None = None
    _face_kinds: tuple = field(init=False, repr=False, default=())

    def __post_init__(self):
        if self.num_layers < 0:
            raise ValueError("BoundarySpec.num_layers must be >= 0.")
        if self.strength < 0:
            raise ValueError("BoundarySpec.strength must be >= 0.")
        default = _normalize_kind(self.kind) if self.kind != "mixed" else "none"
        faces = []
        for axis in _AXES:
            lo = getattr(self, f"{axis}_low")
            hi = getattr(self, f"{axis}_high")
            lo = default if lo is None else _normalize_kind(lo)
            hi = default if hi is None else _normalize_kind(hi)
            if (lo == "periodic") != (hi == "periodic"):
                raise ValueError("periodic boundaries must be set on both faces of an axis.")
            faces.append((lo, hi))
        object.__setattr__(self, "_face_kinds", tuple(faces))

    @classmethod
    def pml(cls, num_layers: int = 15, strength: float = 1e6) -> "BoundarySpec":
        return cls(kind="pml", num_layers=int(num_layers), strength=float(strength))

    @classmethod
    def none(cls) -> "BoundarySpec":
        return cls(kind="none")

    @classmethod
    def periodic(cls) -> "BoundarySpec":
        return cls(kind="periodic")

    @classmethod
    def pec(cls) -> "BoundarySpec":
        return cls(kind="pec")

    @classmethod
    def pmc(cls) -> "BoundarySpec":
        return cls(kind="pmc")

    @classmethod
    def faces(cls, *, default="none", num_layers=0, strength=0.0, x=None, y=None, z=None,
              x_low=None, x_high=None, y_low=None, y_high=None, z_low=None, z_high=None) -> "BoundarySpec":
        def _pair(v):
            if v is None:
                return (None, None)
            if isinstance(v, str):
                return (v, v)
            return (v[0], v[1])
        xl, xh = _pair(x)
        yl, yh = _pair(y)
        zl, zh = _pair(z)
        return cls(kind=default, num_layers=int(num_layers), strength=float(strength),
                   x_low=x_low or xl, x_high=x_high or xh, y_low=y_low or yl, y_high=y_high or yh,
                   z_low=z_low or zl, z_high=z_high or zh)

    def face_kinds(self):
        return self._face_kinds

    def axis_face_kinds(self, axis: str):
        return self._face_kinds[_AXIS_INDEX[axis]]

    def axis_kind(self, axis: str) -> str:
        lo, hi = self.axis_face_kinds(axis)
        return lo if lo == hi else "mixed"

    def face_kind(self, axis: str, side: str) -> str:
        return self._face_kinds[_AXIS_INDEX[axis]][0 if side == "low" else 1]

    def uses_kind(self, kind: str) -> bool:
        return any(k == kind for pair in self._face_kinds for k in pair)

    def pml_layers_for_face(self, axis: str, side: str) -> int:
        return int(self.num_layers) if self.face_kind(axis, side) == "pml" else 0


@dataclass(frozen=True)
class MaterialRegion:
    """Density-parameterised design region (scene.py:567-610).

    ``density`` is a (Nx, Ny, Nz) torch tensor sampled trilinearly inside
    ``geometry`` (a ``Box``); permittivity is ``eps_lo + rho*(eps_hi-eps_lo)``.
    """

    name: str
    geometry: Any
    density: torch.Tensor
    basis: str = "density"
    bounds: tuple = (0.0, 1.0)
    eps_bounds: tuple = (1.0, 1.0)
    mu_bounds: tuple = (1.0, 1.0)
    filter_radius: float | None = None
    projection_beta: float | None = None
    symmetry: str | None = None

    def __post_init__(self):
        if str(self.basis).lower() != "density":
            raise ValueError("MaterialRegion currently supports basis='density' only.")
        if not isinstance(self.density, torch.Tensor):
            raise TypeError("MaterialRegion.density must be a torch.Tensor.")
        if self.density.ndim != 3:
            raise ValueError("MaterialRegion.density must have shape (Nx, Ny, Nz).")
        b = (float(self.bounds[0]), float(self.bounds[1]))
        if b[1] <= b[0]:
            raise ValueError("MaterialRegion.bounds must satisfy upper > lower.")
        object.__setattr__(self, "bounds", b)
        object.__setattr__(self, "eps_bounds", (float(self.eps_bounds[0]), float(self.eps_bounds[1])))
        object.__setattr__(self, "mu_bounds", (float(self.mu_bounds[0]), float(self.mu_bounds[1])))
        if self.filter_radius is not None and float(self.filter_radius) <= 0.0:
            raise ValueError("MaterialRegion.filter_radius must be > 0 when provided.")
        if self.projection_beta is not None and float(self.projection_beta) <= 0.0:
            raise ValueError("MaterialRegion.projection_beta must be > 0 when provided.")


class SceneModule(torch.nn.Module):
    """Trainable scene factory (scene.py:613-615): subclasses implement ``to_scene``."""

    def to_scene(self) -> "Scene":
        raise NotImplementedError("SceneModule subclasses must implement to_scene().")


class Scene:
    """Declarative 3-D scene (scene.py:748-1212), FDTD subset."""

    def __init__(self, domain: Domain | None = None, grid: GridSpec | None = None,
                 boundary: BoundarySpec | None = None, *, structures=None, sources=None, monitors=None,
                 material_regions=None, metadata=None, device="cuda", verbose=False, **unsupported):
        for key, value in unsupported.items():
            if value:
                raise NotImplementedError(f"Scene argument {key!r} is outside this FDTD runtime.")
        from .geometry import Structure

        self.structures = list(structures or [])
        if any(not isinstance(s, Structure) for s in self.structures):
            raise TypeError("Scene structures must be Structure instances.")
        self.sources = list(sources or [])
        self.monitors = list(monitors or [])
        self.material_regions = list(material_regions or [])
        self.metadata = dict(metadata or {})
        self.device = str(torch.device(device))
        self.verbose = bool(verbose)
        self.domain = domain if domain is not None else Domain.from_domain_range((-1, 1, -1, 1, -1, 1))
        self.grid = grid if grid is not None else GridSpec.uniform(0.05)
        self.boundary = boundary if boundary is not None else BoundarySpec.none()

    # builder helpers, as in the reference
    def add_structure(self, structure):
        self.structures.append(structure)
        return self

    def add_source(self, source):
        self.sources.append(source)
        return self

    def add_monitor(self, monitor):
        self.monitors.append(monitor)
        return self

    def add_material_region(self, region):
        self.material_regions.append(region)
        return self

    @property
    def pml_thickness(self) -> int:
        return int(self.boundary.num_layers) if self.boundary.uses_kind("pml") else 0

    @property
    def pml_strength(self) -> float:
        return float(self.boundary.strength) if self.boundary.uses_kind("pml") else 0.0

    def pml_thickness_for_face(self, axis: str, side: str) -> int:
        return self.boundary.pml_layers_for_face(axis, side)

    @property
    def domain_size(self) -> float:
        return max(hi - lo for lo, hi in self.domain.bounds)


def _axis_nodes64(grid: GridSpec, axis: str, lo: float, hi: float) -> np.ndarray:
    """Endpoint-inclusive node coordinates (scene.py:1216-1242)."""
    coords = grid.axis_coords(axis)
    if coords is None:
        step = float(getattr(grid, f"d{axis}"))
        count = max(1, int(math.ceil((float(hi) - float(lo)) / step)))
        return np.linspace(float(lo), float(hi), count + 1, endpoint=True, dtype=np.float64)
    span = float(hi) - float(lo)
    tol = 1e-9 * span
    if abs(coords[0] - lo) > tol or abs(coords[-1] - hi) > tol:
        raise ValueError(f"GridSpec.custom {axis}_coords must span the domain exactly.")
    return np.asarray(coords, dtype=np.float64)


def _axis_dual_spacing64(primal: np.ndarray, kind: str) -> np.ndarray:
    """Dual spacing of length N (scene.py:1245-1257)."""
    dual = np.empty(primal.size + 1, dtype=np.float64)
    dual[1:-1] = 0.5 * (primal[:-1] + primal[1:])
    if kind == "periodic":
        wrap = 0.5 * (primal[0] + primal[-1])
        dual[0] = wrap
        dual[-1] = wrap
    else:
        dual[0] = primal[0]
        dual[-1] = primal[-1]
    return dual


def build_axis_grid64(grid: GridSpec, axis: str, lo: float, hi: float, kind: str, pml_layers=(0, 0)):
    """Nodes (with PML extension), primal spacing, half points and dual spacing
    (scene.py:1260-1281)."""
    nodes = _axis_nodes64(grid, axis, lo, hi)
    low_layers, high_layers = (int(v) for v in pml_layers)
    if low_layers or high_layers:
        low_step = float(nodes[1] - nodes[0])
        high_step = float(nodes[-1] - nodes[-2])
        low_nodes = nodes[0] - low_step * np.arange(low_layers, 0, -1, dtype=np.float64)
        high_nodes = nodes[-1] + high_step * np.arange(1, high_layers + 1, dtype=np.float64)
        nodes = np.concatenate((low_nodes, nodes, high_nodes))
    primal = np.diff(nodes)
    half = nodes[:-1] + 0.5 * primal
    dual = _axis_dual_spacing64(primal, kind)
    return nodes, primal, half, dual


class PreparedScene:
    """Solver-ready grid state (scene.py:1284-1395)."""

    def __init__(self, scene: Scene):
        self.scene = scene
        self.domain = scene.domain
        self.boundary = scene.boundary
        self.physical_domain_range = scene.domain.domain_range
        self.pml_layers = {}
        axes = {}
        for idx, axis in enumerate(_AXES):
            lo, hi = scene.domain.bounds[idx]
            layers = (scene.pml_thickness_for_face(axis, "low"), scene.pml_thickness_for_face(axis, "high"))
            self.pml_layers[axis] = layers
            axes[axis] = build_axis_grid64(scene.grid, axis, lo, hi, scene.boundary.axis_kind(axis), layers)
        self.x_nodes64, self.dx_primal64, self.x_half64, self.dx_dual64 = axes["x"]
        self.y_nodes64, self.dy_primal64, self.y_half64, self.dy_dual64 = axes["y"]
        self.z_nodes64, self.dz_primal64, self.z_half64, self.dz_dual64 = axes["z"]
        bounds = []
        for idx, axis in enumerate(_AXES):
            primal = axes[axis][1]
            lo_layers, hi_layers = self.pml_layers[axis]
            plo, phi = scene.domain.bounds[idx]
            low_step = float(primal[lo_layers]) if lo_layers < primal.size else float(primal[0])
            high_step = float(primal[max(primal.size - hi_layers - 1, 0)])
            bounds.append((plo - lo_layers * low_step, phi + hi_layers * high_step))
        self.domain_range = (bounds[0][0], bounds[0][1], bounds[1][0], bounds[1][1], bounds[2][0], bounds[2][1])
        self.Nx = len(self.x_nodes64)
        self.Ny = len(self.y_nodes64)
        self.Nz = len(self.z_nodes64)
        self.pml_thickness = scene.pml_thickness
        self.pml_strength = scene.pml_strength

    def nodes(self, axis):
        return {"x": self.x_nodes64, "y": self.y_nodes64, "z": self.z_nodes64}[axis]

    def half(self, axis):
        return {"x": self.x_half64, "y": self.y_half64, "z": self.z_half64}[axis]

    def primal(self, axis):
        return {"x": self.dx_primal64, "y": self.dy_primal64, "z": self.dz_primal64}[axis]

    def dual(self, axis):
        return {"x": self.dx_dual64, "y": self.dy_dual64, "z": self.dz_dual64}[axis]

    @property
    def shape(self):
        return (self.Nx, self.Ny, self.Nz)

    def component_coords(self, component: str):
        """Yee sample coordinates per component (fdtd/coords.py:9-27)."""
        c = component.capitalize() if len(component) == 2 else component
        c = c[0].upper() + c[1].lower()
        x, y, z = self.x_nodes64, self.y_nodes64, self.z_nodes64
        xh, yh, zh = self.x_half64, self.y_half64, self.z_half64
        table = {"Ex": (xh, y, z), "Ey": (x, yh, z), "Ez": (x, y, zh),
                 "Hx": (x, yh, zh), "Hy": (xh, y, zh), "Hz": (xh, yh, z)}
        if c not in table:
            raise ValueError(f"Unsupported field component: {component!r}")
        return table[c]


def prepare_scene(scene) -> PreparedScene:
    if isinstance(scene, PreparedScene):
        return scene
    return PreparedScene(scene)


__all__ = ["Domain", "GridSpec", "BoundarySpec", "MaterialRegion", "SceneModule", "Scene",
           "PreparedScene", "prepare_scene", "build_axis_grid64"]
