"""Material compilation: node permittivity, Yee-edge averaging, update
coefficients and ADE pole templates.

Follows
  * node->edge averaging ``average_node_to_component`` (fdtd/runtime/materials.py:399-407)
    and ``average_node_to_magnetic_component`` (:420-456);
  * semi-implicit lossy coefficients ``_electric_update_coefficients`` /
    ``_magnetic_update_coefficients`` (:942-1008) with CPML's ``pml_decay=None``
    (:1577-1594);
  * ADE constants ``_build_family_dispersive_templates`` (:481-600);
  * ``MaterialRegion`` density preparation, trilinear sampling and blending
    (compiler/materials.py:1310-1470).

All host arithmetic is float64; the device receives float32 copies.  Structure
occupancy is the hard-voxel node rasteriser of ``geometry.rasterize_window`` (the
reference's occupancy lives in the absent ``witwin.core``).
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

from ..constants import EPSILON_0, MU_0
from ..geometry import Box, rasterize_window

E_COMPONENTS = ("Ex", "Ey", "Ez")
H_COMPONENTS = ("Hx", "Hy", "Hz")


def _threaded(fn):
    """Run an averaging rule through torch's multi-threaded CPU kernels when the
    input is a large numpy array (same float64 arithmetic, same result)."""
    def wrapper(node, component):
        if isinstance(node, np.ndarray) and node.size >= (1 << 20):
            return fn(torch.from_numpy(np.ascontiguousarray(node)), component).numpy()
        return fn(node, component)
    wrapper.__doc__ = fn.__doc__
    wrapper.__name__ = fn.__name__
    return wrapper


@_threaded
def average_node_to_component(node, component):
    """Two-node average onto the E edge (materials.py:399-407)."""
    if component == "Ex":
        return 0.5 * (node[:-1, :, :] + node[1:, :, :])
    if component == "Ey":
        return 0.5 * (node[:, :-1, :] + node[:, 1:, :])
    if component == "Ez":
        return 0.5 * (node[:, :, :-1] + node[:, :, 1:])
    raise ValueError(component)


@_threaded
def average_node_to_magnetic_component(node, component):
    """Four-node average onto the H face (materials.py:420-456)."""
    if component == "Hx":
        return 0.25 * (node[:, :-1, :-1] + node[:, :-1, 1:] + node[:, 1:, :-1] + node[:, 1:, 1:])
    if component == "Hy":
        return 0.25 * (node[:-1, :, :-1] + node[:-1, :, 1:] + node[1:, :, :-1] + node[1:, :, 1:])
    if component == "Hz":
        return 0.25 * (node[:-1, :-1, :] + node[:-1, 1:, :] + node[1:, :-1, :] + node[1:, 1:, :])
    raise ValueError(component)


def component_shape(component, shape):
    nx, ny, nz = shape
    return {"Ex": (nx - 1, ny, nz), "Ey": (nx, ny - 1, nz), "Ez": (nx, ny, nz - 1),
            "Hx": (nx, ny - 1, nz - 1), "Hy": (nx - 1, ny, nz - 1), "Hz": (nx - 1, ny - 1, nz)}[component]


class NodeMaterialModel:
    """Node-sampled eps_r, mu_r, sigma_e and pole weights for a prepared scene."""

    def __init__(self, prepared):
        self.prepared = prepared
        shape = prepared.shape
        self.eps_r = np.ones(shape)
        self.mu_r = np.ones(shape)
        self.sigma_e = np.zeros(shape)
        self.poles = []  # list of (kind, pole, node weight array)
        self.region_mask = np.zeros(shape, dtype=bool)
        x, y, z = prepared.x_nodes64, prepared.y_nodes64, prepared.z_nodes64
        pole_index = {}
        for structure in prepared.scene.structures:
            hit = rasterize_window(structure.geometry, x, y, z)
            if hit is None:
                continue
            win, mask = hit
            if not mask.any():
                continue
            m = structure.material
            # every assignment touches the structure's bounding window only
            self.eps_r[win][mask] = m.eps_r
            self.mu_r[win][mask] = m.mu_r
            self.sigma_e[win][mask] = m.sigma_e
            # a later structure replaces whatever poles an earlier one put on these nodes
            for _, _, weight in self.poles:
                weight[win][mask] = 0.0
            for kind, plist in (("debye", m.debye_poles), ("drude", m.drude_poles), ("lorentz", m.lorentz_poles)):
                for pole in plist:
                    key = (kind, pole)
                    if key not in pole_index:
                        pole_index[key] = len(self.poles)
                        self.poles.append((kind, pole, np.zeros(shape)))
                    self.poles[pole_index[key]][2][win][mask] = 1.0
        self.poles = [p for p in self.poles if p[2].any()]

    @property
    def dispersive(self):
        return bool(self.poles)


def prepare_region_density(prepared, region, density=None) -> torch.Tensor:
    """Normalise, filter and project a region density on its design grid
    (compiler/materials.py:1310-1360), on the density's device in float64.
    ``density`` defaults to ``region.density`` (the design variable)."""
    if not isinstance(region.geometry, Box):
        raise ValueError("MaterialRegion currently supports Box geometry only.")
    density = (region.density if density is None else density).to(dtype=torch.float64)
    lo, hi = region.bounds
    if not np.isclose(lo, 0.0) or not np.isclose(hi, 1.0):
        density = (density - lo) / max(hi - lo, 1e-12)
    density = density.clamp(0.0, 1.0)
    if region.filter_radius is not None:
        spacing = [float(np.min(prepared.primal(a))) for a in "xyz"]
        kernel = tuple(max(1, 2 * int(np.ceil(region.filter_radius / s)) + 1) for s in spacing)
        kernel = tuple(min(k, n if n % 2 == 1 else max(1, n - 1)) for k, n in zip(kernel, density.shape))
        if kernel != (1, 1, 1):
            density = F.avg_pool3d(density[None, None], kernel_size=kernel, stride=1,
                                   padding=tuple(k // 2 for k in kernel))[0, 0]
    if region.projection_beta is not None:
        beta = float(region.projection_beta)
        mid = density.new_tensor(0.5)
        num = torch.tanh(beta * mid) + torch.tanh(beta * (density - mid))
        den = torch.tanh(beta * mid) + torch.tanh(beta * (1.0 - mid))
        density = num / den
    return density.clamp(0.0, 1.0)


def sample_region_density(region, density, x, y, z):
    """Trilinear ``grid_sample`` of the density at node coordinates inside the
    region's box window (compiler/materials.py:1363-1390); returns the window
    slices and the sampled float64 tensor."""
    (bx0, bx1), (by0, by1), (bz0, bz1) = region.geometry.bounds
    i0, i1 = np.searchsorted(x, bx0, "left"), np.searchsorted(x, bx1, "right")
    j0, j1 = np.searchsorted(y, by0, "left"), np.searchsorted(y, by1, "right")
    k0, k1 = np.searchsorted(z, bz0, "left"), np.searchsorted(z, bz1, "right")
    if i1 <= i0 or j1 <= j0 or k1 <= k0:
        return None, None
    cx, cy, cz = region.geometry.position
    sx, sy, sz = region.geometry.size
    dev = density.device
    xs = torch.as_tensor(2.0 * (x[i0:i1] - cx) / sx, dtype=torch.float64, device=dev)
    ys = torch.as_tensor(2.0 * (y[j0:j1] - cy) / sy, dtype=torch.float64, device=dev)
    zs = torch.as_tensor(2.0 * (z[k0:k1] - cz) / sz, dtype=torch.float64, device=dev)
    gx, gy, gz = torch.meshgrid(xs, ys, zs, indexing="ij")
    grid = torch.stack((gx, gy, gz), dim=-1)[None]
    texture = density.permute(2, 1, 0)[None, None]
    sampled = F.grid_sample(texture, grid, mode="bilinear", padding_mode="border", align_corners=False)[0, 0]
    return (slice(i0, i1), slice(j0, j1), slice(k0, k1)), sampled


def node_permittivity(prepared, model: NodeMaterialModel, region_densities=None, device=None) -> torch.Tensor:
    """Node relative permittivity as a float64 torch tensor with the design regions
    blended in (compiler/materials.py:1393-1470): each region's density is
    normalised, filtered and projected (``prepare_region_density``), sampled
    trilinearly at the nodes of its box, mapped to eps_lo + rho (eps_hi - eps_lo)
    and blended with the region's hard occupancy, which also clears the
    conductivity and the dispersive poles there.  ``region_densities`` (default:
    each region's ``density``) are the design variables; the autograd graph runs
    back to them.  Everything runs on ``device`` (default: the first density's
    device), so a HIP-resident density is compiled on the GPU."""
    regions = list(prepared.scene.material_regions)
    densities = list(region_densities) if region_densities is not None else [r.density for r in regions]
    if device is None:
        device = densities[0].device if densities else torch.device("cpu")
    eps = torch.as_tensor(model.eps_r, dtype=torch.float64, device=device)
    if not regions:
        return eps
    eps = eps.clone()
    x, y, z = prepared.x_nodes64, prepared.y_nodes64, prepared.z_nodes64
    for region, dens in zip(regions, densities):
        dens = prepare_region_density(prepared, region, dens.to(device))
        window, sampled = sample_region_density(region, dens, x, y, z)
        if window is None:
            continue
        occ_np = region.geometry.contains(x[window[0]][:, None, None], y[window[1]][None, :, None],
                                          z[window[2]][None, None, :])
        occ = torch.as_tensor(occ_np, dtype=torch.float64, device=device)
        lo_e, hi_e = region.eps_bounds
        value = lo_e + sampled * (hi_e - lo_e)
        eps[window] = eps[window] * (1.0 - occ) + occ * value
        hard = np.broadcast_to(np.asarray(occ_np, dtype=bool), sampled.shape)
        model.region_mask[window] |= hard
        # the region replaces the medium: no conductivity, no poles (materials.py:1471-1484)
        model.sigma_e[window][hard] = 0.0
        for _, _, weight in model.poles:
            weight[window][hard] = 0.0
    return eps


def electric_coefficients(eps_rel_edge, sigma_edge, dt: float):
    """Semi-implicit lossy E coefficients (materials.py:942-970): with
    h = sigma dt / (2 eps), decay = (1-h)/(1+h) and curl = (dt/eps)/(1+h); sigma = 0
    gives exactly (1, dt/eps).  Works on numpy arrays and on float64 torch
    tensors (the adjoint chains coefficient gradients through it)."""
    if isinstance(eps_rel_edge, torch.Tensor):
        sigma_edge = torch.as_tensor(sigma_edge, dtype=eps_rel_edge.dtype, device=eps_rel_edge.device)
    eps = eps_rel_edge * EPSILON_0
    half = 0.5 * sigma_edge * dt / eps
    return (1.0 - half) / (1.0 + half), (dt / eps) / (1.0 + half)


def _uniform_node(arr):
    """The value of a node array that holds one value everywhere, else None."""
    a = np.asarray(arr)
    v = a.flat[0]
    return float(v) if np.all(a == v) else None


def update_coefficients(prepared, model: NodeMaterialModel, eps_node: np.ndarray, dt: float):
    """Per-component (decay, curl) in float64 (materials.py:942-1008, :1577-1594).
    A coefficient that is uniform over its component (lossless media, uniform mu)
    is returned as a zero-stride broadcast of its value: the same numbers, with no
    full-size array behind them."""
    coef = {}
    eps_edges = {}
    lossless = _uniform_node(model.sigma_e) == 0.0
    for comp in E_COMPONENTS:
        eps_rel = average_node_to_component(eps_node, comp)
        eps_edges[comp] = eps_rel * EPSILON_0
        if lossless:
            # sigma = 0: decay = (1 - 0)/(1 + 0) = 1 and curl = (dt/eps)/1 exactly
            coef[comp] = (np.broadcast_to(np.float64(1.0), eps_rel.shape), dt / eps_edges[comp])
        else:
            coef[comp] = electric_coefficients(eps_rel, average_node_to_component(model.sigma_e, comp), dt)
    mu_value = _uniform_node(model.mu_r)
    for comp in H_COMPONENTS:
        shape = component_shape(comp, model.mu_r.shape)
        if mu_value is not None:
            coef[comp] = (np.broadcast_to(np.float64(1.0), shape), np.broadcast_to(np.float64(dt / (mu_value * MU_0)), shape))
        else:
            mu = average_node_to_magnetic_component(model.mu_r, comp) * MU_0
            coef[comp] = (np.ones_like(mu), dt / mu)
    return coef, eps_edges


def dispersive_templates(model: NodeMaterialModel, eps_edges, dt: float):
    """ADE constants per E component (materials.py:481-600): decay/restoring
    scalars and the edge-averaged drive arrays, plus 1/eps_inf."""
    out = {}
    for comp in E_COMPONENTS:
        entries = []
        for kind, pole, weight_node in model.poles:
            weight = average_node_to_component(weight_node, comp)
            if kind == "debye":
                decay = (2.0 * pole.tau - dt) / (2.0 * pole.tau + dt)
                scale = 2.0 * EPSILON_0 * pole.delta_eps * dt / (2.0 * pole.tau + dt)
                entries.append({"kind": "debye", "decay": decay, "restoring": 0.0, "drive": weight * scale,
                                "drive_scale": scale})
            elif kind == "drude":
                wp = 2.0 * np.pi * pole.plasma_frequency
                g = 2.0 * np.pi * pole.gamma
                denom = 2.0 + g * dt
                scale = 2.0 * EPSILON_0 * wp * wp * dt / denom
                entries.append({"kind": "drude", "decay": (2.0 - g * dt) / denom, "restoring": 0.0,
                                "drive": weight * scale, "drive_scale": scale})
            else:
                w0 = 2.0 * np.pi * pole.resonance_frequency
                g = 2.0 * np.pi * pole.gamma
                denom = 2.0 + g * dt
                scale = 2.0 * EPSILON_0 * pole.delta_eps * w0 * w0 * dt / denom
                entries.append({"kind": "lorentz", "decay": (2.0 - g * dt) / denom,
                                "restoring": 2.0 * w0 * w0 * dt / denom,
                                "drive": weight * scale, "drive_scale": scale})
        out[comp] = {"inv_eps": 1.0 / eps_edges[comp], "poles": entries}
    return out


def material_characteristic_frequency(scene) -> float:
    """Largest pole frequency in Hz (fdtd/runtime/initialization.py:24-69)."""
    best = 0.0
    for s in scene.structures:
        m = s.material
        for p in m.debye_poles:
            best = max(best, 1.0 / (2.0 * np.pi * p.tau))
        for p in m.drude_poles:
            best = max(best, p.plasma_frequency, p.gamma)
        for p in m.lorentz_poles:
            best = max(best, p.resonance_frequency, p.gamma)
    return best


def pole_weight_codes(poles):
    """Per-sample byte of one E component with two bits per pole: 0 where pole p
    does not act, 1 where its drive is half its ``drive_scale`` (an edge with one
    end node in the medium), 2 where it is the whole scale.  Node occupancy is 0 or
    1, so every drive is one of those values and the step kernels rebuild it from
    the code (exactly: halving a float32 is exact).  None without poles, with more
    than four, or when some drive is not one of the three values."""
    if not poles or len(poles) > 4:
        return None
    codes = np.zeros(np.shape(poles[0]["drive"]), dtype=np.uint8)
    for p, pole in enumerate(poles):
        drive = np.asarray(pole["drive"], dtype=np.float32)
        full = np.float32(pole["drive_scale"])
        half = np.float32(0.5) * full
        code = np.where(drive == full, 2, np.where(drive == half, 1, 0)).astype(np.uint8)
        if np.any((code == 0) & (drive != 0)):
            return None
        codes |= code << np.uint8(2 * p)
    return codes


def pole_activity(poles):
    """Per-sample activity mask of the poles on one E component: bit p set where
    pole p's drive is non-zero.  Where a pole's drive is zero its state starts at
    zero and stays zero and it adds no current, so the step kernels skip it there
    and read the drive arrays only where some pole acts.  None without poles."""
    mask = None
    for p, pole in enumerate(poles):
        # the device holds the drives in float32: test the values it will read
        bit = (np.asarray(pole["drive"], dtype=np.float32) != 0).astype(np.uint8) << np.uint8(p)
        mask = bit if mask is None else (mask | bit)
    return mask
