"""Seeded workload scenes: the five benchmark/parity configurations.

Each factory returns ``(scene, run_kwargs)`` where run_kwargs carries the
frequencies, step count and time step for ``Simulation.fdtd`` / ``compile_plan``.
Sizes are physical cells per axis; the Yee grid adds 1 + 2*PML nodes per axis
(reference convention, scene.py:1260-1281).  ``scale`` shrinks a configuration
for parity tests (cells and feature sizes scale together, steps are given
explicitly by the caller).
"""

from __future__ import annotations

import numpy as np
import torch

from .constants import C_0
from .geometry import Box, Sphere, Structure
from .media import DrudePole, LorentzPole, Material
from .monitors import FluxMonitor, PlaneMonitor, PointMonitor
from .scene import BoundarySpec, Domain, GridSpec, MaterialRegion, Scene
from .sources import TFSF, GaussianPulse, PlaneWave, PointDipole

LAMBDA0 = 1.55e-6
F0 = C_0 / LAMBDA0
DX = LAMBDA0 / 20.0
COURANT = 0.99


def _domain(cells, dx=DX):
    ext = [0.5 * n * dx for n in cells]
    return Domain(tuple((-e, e) for e in ext))


def vacuum_dipole(cells=64, pml=10, steps=500, device="cuda"):
    """C1: vacuum cube, CPML 10, centred Ez ideal dipole, 3-frequency DFT + point monitor."""
    scene = Scene(domain=_domain((cells,) * 3), grid=GridSpec.uniform(DX), boundary=BoundarySpec.pml(pml),
                  device=device)
    scene.add_source(PointDipole((0.0, 0.0, 0.0), "Ez", profile="ideal", source_time=GaussianPulse(F0, F0 / 4)))
    scene.add_monitor(PointMonitor("probe", (4 * DX, 0.0, 0.0), fields=("Ez",)))
    freqs = (0.9 * F0, F0, 1.1 * F0)
    return scene, {"frequencies": freqs, "time_steps": steps, "courant": COURANT, "full_field_dft": True}


def dielectric_boxes(cells=256, pml=12, n_boxes=40, steps=2000, seed=0, box_range=(8, 64), device="cuda"):
    """C2: dielectric boxes (eps ~ U[1,12]), +z soft plane wave, two z-plane flux
    monitors at 8 frequencies over f0 +- 20%."""
    rng = np.random.default_rng(seed)
    scene = Scene(domain=_domain((cells,) * 3), grid=GridSpec.uniform(DX), boundary=BoundarySpec.pml(pml),
                  device=device)
    half = 0.5 * cells * DX
    for _ in range(n_boxes):
        size = rng.integers(box_range[0], box_range[1] + 1, size=3) * DX
        centre = rng.uniform(-half, half, size=3)
        eps = rng.uniform(1.0, 12.0)
        scene.add_structure(Structure(Box(tuple(centre), tuple(size)), Material(eps_r=float(eps))))
    scene.add_source(PlaneWave((0, 0, 1), (1, 0, 0), source_time=GaussianPulse(F0, F0 / 4)))
    freqs = tuple(np.linspace(0.8 * F0, 1.2 * F0, 8))
    scene.add_monitor(FluxMonitor("flux_lo", axis="z", position=-0.3 * half, frequencies=freqs))
    scene.add_monitor(FluxMonitor("flux_hi", axis="z", position=0.35 * half, frequencies=freqs))
    return scene, {"frequencies": freqs, "time_steps": steps, "courant": COURANT}


def dispersive_spheres(cells=(384, 384, 256), pml=12, fill=0.30, steps=4000, seed=0, radius=(6, 24),
                       device="cuda", injection="tfsf", pulse_delay=None):
    """C3: Drude + 2-pole Lorentz spheres filling ~30% of cells, the rest eps ~ U[1,4]
    boxes; TFSF slab plane wave; flux monitors at 16 frequencies."""
    rng = np.random.default_rng(seed)
    scene = Scene(domain=_domain(cells), grid=GridSpec.uniform(DX), boundary=BoundarySpec.pml(pml), device=device)
    half = np.array(cells) * 0.5 * DX
    # a TFSF slab needs vacuum on its two faces (4 cells inside the physical z
    # span): every structure stays within |z| <= half_z - 6 cells
    zlim = half[2] - 6 * DX if injection == "tfsf" else np.inf
    # background: eps in [1,4] slabs of random boxes
    for _ in range(12):
        size = rng.integers(8, 64, size=3) * DX
        centre = rng.uniform(-half, half)
        z0, z1 = max(centre[2] - 0.5 * size[2], -zlim), min(centre[2] + 0.5 * size[2], zlim)
        if z1 <= z0:
            continue
        centre[2], size[2] = 0.5 * (z0 + z1), z1 - z0
        scene.add_structure(Structure(Box(tuple(centre), tuple(size)), Material(eps_r=float(rng.uniform(1, 4)))))
    two_pi = 2.0 * np.pi
    metal = Material(eps_r=2.0, drude_poles=(DrudePole(1.2e16 / two_pi, 1e14 / two_pi),),
                     lorentz_poles=(LorentzPole(1.5, 4e15 / two_pi, 2e14 / two_pi),
                                    LorentzPole(0.8, 6e15 / two_pi, 2e14 / two_pi)))
    volume = float(np.prod(cells))
    filled = 0.0
    while filled < fill * volume:
        r = rng.uniform(radius[0], radius[1])
        centre = rng.uniform(-half, half)
        centre[2] = float(np.clip(centre[2], -zlim + r * DX, zlim - r * DX)) if np.isfinite(zlim) else centre[2]
        scene.add_structure(Structure(Sphere(tuple(centre), r * DX), metal))
        filled += 4.0 / 3.0 * np.pi * r ** 3 * 0.85  # overlaps and clipping lose some volume
    # slab faces 4 cells inside the physical z span
    inj = TFSF.slab(axis="z", bounds=(-half[2] + 4 * DX, half[2] - 4 * DX)) if injection == "tfsf" else "soft"
    scene.add_source(PlaneWave((0, 0, 1), (1, 0, 0), source_time=GaussianPulse(F0, F0 / 4, delay=pulse_delay),
                               injection=inj))
    freqs = tuple(np.linspace(0.7 * F0, 1.3 * F0, 16))
    scene.add_monitor(FluxMonitor("flux_t", axis="z", position=0.4 * half[2], frequencies=freqs))
    # The explicit ADE is unstable at Courant 0.99 for omega_p*dt ~ 1.8, so this
    # configuration keeps the reference auto_dt, which resolves the plasma
    # frequency (fdtd/runtime/initialization.py:106-130).
    return scene, {"frequencies": freqs, "time_steps": steps}


def design_region(cells=(200, 200, 120), design=(64, 64, 32), pml=12, steps=3000, seed=0, device="cuda",
                  theta=None):
    """C4: sigmoid-parameterised design region, Ez dipole, single-frequency flux
    objective through a 40x40 plane."""
    rng = np.random.default_rng(seed)
    scene = Scene(domain=_domain(cells), grid=GridSpec.uniform(DX), boundary=BoundarySpec.pml(pml), device=device)
    if theta is None:
        theta = torch.as_tensor(rng.standard_normal(design), dtype=torch.float64)
    density = torch.sigmoid(theta)
    dsize = tuple(n * DX for n in design)
    scene.add_material_region(MaterialRegion("design", Box((0, 0, 0), dsize), density, eps_bounds=(1.0, 12.0)))
    zsrc = -0.5 * dsize[2] - 6 * DX
    scene.add_source(PointDipole((0, 0, zsrc), "Ez", profile="ideal", source_time=GaussianPulse(F0, F0 / 4)))
    plane = 0.5 * dsize[2] + 6 * DX
    scene.add_monitor(PlaneMonitor("objective", axis="z", position=plane, fields=("Ex", "Ey", "Hx", "Hy"),
                                   frequencies=(F0,)))
    return scene, {"frequencies": (F0,), "time_steps": steps, "courant": COURANT, "theta": theta}


def batched_scenes(count=32, smallest=(64, 64, 64), largest=(192, 192, 128), design=(32, 32, 16), pml=10,
                   steps=1500, seed=0, device="cuda", n_boxes=6):
    """C5: ``count`` independent scenes of seeded uneven sizes between ``smallest``
    and ``largest`` cells, each with random dielectric boxes (eps ~ U[1,12]), a
    sigmoid-parameterised design region, one Ez dipole below it and a
    single-frequency flux plane above it.  Returns a list of (scene, theta) and
    the shared run kwargs (every scene runs the same steps at Courant 0.99)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array(smallest), np.array(largest)
    out = []
    dsize = np.array(design) * DX
    for _ in range(count):
        cells = tuple(int(v) for v in np.round(lo + rng.uniform(0, 1) * (hi - lo)))
        scene = Scene(domain=_domain(cells), grid=GridSpec.uniform(DX), boundary=BoundarySpec.pml(pml), device=device)
        half = np.array(cells) * 0.5 * DX
        for _ in range(n_boxes):
            size = rng.integers(4, 16, size=3) * DX
            centre = rng.uniform(-0.8 * half, 0.8 * half)
            # keep boxes clear of the source and monitor planes (|z| > design/2 + 4 cells)
            if abs(centre[2]) < 0.5 * dsize[2] + 8 * DX:
                continue
            scene.add_structure(Structure(Box(tuple(centre), tuple(size)), Material(eps_r=float(rng.uniform(1, 12)))))
        theta = torch.as_tensor(rng.standard_normal(design), dtype=torch.float64)
        scene.add_material_region(MaterialRegion("design", Box((0, 0, 0), tuple(dsize)), torch.sigmoid(theta),
                                                 eps_bounds=(1.0, 12.0)))
        zsrc = -0.5 * dsize[2] - 4 * DX
        scene.add_source(PointDipole((0, 0, zsrc), "Ez", profile="ideal", source_time=GaussianPulse(F0, F0 / 4)))
        scene.add_monitor(FluxMonitor("objective", axis="z", position=0.5 * dsize[2] + 4 * DX, frequencies=(F0,)))
        out.append((scene, theta))
    return out, {"frequencies": (F0,), "time_steps": steps, "courant": COURANT}


__all__ = ["vacuum_dipole", "dielectric_boxes", "dispersive_spheres", "design_region", "batched_scenes", "F0", "DX",
           "LAMBDA0"]
