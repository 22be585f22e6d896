"""MI355X-native FDTD runtime with the witwin.maxwell ``Scene -> Simulation.fdtd(...) -> Result`` API.

The time-step loop (Yee E/H, CPML psi, Debye/Drude/Lorentz ADE, soft/TFSF sources,
running DFT) and its adjoint run as hand-written HIP kernels for gfx950 behind a
C ABI (``include/maxwell_fdtd.h``); this package is the host side.
"""

from .constants import C_0, EPSILON_0, ETA_0, MU_0
from .geometry import Box, Sphere, Structure
from .media import DebyePole, DrudePole, LorentzPole, Material
from .monitors import FluxMonitor, PlaneMonitor, PointMonitor
from .scene import BoundarySpec, Domain, GridSpec, MaterialRegion, Scene, SceneModule, prepare_scene
from .sources import CW, TFSF, GaussianPulse, PlaneWave, PointDipole, RickerWavelet

__all__ = [
    "C_0", "EPSILON_0", "ETA_0", "MU_0",
    "Box", "Sphere", "Structure",
    "Material", "DebyePole", "DrudePole", "LorentzPole",
    "PointMonitor", "PlaneMonitor", "FluxMonitor",
    "Scene", "SceneModule", "MaterialRegion", "Domain", "GridSpec", "BoundarySpec", "prepare_scene",
    "GaussianPulse", "CW", "RickerWavelet", "PointDipole", "PlaneWave", "TFSF",
]


def __getattr__(name):
    # Simulation/Result pull in the native runtime; import them lazily so that the
    # scene API stays usable for compile-only work.
    if name in ("Simulation", "PreparedSimulation", "FDTDConfig", "TimeConfig", "SpectralSampler", "Result", "run",
                "calculate_required_steps"):
        from . import simulation
        return getattr(simulation, name)
    raise AttributeError(name)
