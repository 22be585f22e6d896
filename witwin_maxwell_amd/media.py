"""Linear media: isotropic permittivity/permeability, conductivity and
Debye/Drude/Lorentz poles.

Pole types follow witwin/maxwell/media.py:125-215: frequencies and damping are
stored in Hz and converted with 2*pi in the ADE coefficients
(fdtd/runtime/materials.py:544-589).  Lorentz damping uses the
``omega0^2 - omega^2 - i*gamma*omega`` convention (media.py:209-215).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def _nonneg(v, name):
    v = float(v)
    if v < 0 or not np.isfinite(v):
        raise ValueError(f"{name} must be finite and >= 0")
    return v


def _pos(v, name):
    v = float(v)
    if v <= 0 or not np.isfinite(v):
        raise ValueError(f"{name} must be finite and > 0")
    return v


@dataclass(frozen=True)
class DebyePole:
    delta_eps: float
    tau: float

    def __post_init__(self):
        object.__setattr__(self, "delta_eps", _nonneg(self.delta_eps, "delta_eps"))
        object.__setattr__(self, "tau", _pos(self.tau, "tau"))

    def susceptibility(self, angular_frequency: float) -> complex:
        return self.delta_eps / (1.0 - 1j * float(angular_frequency) * self.tau)


@dataclass(frozen=True)
class DrudePole:
    plasma_frequency: float
    gamma: float

    def __post_init__(self):
        object.__setattr__(self, "plasma_frequency", _pos(self.plasma_frequency, "plasma_frequency"))
        object.__setattr__(self, "gamma", _nonneg(self.gamma, "gamma"))

    def susceptibility(self, angular_frequency: float) -> complex:
        w = float(angular_frequency)
        wp = 2.0 * np.pi * self.plasma_frequency
        g = 2.0 * np.pi * self.gamma
        return -(wp * wp) / (w * w + 1j * g * w)


@dataclass(frozen=True)
class LorentzPole:
    delta_eps: float
    resonance_frequency: float
    gamma: float

    def __post_init__(self):
        object.__setattr__(self, "delta_eps", float(self.delta_eps))
        object.__setattr__(self, "resonance_frequency", _pos(self.resonance_frequency, "resonance_frequency"))
        object.__setattr__(self, "gamma", _nonneg(self.gamma, "gamma"))

    def susceptibility(self, angular_frequency: float) -> complex:
        w = float(angular_frequency)
        w0 = 2.0 * np.pi * self.resonance_frequency
        g = 2.0 * np.pi * self.gamma
        return self.delta_eps * w0 * w0 / (w0 * w0 - w * w - 1j * g * w)


class Material:
    """Isotropic linear medium (media.py:626+, FDTD subset).

    ``eps_r`` is the instantaneous (infinite-frequency) relative permittivity when
    poles are present."""

    def __init__(self, eps_r: float = 1.0, mu_r: float = 1.0, sigma_e: float = 0.0, name: str | None = None,
                 *, debye_poles=(), drude_poles=(), lorentz_poles=()):
        self.eps_r = _pos(eps_r, "eps_r")
        self.mu_r = _pos(mu_r, "mu_r")
        self.sigma_e = _nonneg(sigma_e, "sigma_e")
        self.name = name
        self.debye_poles = tuple(debye_poles)
        self.drude_poles = tuple(drude_poles)
        self.lorentz_poles = tuple(lorentz_poles)
        for p in self.debye_poles:
            if not isinstance(p, DebyePole):
                raise TypeError("debye_poles must hold DebyePole")
        for p in self.drude_poles:
            if not isinstance(p, DrudePole):
                raise TypeError("drude_poles must hold DrudePole")
        for p in self.lorentz_poles:
            if not isinstance(p, LorentzPole):
                raise TypeError("lorentz_poles must hold LorentzPole")

    @property
    def is_dispersive(self) -> bool:
        return bool(self.debye_poles or self.drude_poles or self.lorentz_poles)

    def relative_permittivity(self, frequency: float) -> complex:
        w = 2.0 * np.pi * float(frequency)
        eps = complex(self.eps_r)
        for p in self.debye_poles + self.drude_poles + self.lorentz_poles:
            eps += p.susceptibility(w)
        return eps

    @classmethod
    def drude(cls, eps_inf=1.0, plasma_frequency=1.0, gamma=0.0, name=None):
        return cls(eps_r=eps_inf, drude_poles=(DrudePole(plasma_frequency, gamma),), name=name)

    @classmethod
    def lorentz(cls, eps_inf=1.0, delta_eps=1.0, resonance_frequency=1.0, gamma=0.0, name=None):
        return cls(eps_r=eps_inf, lorentz_poles=(LorentzPole(delta_eps, resonance_frequency, gamma),), name=name)

    @classmethod
    def debye(cls, eps_inf=1.0, delta_eps=1.0, tau=1.0, name=None):
        return cls(eps_r=eps_inf, debye_poles=(DebyePole(delta_eps, tau),), name=name)

    def __repr__(self):
        return (f"Material(eps_r={self.eps_r}, mu_r={self.mu_r}, sigma_e={self.sigma_e}, "
                f"debye={len(self.debye_poles)}, drude={len(self.drude_poles)}, lorentz={len(self.lorentz_poles)})")


__all__ = ["Material", "DebyePole", "DrudePole", "LorentzPole"]
