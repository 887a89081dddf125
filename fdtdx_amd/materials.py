"""Materials and dispersion poles (reference: src/fdtdx/materials.py:63-222,
src/fdtdx/dispersion.py:371-456, 789-857).

Permittivity/permeability are stored as 9-tuples (row-major 3x3) like the reference; a scalar
or a 3-tuple (diagonal) is expanded on construction.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def _to9(v) -> tuple:
    if np.isscalar(v):
        v = float(v)
        return (v, 0.0, 0.0, 0.0, v, 0.0, 0.0, 0.0, v)
    v = tuple(float(x) for x in v)
    if len(v) == 3:
        return (v[0], 0.0, 0.0, 0.0, v[1], 0.0, 0.0, 0.0, v[2])
    if len(v) == 9:
        return v
    raise ValueError(f"material property must be a scalar, 3-tuple or 9-tuple, got {v}")


def _unit(u):
    """Pole._validate_orientation (dispersion.py:245-265): a finite non-zero 3-vector, normalised."""
    if u is None:
        return None
    u = tuple(float(x) for x in u)
    if len(u) != 3:
        raise ValueError(f"Pole orientation must be a 3-tuple (x, y, z), got {u!r}.")
    if not all(np.isfinite(u)):
        raise ValueError(f"Pole orientation components must be finite, got {u!r}.")
    norm = float(np.sqrt(sum(x * x for x in u)))
    if norm == 0.0:
        raise ValueError("Pole orientation must be a non-zero vector.")
    return tuple(x / norm for x in u)


def _axis3(v) -> tuple:
    if np.isscalar(v):
        return (float(v),) * 3
    v = tuple(float(x) for x in v)
    if len(v) != 3:
        raise ValueError(f"per-axis pole parameter must be a scalar or 3-tuple, got {v}")
    return v


@dataclass(frozen=True)
class LorentzPole:
    """chi(w) = d_eps * w0^2 / (w0^2 - w^2 - i*gamma*w)."""

    resonance_frequency: float | tuple
    damping: float | tuple
    delta_epsilon: float | tuple
    # a 1-D oscillator along this unit vector (coupling K u u^T, dispersion.py:85-110), or None
    orientation: tuple | None = None

    def __post_init__(self):
        object.__setattr__(self, "orientation", _unit(self.orientation))
        if self.orientation is not None and not all(
                np.isscalar(x) for x in (self.resonance_frequency, self.damping, self.delta_epsilon)):
            raise ValueError("an oriented pole takes scalar parameters")

    @property
    def is_oriented(self):
        return self.orientation is not None

    @property
    def omega_0_axes(self):
        return _axis3(self.resonance_frequency)

    @property
    def gamma_axes(self):
        return _axis3(self.damping)

    @property
    def coupling_sq_axes(self):
        w, de = self.omega_0_axes, _axis3(self.delta_epsilon)
        return tuple(de[a] * w[a] ** 2 for a in range(3))

    @property
    def is_isotropic(self):
        return self.orientation is None and all(
            np.isscalar(x) for x in (self.resonance_frequency, self.damping, self.delta_epsilon))

    def susceptibility(self, omega):
        w0, g, de = self.omega_0_axes[0], self.gamma_axes[0], _axis3(self.delta_epsilon)[0]
        return de * w0**2 / (w0**2 - omega**2 - 1j * g * omega)


@dataclass(frozen=True)
class DrudePole:
    """chi(w) = -wp^2 / (w^2 + i*gamma*w)."""

    plasma_frequency: float | tuple
    damping: float | tuple
    orientation: tuple | None = None

    def __post_init__(self):
        object.__setattr__(self, "orientation", _unit(self.orientation))
        if self.orientation is not None and not (np.isscalar(self.plasma_frequency) and np.isscalar(self.damping)):
            raise ValueError("an oriented pole takes scalar parameters")

    @property
    def is_oriented(self):
        return self.orientation is not None

    @property
    def omega_0_axes(self):
        return (0.0, 0.0, 0.0)

    @property
    def gamma_axes(self):
        return _axis3(self.damping)

    @property
    def coupling_sq_axes(self):
        wp = _axis3(self.plasma_frequency)
        return tuple(x**2 for x in wp)

    @property
    def is_isotropic(self):
        return self.orientation is None and np.isscalar(self.plasma_frequency) and np.isscalar(self.damping)

    def susceptibility(self, omega):
        wp, g = _axis3(self.plasma_frequency)[0], self.gamma_axes[0]
        return -(wp**2) / (omega**2 + 1j * g * omega)


@dataclass(frozen=True)
class DispersionModel:
    poles: tuple = ()

    def __post_init__(self):
        object.__setattr__(self, "poles", tuple(self.poles))


@dataclass(frozen=True)
class Material:
    permittivity: tuple = 1.0
    permeability: tuple = 1.0
    electric_conductivity: tuple = 0.0
    magnetic_conductivity: tuple = 0.0
    dispersion: DispersionModel | None = None

    def __post_init__(self):
        object.__setattr__(self, "permittivity", _to9(self.permittivity))
        object.__setattr__(self, "permeability", _to9(self.permeability))
        object.__setattr__(self, "electric_conductivity", _to9(self.electric_conductivity))
        object.__setattr__(self, "magnetic_conductivity", _to9(self.magnetic_conductivity))

    @property
    def is_isotropic_permittivity(self) -> bool:
        p = self.permittivity
        return p[0] == p[4] == p[8] and not any(p[i] for i in (1, 2, 3, 5, 6, 7))

    @property
    def is_diagonal_permittivity(self) -> bool:
        p = self.permittivity
        return not any(p[i] for i in (1, 2, 3, 5, 6, 7))

    @property
    def is_magnetic(self) -> bool:
        return self.permeability != _to9(1.0)

    @property
    def is_conductive(self) -> bool:
        return any(self.electric_conductivity) or any(self.magnetic_conductivity)

    @property
    def num_poles(self) -> int:
        return 0 if self.dispersion is None else len(self.dispersion.poles)


def compute_pole_coefficients_per_axis(poles, dt: float):
    """dispersion.py:789-857: c1 = (2 - w0^2 dt^2)/D, c2 = -(1 - gamma dt/2)/D, c3 = K dt^2/D,
    D = 1 + gamma dt / 2, per pole and axis (float64)."""
    n = len(poles)
    c1 = np.zeros((n, 3)); c2 = np.zeros((n, 3)); c3 = np.zeros((n, 3))
    for i, p in enumerate(poles):
        w0, g, K = p.omega_0_axes, p.gamma_axes, p.coupling_sq_axes
        for ax in range(3):
            if K[ax] != 0.0 and w0[ax] * dt >= 2.0:
                raise ValueError(
                    f"Pole {i} ({type(p).__name__}) has omega_0 * dt = {w0[ax] * dt:.4g} >= 2; the ADE recurrence "
                    "roots leave the unit circle (requires omega_0 * dt < 2)."
                )
            D = 1.0 + 0.5 * g[ax] * dt
            c1[i, ax] = (2.0 - (w0[ax] ** 2) * (dt**2)) / D
            c2[i, ax] = -(1.0 - 0.5 * g[ax] * dt) / D
            c3[i, ax] = (K[ax] * dt**2) / D
    return c1, c2, c3


def compute_pole_coefficients_tensor(poles, dt: float):
    """dispersion.py:888-952: c1, c2 per pole and axis as compute_pole_coefficients_per_axis, and the
    coupling c3 as a row-major 3x3 tensor per pole: (K dt^2 / D) u u^T for a pole oriented along u,
    the per-axis c3 on the diagonal otherwise (float64)."""
    c1, c2, c3d = compute_pole_coefficients_per_axis(poles, dt)
    c3 = np.zeros((len(poles), 9))
    for i, p in enumerate(poles):
        if getattr(p, "is_oriented", False):
            K = p.coupling_sq_axes[0]
            if K < 0.0:
                raise ValueError(f"Pole {i} ({type(p).__name__}) has negative coupling K = {K:.4g}; oriented poles "
                                 "require K >= 0 so the coupling tensor K u u^T stays positive semi-definite.")
            D = 1.0 + 0.5 * p.gamma_axes[0] * dt
            u = np.asarray(p.orientation, np.float64)
            c3[i] = ((K * dt**2 / D) * np.outer(u, u)).reshape(-1)
        else:
            for ax in range(3):
                c3[i, 4 * ax] = c3d[i, ax]
    return c1, c2, c3
