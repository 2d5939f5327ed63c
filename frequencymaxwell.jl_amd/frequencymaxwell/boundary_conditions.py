"""Boundary conditions: the per-axis parameters that feed the Green operator.

Mirrors src/solvers/boundary_conditions.jl (types :95-395, queries :450-615,
attenuation windows :632-658).
"""

from __future__ import annotations

import enum

import numpy as np


class AttenuationProfile(enum.Enum):
    """AttenuationProfile (boundary_conditions.jl:228-232)."""

    TanhProfile = 0
    ExponentialProfile = 1
    TukeyProfile = 2


TanhProfile = AttenuationProfile.TanhProfile
ExponentialProfile = AttenuationProfile.ExponentialProfile
TukeyProfile = AttenuationProfile.TukeyProfile


class AbstractBoundaryCondition:
    """AbstractBoundaryCondition (boundary_conditions.jl:95)."""


class PeriodicBoundaryCondition(AbstractBoundaryCondition):
    """PeriodicBoundaryCondition (boundary_conditions.jl:157-181): no padding, no shift."""

    def __eq__(self, other):
        return isinstance(other, PeriodicBoundaryCondition)

    def __hash__(self):
        return hash("periodic")

    def __repr__(self):
        return "PeriodicBoundaryCondition{Float64}()"


class AbsorbingBoundaryCondition(AbstractBoundaryCondition):
    """AbsorbingBoundaryCondition (boundary_conditions.jl:335-395).

    Raises ``ValueError`` (Julia ``ArgumentError``) on the reference's checks.
    """

    def __init__(self, *, thickness: float, attenuation_thickness: float,
                 sharpness: float = 0.9, profile: AttenuationProfile = TanhProfile):
        if not thickness > 0:
            raise ValueError(f"thickness must be positive, got: {thickness}")
        if not attenuation_thickness > 0:
            raise ValueError(f"attenuation_thickness must be positive, got: {attenuation_thickness}")
        if not attenuation_thickness <= thickness:
            raise ValueError(f"attenuation_thickness ({attenuation_thickness}) must be ≤ thickness ({thickness})")
        if not (0 <= sharpness <= 1):
            raise ValueError(f"sharpness must be in [0, 1], got: {sharpness}")
        self.thickness = float(thickness)
        self.attenuation_thickness = float(attenuation_thickness)
        self.sharpness = float(sharpness)
        self.profile = AttenuationProfile(profile)

    def __eq__(self, other):
        return (isinstance(other, AbsorbingBoundaryCondition)
                and (self.thickness, self.attenuation_thickness, self.sharpness, self.profile)
                == (other.thickness, other.attenuation_thickness, other.sharpness, other.profile))

    def __hash__(self):
        return hash((self.thickness, self.attenuation_thickness, self.sharpness, self.profile))

    def __repr__(self):
        return (f"AbsorbingBoundaryCondition{{Float64}}:\n  thickness: {self.thickness} m\n"
                f"  attenuation_thickness: {self.attenuation_thickness} m\n"
                f"  sharpness: {self.sharpness}\n  profile: {self.profile.name}\n"
                f"  subpixel_shift: {subpixel_shift(self)}")


def _round_half_even(x: float) -> int:
    """Julia ``round(Int, x)``."""
    return int(np.rint(x))


def padding_pixels(bc: AbstractBoundaryCondition, resolution: float) -> int:
    """boundary_conditions.jl:450-459: ``round(Int, thickness / (2 resolution))``."""
    if isinstance(bc, PeriodicBoundaryCondition):
        return 0
    return _round_half_even(bc.thickness / (float(resolution) * 2))


def subpixel_shift(bc: AbstractBoundaryCondition) -> float:
    """boundary_conditions.jl:518-527: 0 (periodic) or 0.25 (absorbing)."""
    return 0.0 if isinstance(bc, PeriodicBoundaryCondition) else 0.25


def requires_padding(bc: AbstractBoundaryCondition) -> bool:
    """boundary_conditions.jl:567-573."""
    return isinstance(bc, AbsorbingBoundaryCondition)


def requires_averaging(bc: AbstractBoundaryCondition) -> bool:
    """boundary_conditions.jl:609-615."""
    return isinstance(bc, AbsorbingBoundaryCondition)


def create_attenuation_window(profile: AttenuationProfile, L: int) -> np.ndarray:
    """_create_attenuation_window (boundary_conditions.jl:632-658).

    The tanh window is ``(tanh(x)/tanh(2.5) - tanh(-2.5)) / 2`` over
    ``x = range(-2.5, 2.5, length=L)``: it runs from -0.00669 to 0.99331.
    """
    if L <= 0:
        return np.zeros(0)
    profile = AttenuationProfile(profile)
    if profile is TanhProfile:
        t = np.tanh(np.linspace(-2.5, 2.5, L))
        return (t / np.tanh(2.5) - np.tanh(-2.5)) / 2
    x = np.linspace(0.0, 1.0, L)
    if profile is ExponentialProfile:
        return np.exp(-(((1.0 - x) / 0.3) ** 2))
    return 0.5 * (1 + np.cos(np.pi * (1 - x)))


def convert_boundary(bc) -> AbstractBoundaryCondition:
    """_convert_boundary_type (boundary_conditions.jl:685-698)."""
    if isinstance(bc, AbstractBoundaryCondition):
        return bc
    raise ValueError(f"not a boundary condition: {bc!r}")
