"""ElectromagneticField container.

Mirrors src/fields/electromagnetic_field.jl:46-90 (container), :252-266 (+),
:288-303 (crop_to_ROI), :324-333 (to_host).
"""

from __future__ import annotations

import math

import numpy as np


class ElectromagneticField:
    """E and H, each (nx, ny, nz, 3) complex, with resolution and wavelength."""

    def __init__(self, E, H, resolution, wavelength: float):
        E = np.asarray(E)
        H = np.asarray(H)
        if E.shape != H.shape:
            raise ValueError("Efield and Hfield arrays must have same size")
        if E.ndim == 4 and E.shape[3] != 3:
            raise ValueError("field arrays must have 3 components in last dimension")
        if not all(r > 0 for r in resolution):
            raise ValueError("resolution must be positive")
        if not wavelength > 0:
            raise ValueError("wavelength must be positive")
        self.E = E
        self.H = H
        self.resolution = tuple(float(r) for r in resolution)
        self.wavelength = float(wavelength)

    def __add__(self, other: "ElectromagneticField") -> "ElectromagneticField":
        if self.E.shape != other.E.shape:
            raise ValueError(f"Fields must have same grid size: {self.E.shape} vs {other.E.shape}")
        if self.resolution != other.resolution:
            raise ValueError("Fields must have same resolution")
        if not math.isclose(self.wavelength, other.wavelength, rel_tol=1e-10, abs_tol=0.0):
            raise ValueError("Fields must have same wavelength")
        return ElectromagneticField(self.E + other.E, self.H + other.H, self.resolution,
                                    self.wavelength)


def crop_array(a: np.ndarray, roi) -> np.ndarray:
    """Crop ``a`` to 1-based inclusive ROI bounds (convergent_born.jl:951-960)."""
    return np.asfortranarray(a[roi[0] - 1:roi[1], roi[2] - 1:roi[3], roi[4] - 1:roi[5], ...])


def crop_to_ROI(field: ElectromagneticField, solver) -> ElectromagneticField:
    """electromagnetic_field.jl:288-303."""
    roi = solver.compute_ROI()
    return ElectromagneticField(crop_array(field.E, roi), crop_array(field.H, roi),
                                solver.resolution, field.wavelength)
