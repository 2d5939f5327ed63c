"""ElectromagneticField container and field analysis.

Mirrors src/fields/electromagnetic_field.jl:46-90 (container), :93-100
(domain_size), :112-237 (field_energy, poynting_vector, field_intensity,
extract_plane), :252-266 (+), :288-303 (crop_to_ROI), :324-333 (to_host).
The GPU versions of the analysis work on fields kept on the device
(``engine.DeviceField``).
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


EPS0 = 8.854187817e-12   # vacuum permittivity used by field_energy (:113)
MU0 = 4 * math.pi * 1e-7  # vacuum permeability (:114)


def field_domain_size(field: ElectromagneticField):
    """domain_size(EMfield) (:97-100)."""
    return tuple(field.E.shape[i] * field.resolution[i] for i in range(3))


def _abs2(a):
    """Julia's abs2: re^2 + im^2."""
    return a.real * a.real + a.imag * a.imag


def field_energy(field: ElectromagneticField) -> float:
    """(eps0 sum|E|^2 + mu0 sum|H|^2) / 2 * voxel volume (:112-127)."""
    e = float(np.sum(_abs2(field.E)))
    h = float(np.sum(_abs2(field.H)))
    return (EPS0 * e + MU0 * h) / 2 * (field.resolution[0] * field.resolution[1] * field.resolution[2])


def poynting_vector(field: ElectromagneticField) -> np.ndarray:
    """Complex S = E x conj(H), (nx, ny, nz, 3) (:135-154)."""
    if field.E.ndim != 4:
        raise ValueError("poynting_vector requires 4D field arrays (3D space + 3 components)")
    E, H = field.E, np.conj(field.H)
    S = np.empty_like(field.E)
    S[..., 0] = E[..., 1] * H[..., 2] - E[..., 2] * H[..., 1]
    S[..., 1] = E[..., 2] * H[..., 0] - E[..., 0] * H[..., 2]
    S[..., 2] = E[..., 0] * H[..., 1] - E[..., 1] * H[..., 0]
    return S


_COMPONENTS = {"total": 0, "x": 1, "y": 2, "z": 3}


def field_intensity(field: ElectromagneticField, component="total") -> np.ndarray:
    """|E|^2 summed over components, or of one component (:162-180)."""
    if field.E.ndim != 4:
        raise ValueError("field_intensity requires 4D field arrays (3D space + 3 components)")
    c = str(component).lstrip(":")
    if c not in _COMPONENTS:
        raise ValueError("component must be :total, :x, :y, or :z")
    if c == "total":
        a = _abs2(field.E)
        return np.asfortranarray(a[..., 0] + a[..., 1] + a[..., 2])
    return np.asfortranarray(_abs2(field.E[..., _COMPONENTS[c] - 1]))


def extract_plane(field: ElectromagneticField, plane_axis: int, plane_index: int) -> ElectromagneticField:
    """The plane plane_index (1-based) normal to plane_axis (1, 2, 3), kept 4-D (:192-235)."""
    if field.E.ndim != 4:
        raise ValueError("extract_plane requires 4D field arrays (3D space + 3 components)")
    if not 1 <= plane_axis <= 3:
        raise ValueError("plane_axis must be 1, 2, or 3")
    if not 1 <= plane_index <= field.E.shape[plane_axis - 1]:
        raise ValueError("plane_index out of bounds")
    sl = [slice(None)] * 4
    sl[plane_axis - 1] = slice(plane_index - 1, plane_index)
    return ElectromagneticField(np.asfortranarray(field.E[tuple(sl)]),
                                np.asfortranarray(field.H[tuple(sl)]), field.resolution,
                                field.wavelength)
