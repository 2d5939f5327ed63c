"""Plane-wave source and the incident field on the padded grid.

Mirrors src/sources/plane_wave.jl (PlaneWaveSource :63-116, source_wavelength
:125-141, generate_incident_field :198-268) and src/sources/field_generation.jl
:42-78 (a vector of sources is a coherent sum: one right-hand side).
"""

from __future__ import annotations

import math

import numpy as np

from .boundary_conditions import padding_pixels
from .fields import ElectromagneticField

Z0 = 377.0  # plane_wave.jl:232 uses 377 Ohm (not 376.73)


class AbstractCurrentSource:
    """AbstractCurrentSource (src/core/types.jl:22)."""


class PlaneWaveSource(AbstractCurrentSource):
    """PlaneWaveSource(; wavelength, polarization, k_vector, amplitude=1, phase=0)."""

    def __init__(self, *, wavelength: float, polarization, k_vector, amplitude: float = 1.0,
                 phase: float = 0.0):
        pol = np.asarray(polarization, dtype=np.complex128).reshape(3)
        kv = np.asarray(k_vector, dtype=np.float64).reshape(3)
        if not wavelength > 0:
            raise ValueError("wavelength must be positive")
        if not amplitude >= 0:
            raise ValueError("amplitude must be non-negative")
        if not abs(np.linalg.norm(kv) - 1) < 1e-10:
            raise ValueError("k_vector must be normalized (|k| = 1)")
        if not abs(np.real(np.vdot(pol, kv))) < 1e-10:
            raise ValueError("polarization must be transverse to k_vector")
        self.wavelength = float(wavelength)
        self.polarization = pol
        self.k_vector = kv
        self.amplitude = float(amplitude)
        self.phase = float(phase)

    def __repr__(self):
        return (f"PlaneWaveSource{{Float64}}:\n  wavelength: {self.wavelength}\n"
                f"  polarization: {self.polarization}\n  k_vector: {self.k_vector}\n"
                f"  amplitude: {self.amplitude}\n  phase: {self.phase}")


def source_wavelength(sources) -> float:
    """plane_wave.jl:125 / :132-141 (all sources of a vector share one wavelength)."""
    if isinstance(sources, AbstractCurrentSource):
        return sources.wavelength
    lam = sources[0].wavelength
    for s in sources[1:]:
        if s.wavelength != lam:
            raise ValueError("All sources must have the same wavelength for coherent superposition")
    return lam


def plane_wave_vectors(src: PlaneWaveSource, permittivity_bg: float):
    """Unit k, transverse E0 and H0 of plane_wave.jl:218-232.

    The projection uses Julia's conjugating ``dot(E0, k_hat)``.
    """
    k0 = 2 * math.pi / src.wavelength
    k_bg = k0 * math.sqrt(permittivity_bg)
    k_hat = src.k_vector / np.linalg.norm(src.k_vector)
    e0 = src.polarization / np.linalg.norm(src.polarization)
    e_perp = e0 - np.vdot(e0, k_hat) * k_hat
    e_perp = e_perp / np.linalg.norm(e_perp)
    h_perp = np.cross(k_hat, e_perp) / (Z0 / math.sqrt(permittivity_bg))
    return k_bg * k_hat, e_perp, h_perp


def _incident_one(src: PlaneWaveSource, solver, permittivity_bg: float):
    g = solver.grid_size
    res = solver.resolution
    pad = tuple(padding_pixels(solver.boundary_conditions[d], res[d]) for d in range(3))
    n = tuple(g[d] + 2 * pad[d] for d in range(3))
    k_vec, e_perp, h_perp = plane_wave_vectors(src, permittivity_bg)
    center = [(g[d] - 1) * res[d] / 2 for d in range(3)]
    r = [(np.arange(n[d]) - pad[d]) * res[d] - center[d] for d in range(3)]
    phase = (k_vec[0] * r[0][:, None, None] + k_vec[1] * r[1][None, :, None]
             + k_vec[2] * r[2][None, None, :]) + src.phase
    amp = src.amplitude * np.exp(1j * phase)
    E = np.empty(n + (3,), dtype=np.complex128, order="F")
    H = np.empty(n + (3,), dtype=np.complex128, order="F")
    for c in range(3):
        E[..., c] = amp * e_perp[c]
        H[..., c] = amp * h_perp[c]
    return ElectromagneticField(E, H, res, src.wavelength)


def generate_incident_field(sources, solver, *, permittivity_bg: float = 1.0) -> ElectromagneticField:
    """generate_incident_field: E, H of the plane wave(s) on the padded grid.

    Position origin: centre of the unpadded grid (plane_wave.jl:235-244).
    """
    if isinstance(sources, AbstractCurrentSource):
        return _incident_one(sources, solver, permittivity_bg)
    if len(sources) == 0:
        raise ValueError("sources vector cannot be empty")
    lam = sources[0].wavelength
    for s in sources[1:]:
        if not math.isclose(s.wavelength, lam, rel_tol=1e-10, abs_tol=0.0):
            raise ValueError("All sources must have the same wavelength for coherent interference")
    total = _incident_one(sources[0], solver, permittivity_bg)
    for s in sources[1:]:
        total = total + _incident_one(s, solver, permittivity_bg)
    return total
