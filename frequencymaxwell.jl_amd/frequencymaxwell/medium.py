"""Medium container and the bead phantom.

Mirrors src/medium/types.jl:160-273 (Medium, accessors) and
src/medium/phantoms.jl:58-106, 212-247 (phantom_bead, _fill_sphere!).
"""

from __future__ import annotations

import warnings

import numpy as np


class Medium:
    """Complex relative permittivity eps(r) on the grid plus the real background eps_bg."""

    def __init__(self, permittivity, permittivity_bg: float):
        if not permittivity_bg > 0:  # types.jl:169-170
            raise ValueError(f"permittivity_bg must be positive, got: {permittivity_bg}")
        eps = np.asfortranarray(permittivity, dtype=np.complex128)
        if eps.ndim != 3:
            raise ValueError("permittivity must be a 3-D array")
        self.permittivity = eps
        self.permittivity_bg = float(permittivity_bg)

    def __repr__(self):
        a = np.abs(self.permittivity)
        g = self.permittivity.shape
        return (f"Medium{{Float64}}\n  Grid size: {g[0]} × {g[1]} × {g[2]}\n"
                f"  Background: εᵣ = {self.permittivity_bg:.4f}, n = {np.sqrt(self.permittivity_bg):.4f}\n"
                f"  Permittivity range: {a.min():.4f} to {a.max():.4f}")


def permittivity(m: Medium) -> np.ndarray:
    """types.jl:205."""
    return m.permittivity


def permittivity_bg(m: Medium) -> float:
    """types.jl:226."""
    return m.permittivity_bg


def grid_size(m: Medium) -> tuple:
    """types.jl:245."""
    return tuple(int(s) for s in m.permittivity.shape)


def _fill_sphere(ph: np.ndarray, center, radius: float, value: complex) -> None:
    """_fill_sphere! (phantoms.jl:212-247): 1-based voxel centres, half-even bbox rounding."""
    g = ph.shape
    lo = [max(1, int(np.rint(center[d] - radius))) for d in range(3)]
    hi = [min(g[d], int(np.rint(center[d] + radius))) for d in range(3)]
    if any(hi[d] < lo[d] for d in range(3)):
        return
    i = np.arange(lo[0], hi[0] + 1)[:, None, None]
    j = np.arange(lo[1], hi[1] + 1)[None, :, None]
    k = np.arange(lo[2], hi[2] + 1)[None, None, :]
    inside = (i - center[0]) ** 2 + (j - center[1]) ** 2 + (k - center[2]) ** 2 <= radius ** 2
    sub = ph[lo[0] - 1:hi[0], lo[1] - 1:hi[1], lo[2] - 1:hi[2]]
    sub[inside] = value


def phantom_bead(grid_size, permittivity_profile, pixel_radius: float, *,
                 permittivity_bg: float = 1.0, bead_distance: float | None = None) -> Medium:
    """phantom_bead (phantoms.jl:58-106): beads along z around the grid centre (g+1)/2."""
    grid_size = tuple(int(g) for g in grid_size)
    if not all(g > 0 for g in grid_size):
        raise ValueError("grid_size must have positive dimensions")
    if not pixel_radius > 0:
        raise ValueError("pixel_radius must be positive")
    if len(permittivity_profile) == 0:
        raise ValueError("permittivity_profile cannot be empty")
    if not permittivity_bg > 0:
        raise ValueError("permittivity_bg must be positive")
    if bead_distance is None:
        bead_distance = 2 * pixel_radius
    nb = len(permittivity_profile)
    ph = np.full(grid_size, complex(permittivity_bg), dtype=np.complex128, order="F")
    center = [(g + 1) / 2 for g in grid_size]
    for b in range(1, nb + 1):
        zoff = (b - (nb + 1) / 2) * bead_distance
        c = (center[0], center[1], center[2] + zoff)
        if c[2] < 1 or c[2] > grid_size[2]:
            warnings.warn(f"Bead {b} is outside grid bounds, skipping")
            continue
        _fill_sphere(ph, c, pixel_radius, complex(permittivity_profile[(b - 1) % nb]))
    return Medium(ph, permittivity_bg)
