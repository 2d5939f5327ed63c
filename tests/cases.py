"""Shared problem builders for the parity tests (oracle and product on identical inputs)."""

from __future__ import annotations

import numpy as np

RES = (50e-9, 50e-9, 50e-9)
N_BG = 1.33


def random_eps(shape, seed, n_lo=1.33, n_hi=1.5, frac=0.5):
    """Random permittivity: background water with a random blob of higher index."""
    rng = np.random.default_rng(seed)
    n = np.full(shape, n_lo)
    mask = rng.random(shape) < frac
    n[mask] = n_lo + (n_hi - n_lo) * rng.random(int(mask.sum()))
    return np.asfortranarray((n ** 2).astype(np.complex128))


def random_field(shape, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(tuple(shape) + (3,)) + 1j * rng.standard_normal(tuple(shape) + (3,))
    return np.asfortranarray(f)


def oracle_setup(oracle, grid, bc, seed, wavelength=500e-9, n_hi=1.5):
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bc)
    eps = random_eps(grid, seed, n_hi=n_hi)
    init = oracle.initialize_solver(s, eps, N_BG ** 2, wavelength)
    return s, eps, init


def spec_from(fm, init, res=RES):
    g = init.green
    return fm.GreenSpec(shape=init.potential.shape, resolution=res, shifts=g.shifts,
                        use_averaging=g.use_averaging, k_square=g.k_square,
                        eps_imag=init.eps_imag)


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d
