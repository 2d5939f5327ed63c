"""Shared problem builders for the parity tests (oracle and product on identical inputs)."""

from __future__ import annotations

import math

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


# scene builders (boundary conditions, coherent plane waves, random media) for
# the device-setup tests: matching host-mirror and oracle objects
def scene_bcs(fm, oracle, kind):
    if kind == "periodic":
        return fm.PeriodicBoundaryCondition(), oracle.PeriodicBC()
    if kind == "absorbing":
        return (fm.AbsorbingBoundaryCondition(thickness=400e-9, attenuation_thickness=300e-9),
                oracle.AbsorbingBC(thickness=400e-9, attenuation_thickness=300e-9))
    # z only absorbing, exponential profile, sharpness 0.7; x, y periodic
    f = fm.AbsorbingBoundaryCondition(thickness=500e-9, attenuation_thickness=400e-9,
                                      sharpness=0.7, profile=fm.ExponentialProfile)
    o = oracle.AbsorbingBC(thickness=500e-9, attenuation_thickness=400e-9, sharpness=0.7,
                           profile="exponential")
    return ((fm.PeriodicBoundaryCondition(), fm.PeriodicBoundaryCondition(), f),
            (oracle.PeriodicBC(), oracle.PeriodicBC(), o))


def scene_sources(fm, oracle, many):
    th, ph = 0.4, 1.1
    k = [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]
    p1 = [-math.sin(ph), math.cos(ph), 0.0]
    specs = [dict(polarization=[1, 0, 0], k_vector=[0, 0, 1])]
    if many:
        specs = [dict(polarization=p1, k_vector=k, amplitude=0.7, phase=0.3),
                 dict(polarization=[0, 1j, 0], k_vector=[0, 0, 1], amplitude=1.2, phase=-1.0)]
    return ([fm.PlaneWaveSource(wavelength=500e-9, **s) for s in specs],
            [oracle.PlaneWave(wavelength=500e-9, **s) for s in specs])


def scene_medium(fm, grid, seed):
    rng = np.random.default_rng(seed)
    n = 1.33 + 0.2 * rng.random(grid)
    eps = np.asfortranarray((n ** 2 + 0.01j * rng.random(grid)).astype(np.complex128))
    return fm.Medium(eps, N_BG ** 2)
