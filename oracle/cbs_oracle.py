"""FP64 CPU oracle for the convergent-Born-series (CBS) solve of FrequencyMaxwell.jl.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py`` -- never by the product package, which must run on the HIP library
or fail.

It is a project-written restatement, in NumPy/SciPy, of the reference's
algorithm.  Every function cites the reference file:line it follows (paths are
relative to the FrequencyMaxwell.jl source tree).  Arrays use the reference's
Julia layout: complex128, Fortran order, fields shaped ``(nx, ny, nz, 3)`` with
x fastest and the component slowest.

Pinning (what checks each function independently of its own formula; the
tests live in ``tests/test_oracle.py``):

* boundary conditions, windows, ROI  -- the reference's own unit-test values
  (test/solvers/boundary_conditions.jl:80-136) and closed-form endpoints.
* FFT convention                     -- direct O(N^2) DFT sums.
* Green operator                     -- Bloch-wave eigenfunctions of each
  shifted half (closed-form eigenvalue), and the independent inverse check
  ``M(k) G(k) = I`` with ``M(k) = |k|^2 I - k k^T - k_c^2 I``.
* curl                               -- analytic curl of single Fourier modes,
  including the odd-N index quirk.
* incident field                     -- closed-form plane wave, transversality,
  impedance ``|E|/|H| = 377/sqrt(eps_bg)`` and the reference's own tests
  (test/sources/plane_wave.jl).
* field analysis                     -- a plane wave's Poynting vector
  ``|A|^2/Z`` along k, the energy of constant fields in closed form, the
  right-hand rule of the cross product on unit vectors, planes by index.
* Born / BiCGStab / GMRES solves     -- dense brute-force solve of
  ``(I - G_avg V) x = G_avg S`` on tiny grids, the spectral PDE residual
  ``M(k) x = F[V x + S]`` (periodic), homogeneous and uniform-medium closed
  forms, mirror symmetry.
* Krylov iterates                   -- fixed-budget BiCGStab (shadow r0) and GMRES
  (full and restarted) iterates equal scipy.sparse.linalg's ``bicgstab`` /
  ``gmres`` on the explicit left-preconditioned operator; every GMRES iterate
  is the least-squares minimiser over the Krylov space (brute force); every
  BiCGStab iterate, for either shadow vector, satisfies the Petrov-Galerkin
  characterisation r_k = Omega_k(A) phi_k(A) r0 (brute force).
* What stays unpinned: the stopping point and count convention of Krylov.jl
  itself (third party, not vendored), restated here from its published
  algorithm.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np
import scipy.fft as sfft

# FFT threads: the host's cores, at most 16 (a one-GPU share of a GPU node)
WORKERS = int(os.environ.get("CBS_ORACLE_WORKERS", min(16, os.cpu_count() or 1)))
Z0_REF = 377.0  # convergent_born.jl:797, plane_wave.jl:232 (quirk: 377, not 376.73)


# ---------------------------------------------------------------------------
# Boundary conditions -- src/solvers/boundary_conditions.jl
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class PeriodicBC:
    """PeriodicBoundaryCondition (boundary_conditions.jl:157-181)."""


@dataclass(frozen=True)
class AbsorbingBC:
    """AbsorbingBoundaryCondition (boundary_conditions.jl:335-395)."""

    thickness: float
    attenuation_thickness: float
    sharpness: float = 0.9
    profile: str = "tanh"  # TanhProfile | "exponential" | "tukey" (:228-232)

    def __post_init__(self):  # validation, boundary_conditions.jl:348-357
        if not self.thickness > 0:
            raise ValueError("thickness must be positive")
        if not self.attenuation_thickness > 0:
            raise ValueError("attenuation_thickness must be positive")
        if not self.attenuation_thickness <= self.thickness:
            raise ValueError("attenuation_thickness must be <= thickness")
        if not (0 <= self.sharpness <= 1):
            raise ValueError("sharpness must be in [0, 1]")


def julia_round(x: float) -> int:
    """``round(Int, x)``: round half to even (quirk Q10)."""
    return int(np.rint(x))


def padding_pixels(bc, resolution: float) -> int:
    """boundary_conditions.jl:450-459."""
    if isinstance(bc, PeriodicBC):
        return 0
    return julia_round(bc.thickness / (resolution * 2))


def subpixel_shift(bc) -> float:
    """boundary_conditions.jl:518-527."""
    return 0.0 if isinstance(bc, PeriodicBC) else 0.25


def requires_padding(bc) -> bool:
    """boundary_conditions.jl:567-573."""
    return isinstance(bc, AbsorbingBC)


def requires_averaging(bc) -> bool:
    """boundary_conditions.jl:609-615."""
    return isinstance(bc, AbsorbingBC)


def _julia_range(a: float, b: float, n: int) -> np.ndarray:
    """``range(a, b, length=n)``."""
    return np.linspace(a, b, n)


def attenuation_window(profile: str, L: int) -> np.ndarray:
    """_create_attenuation_window, boundary_conditions.jl:632-658.

    Tanh: ``(tanh(x)/tanh(2.5) - tanh(-2.5))/2`` -- the operator precedence of
    the reference makes the window end at -0.00669 / 0.99331 (quirk Q2).
    """
    if L <= 0:
        return np.zeros(0)
    if profile == "tanh":
        t = np.tanh(_julia_range(-2.5, 2.5, L))
        return (t / np.tanh(2.5) - np.tanh(-2.5)) / 2
    if profile == "exponential":
        x = _julia_range(0.0, 1.0, L)
        w = 0.3
        return np.exp(-(((1.0 - x) / w) ** 2))
    if profile == "tukey":
        x = _julia_range(0.0, 1.0, L)
        return 0.5 * (1 + np.cos(np.pi * (1 - x)))
    raise ValueError(f"Unknown attenuation profile: {profile}")


# ---------------------------------------------------------------------------
# Solver configuration and potential -- src/solvers/convergent_born.jl
# ---------------------------------------------------------------------------

@dataclass
class Solver:
    """The configuration fields of ConvergentBornSolver (convergent_born.jl:87-137, 191-233)."""

    resolution: tuple
    grid_size: tuple
    boundary_conditions: object  # one BC or a 3-tuple
    iterations_max: int = -1
    tolerance: float = 1e-6

    def __post_init__(self):
        bcs = self.boundary_conditions
        if not isinstance(bcs, (tuple, list)):
            bcs = (bcs, bcs, bcs)  # :207-217
        self.bcs = tuple(bcs)
        self.resolution = tuple(float(r) for r in self.resolution)
        self.grid_size = tuple(int(g) for g in self.grid_size)
        if not all(r > 0 for r in self.resolution):  # :117
            raise ValueError("all resolution components must be positive")
        if not all(g > 0 for g in self.grid_size):  # :119
            raise ValueError("all grid_size components must be positive")
        if not (0 <= self.tolerance <= 1):  # :121
            raise ValueError("tolerance must be in [0, 1]")

    @property
    def padding(self):
        return tuple(padding_pixels(self.bcs[d], self.resolution[d]) for d in range(3))

    @property
    def padded_size(self):
        return tuple(self.grid_size[d] + 2 * self.padding[d] for d in range(3))


def compute_roi(s: Solver):
    """compute_ROI, convergent_born.jl:250-260 (1-based inclusive bounds)."""
    p = s.padding
    g = s.grid_size
    return (p[0] + 1, p[0] + g[0], p[1] + 1, p[1] + g[1], p[2] + 1, p[2] + g[2])


def pad_array(arr: np.ndarray, pad) -> np.ndarray:
    """_pad_array, convergent_born.jl:807-858.  ``:replicate`` zero-fills (quirk Q1)."""
    if all(p == 0 for p in pad):
        return arr
    shape = tuple(arr.shape[d] + 2 * pad[d] for d in range(3)) + arr.shape[3:]
    out = np.zeros(shape, dtype=arr.dtype, order="F")
    out[pad[0]:pad[0] + arr.shape[0], pad[1]:pad[1] + arr.shape[1],
        pad[2]:pad[2] + arr.shape[2], ...] = arr
    return out


def crop_to_roi(arr: np.ndarray, s: Solver) -> np.ndarray:
    """crop_to_ROI, convergent_born.jl:951-960 / fields/electromagnetic_field.jl:288-303."""
    r = compute_roi(s)
    return np.asfortranarray(arr[r[0] - 1:r[1], r[2] - 1:r[3], r[4] - 1:r[5], ...])


def apply_attenuation_to_potential(s: Solver, potential: np.ndarray) -> None:
    """_apply_attenuation_to_potential!, convergent_born.jl:893-944 (in place)."""
    roi = compute_roi(s)
    for dim in range(3):
        bc = s.bcs[dim]
        if not requires_padding(bc):
            continue
        if isinstance(bc, AbsorbingBC):
            max_L = padding_pixels(bc, s.resolution[dim])
            att_px = julia_round(bc.attenuation_thickness / (s.resolution[dim] * 2))
            L = min(max_L, att_px)
            if max_L == 0:
                continue
            window = attenuation_window(bc.profile, L)
            window = window * bc.sharpness + (1 - bc.sharpness)
            roi_size = roi[2 * dim + 1] - roi[2 * dim] + 1
            remaining = max_L - L
            filt = np.concatenate([window, np.ones(roi_size + 2 * remaining), window[::-1]])
            shape = [1, 1, 1]
            shape[dim] = -1
            potential *= filt.reshape(shape, order="F")


@dataclass
class GreenParams:
    """DyadicGreen (dyadic_green.jl:45-99)."""

    k_square: complex
    freq_res: tuple
    shifts: tuple
    use_averaging: bool


@dataclass
class Init:
    potential: np.ndarray
    eps_imag: float
    bornmax: int
    k_bg: float
    green: GreenParams


def initialize_solver(s: Solver, eps_r: np.ndarray, eps_bg: float, wavelength: float) -> Init:
    """_initialize_solver!, convergent_born.jl:619-675."""
    k_bg = 2 * math.pi * math.sqrt(eps_bg) / wavelength  # :630
    v0 = (k_bg ** 2 * (eps_r / eps_bg - 1)).astype(np.complex128)  # :631-632
    pot = pad_array(np.asfortranarray(v0), s.padding)  # :638
    pot = np.array(pot, dtype=np.complex128, order="F", copy=True)
    eps_imag = max(2.0 ** -20, float(np.max(np.abs(pot))) * 1.01)  # :641 (quirk Q6)
    pot -= 1j * eps_imag  # :644
    apply_attenuation_to_potential(s, pot)  # :647
    if s.iterations_max < 0:  # :653-660
        steps = abs(2 * k_bg / eps_imag)
        extent = math.sqrt(sum((pot.shape[d] * s.resolution[d]) ** 2 for d in range(3)))
        bornmax = math.ceil(extent / steps / 2 + 1) * 2 * abs(s.iterations_max)
    else:
        bornmax = s.iterations_max
    k_square = k_bg ** 2 + 1j * eps_imag  # :663
    shape = pot.shape
    green = GreenParams(
        k_square=k_square,
        freq_res=tuple(2 * math.pi / (shape[d] * s.resolution[d]) for d in range(3)),  # dyadic_green.jl:89
        shifts=tuple(subpixel_shift(bc) for bc in s.bcs),  # dyadic_green.jl:72-74
        use_averaging=any(requires_averaging(bc) for bc in s.bcs),  # dyadic_green.jl:77
    )
    return Init(pot, eps_imag, bornmax, k_bg, green)


# ---------------------------------------------------------------------------
# Incident plane wave -- src/sources/plane_wave.jl, src/sources/field_generation.jl
# ---------------------------------------------------------------------------

@dataclass
class PlaneWave:
    """PlaneWaveSource (plane_wave.jl:63-116)."""

    wavelength: float
    polarization: Sequence[complex]
    k_vector: Sequence[float]
    amplitude: float = 1.0
    phase: float = 0.0

    def __post_init__(self):  # plane_wave.jl:79-90
        self.polarization = np.asarray(self.polarization, dtype=np.complex128)
        self.k_vector = np.asarray(self.k_vector, dtype=np.float64)
        if not self.wavelength > 0:
            raise ValueError("wavelength must be positive")
        if not self.amplitude >= 0:
            raise ValueError("amplitude must be non-negative")
        if not abs(np.linalg.norm(self.k_vector) - 1) < 1e-10:
            raise ValueError("k_vector must be normalized (|k| = 1)")
        # Julia dot(a, b) conjugates a
        if not abs(np.real(np.vdot(self.polarization, self.k_vector))) < 1e-10:
            raise ValueError("polarization must be transverse to k_vector")


def incident_field(s: Solver, src: PlaneWave, eps_bg: float = 1.0):
    """generate_incident_field(::PlaneWaveSource), plane_wave.jl:198-268.

    The reference loops over voxels; this evaluates the same per-voxel
    expression on index grids (position origin: centre of the unpadded grid,
    quirk Q9).
    """
    g = s.grid_size
    res = s.resolution
    pad = s.padding
    n = s.padded_size
    k0 = 2 * math.pi / src.wavelength  # :218
    k_bg = k0 * math.sqrt(eps_bg)  # :219
    k_hat = src.k_vector / np.linalg.norm(src.k_vector)  # :222
    k_vec = k_bg * k_hat  # :223
    e0 = src.polarization / np.linalg.norm(src.polarization)  # :224
    e_perp = e0 - np.vdot(e0, k_hat) * k_hat  # :227 (conjugating dot, Q9)
    e_perp = e_perp / np.linalg.norm(e_perp)  # :228
    h_perp = np.cross(k_hat, e_perp) / (Z0_REF / math.sqrt(eps_bg))  # :231-232
    center = [(g[d] - 1) * res[d] / 2 for d in range(3)]  # :235
    r = [(np.arange(1, n[d] + 1) - 1 - pad[d]) * res[d] - center[d] for d in range(3)]  # :240-244
    ph = (k_vec[0] * r[0][:, None, None] + k_vec[1] * r[1][None, :, None]
          + k_vec[2] * r[2][None, None, :]) + src.phase  # :247
    amp = src.amplitude * np.exp(1j * ph)  # :248
    E = np.empty(n + (3,), dtype=np.complex128, order="F")
    H = np.empty(n + (3,), dtype=np.complex128, order="F")
    for c in range(3):
        E[..., c] = amp * e_perp[c]
        H[..., c] = amp * h_perp[c]
    return E, H


def incident_field_sum(s: Solver, sources, eps_bg: float = 1.0):
    """generate_incident_field(::Vector), field_generation.jl:42-78: coherent sum (Q14)."""
    if isinstance(sources, PlaneWave):
        return incident_field(s, sources, eps_bg)
    if len(sources) == 0:
        raise ValueError("sources vector cannot be empty")
    lam = sources[0].wavelength
    for sc in sources[1:]:
        if not np.isclose(sc.wavelength, lam, rtol=1e-10, atol=0):
            raise ValueError("All sources must have the same wavelength")
    E, H = incident_field(s, sources[0], eps_bg)
    for sc in sources[1:]:
        e2, h2 = incident_field(s, sc, eps_bg)
        E = E + e2
        H = H + h2
    return E, H


def source_wavelength(sources) -> float:
    """plane_wave.jl:125, :132-141."""
    if isinstance(sources, PlaneWave):
        return sources.wavelength
    lam = sources[0].wavelength
    for sc in sources[1:]:
        if sc.wavelength != lam:
            raise ValueError("All sources must have the same wavelength")
    return lam


# ---------------------------------------------------------------------------
# Dyadic Green operator -- src/core/dyadic_green.jl
# ---------------------------------------------------------------------------

def fft3(a: np.ndarray) -> np.ndarray:
    """``fft!(src, 1:3)``: unnormalised forward DFT over the three spatial axes."""
    return sfft.fftn(a, axes=(0, 1, 2), workers=WORKERS)


def ifft3(a: np.ndarray) -> np.ndarray:
    """``ifft!(src, 1:3)``: inverse DFT with 1/N."""
    return sfft.ifftn(a, axes=(0, 1, 2), workers=WORKERS)


def green_freq_index(N: int) -> np.ndarray:
    """``rem(i + floor(N/2) - 1, N) - floor(N/2)`` for i = 1..N (dyadic_green.jl:203-208)."""
    i = np.arange(1, N + 1)
    h = N // 2
    return np.remainder(i + h - 1, N) - h


_CACHE: dict = {}


def _green_grids(shape, gp: GreenParams, shifts):
    """k components and denominator of dyadic_green.jl:203-214 (cached per problem)."""
    key = ("g", tuple(shape), gp.freq_res, tuple(shifts), gp.k_square)
    if key not in _CACHE:
        nx, ny, nz = shape
        kx = gp.freq_res[0] * (green_freq_index(nx) + shifts[0])
        ky = gp.freq_res[1] * (green_freq_index(ny) + shifts[1])
        kz = gp.freq_res[2] * (green_freq_index(nz) + shifts[2])
        KX = kx[:, None, None]
        KY = ky[None, :, None]
        KZ = kz[None, None, :]
        den = np.abs(KX ** 2 + KY ** 2 + KZ ** 2) - gp.k_square
        if len(_CACHE) > 64:
            _CACHE.clear()
        _CACHE[key] = (KX, KY, KZ, den)
    return _CACHE[key]


def multiply_green(src: np.ndarray, gp: GreenParams, shifts) -> None:
    """multiply_Green_kernel!, dyadic_green.jl:189-220 (in place, quirk Q4)."""
    KX, KY, KZ, den = _green_grids(src.shape[:3], gp, shifts)
    s1, s2, s3 = src[..., 0].copy(), src[..., 1].copy(), src[..., 2].copy()
    kdotE = (KX * s1 + KY * s2 + KZ * s3) / gp.k_square
    src[..., 0] = (s1 - KX * kdotE) / den
    src[..., 1] = (s2 - KY * kdotE) / den
    src[..., 2] = (s3 - KZ * kdotE) / den


def phase_ramp(shape, shifts, reverse: bool) -> np.ndarray:
    """Per-voxel factor of multiply_phase_ramp_kernel!, dyadic_green.jl:243-266 (quirk Q8)."""
    key = ("r", tuple(shape), tuple(shifts), bool(reverse))
    if key not in _CACHE:
        nx, ny, nz = shape
        i = np.arange(1, nx + 1)[:, None, None]
        j = np.arange(1, ny + 1)[None, :, None]
        k = np.arange(1, nz + 1)[None, None, :]
        phase = (i - 1) * shifts[0] / nx + (j - 1) * shifts[1] / ny + (k - 1) * shifts[2] / nz
        if len(_CACHE) > 64:
            _CACHE.clear()
        _CACHE[key] = np.exp((2j if reverse else -2j) * np.pi * phase)
    return _CACHE[key]


def apply_single_convolution(src: np.ndarray, gp: GreenParams, shifts) -> np.ndarray:
    """_apply_single_convolution!, dyadic_green.jl:161-179.

    With all shifts zero the ramp is exp(0) = 1 exactly, so the multiply is
    skipped (bit-identical).
    """
    shifted = any(s != 0 for s in shifts)
    if shifted:
        ramp = phase_ramp(src.shape[:3], shifts, False)[..., None]
        a = fft3(src * ramp)  # :164-167
    else:
        a = fft3(src)
    multiply_green(a, gp, shifts)  # :170
    a = ifft3(a)  # :173
    if shifted:
        a *= phase_ramp(src.shape[:3], shifts, True)[..., None]  # :176
    return np.asfortranarray(a)


def green_conv(src: np.ndarray, gp: GreenParams) -> np.ndarray:
    """conv!(::DyadicGreen, src), dyadic_green.jl:127-151 (returns the result, quirk Q11)."""
    if not gp.use_averaging:
        return apply_single_convolution(src, gp, gp.shifts)
    plus = apply_single_convolution(src, gp, gp.shifts)  # :140
    minus = apply_single_convolution(src, gp, tuple(-s for s in gp.shifts))  # :143-144
    out = plus + minus  # :147
    out /= 2  # :148
    return np.asfortranarray(out)


# ---------------------------------------------------------------------------
# Curl -- src/core/curl.jl ; magnetic field -- convergent_born.jl:784-800
# ---------------------------------------------------------------------------

def curl_freq_index(N: int) -> np.ndarray:
    """``i <= div(N,2) ? i-1 : i-N-1`` for i = 1..N (curl.jl:84-86, quirk Q3)."""
    i = np.arange(1, N + 1)
    return np.where(i <= N // 2, i - 1, i - N - 1)


def curl_conv(field_: np.ndarray, resolution) -> np.ndarray:
    """conv(::Curl, field), curl.jl:31-36, :55-66, :73-116."""
    nx, ny, nz = field_.shape[:3]
    fr = [2 * math.pi / (field_.shape[d] * resolution[d]) for d in range(3)]
    a = fft3(field_)
    ikx = (1j * fr[0] * curl_freq_index(nx))[:, None, None]
    iky = (1j * fr[1] * curl_freq_index(ny))[None, :, None]
    ikz = (1j * fr[2] * curl_freq_index(nz))[None, None, :]
    s1, s2, s3 = a[..., 0].copy(), a[..., 1].copy(), a[..., 2].copy()
    a[..., 0] = iky * s3 - ikz * s2
    a[..., 1] = ikz * s1 - ikx * s3
    a[..., 2] = ikx * s2 - iky * s1
    return np.asfortranarray(ifft3(a))


def h_scale(wavelength: float) -> complex:
    """``Complex(0,-1) * wavelength / (2pi * 377)``, convergent_born.jl:797."""
    return complex(0, -1) * wavelength / (2 * math.pi * Z0_REF)


def magnetic_field(E: np.ndarray, resolution, wavelength: float) -> np.ndarray:
    """_compute_magnetic_field, convergent_born.jl:784-800."""
    return h_scale(wavelength) * curl_conv(E, resolution)


# ---------------------------------------------------------------------------
# CBS operator and preconditioner -- convergent_born.jl:379-518
# ---------------------------------------------------------------------------

def cbs_operator(x: np.ndarray, V: np.ndarray, gp: GreenParams, alpha=1.0, beta=0.0, y=None):
    """``y = alpha (x - G_avg(V x)) + beta y`` -- CBSLinearOperator, convergent_born.jl:482-512."""
    psi = V[..., None] * x  # :494
    psi = green_conv(psi, gp)  # :497
    psi = x - psi  # :500
    if y is None:
        return np.asfortranarray(alpha * psi)
    return np.asfortranarray(beta * y + alpha * psi)  # :504-505


def cbs_precond(x: np.ndarray, V: np.ndarray, eps_imag: float) -> np.ndarray:
    """``y = (i/eps) V x`` -- CBSPreconditioner ldiv!, convergent_born.jl:398-410."""
    return np.asfortranarray((V[..., None] * x) * (1j / eps_imag))


# ---------------------------------------------------------------------------
# Linear solvers on gamma (I - G_avg V) x = gamma b  (gamma = i V / eps)
# ---------------------------------------------------------------------------

@dataclass
class SolveStats:
    iters: int = 0          # iterations as the method counts them (Krylov.jl ``niter``;
                            # Born: sweeps)
    applications: int = 0   # operator applications (voxel-iterations / N_vox)
    converged: bool = False
    breakdown: bool = False
    resid: float = float("nan")  # ||gamma (b - A x)|| / ||gamma b|| of the returned x
    history: list = field(default_factory=list)


def _stop_threshold(rule: str, tol: float, abstol: float, bnorm: float, r0norm: float) -> float:
    """'rhs': tol*||gamma b|| (the metric's definition).  'krylov': atol + rtol*||r0||
    (Krylov.jl's test, LinearSolve passes abstol = tol*1e-2, convergent_born.jl:743-747)."""
    if rule == "rhs":
        return tol * bnorm
    if rule == "krylov":
        return abstol + tol * r0norm
    raise ValueError(rule)


def _nrm(a):
    return float(np.sqrt(np.vdot(a, a).real))


def solve_born(b, x0, V, eps_imag, gp, tol, maxiters, rule="rhs", abstol=0.0):
    """CBS (preconditioned Richardson) iteration ``x <- x + gamma (b - A x)``.

    Each pass costs one operator application.  The residual of the current
    iterate is measured in the same pass; when it meets the threshold that
    iterate is returned unchanged.  ``maxiters`` bounds the passes.
    """
    precond = lambda v: cbs_precond(v, V, eps_imag)
    bnorm = _nrm(precond(b))
    st = SolveStats()
    x = x0.copy(order="F")
    if bnorm == 0.0:
        st.converged, st.resid = True, 0.0
        return np.zeros_like(x), st
    thr = None
    while True:
        r = precond(b - cbs_operator(x, V, gp))
        st.iters += 1
        st.applications += 1
        rn = _nrm(r)
        st.history.append(rn / bnorm)
        if thr is None:
            thr = _stop_threshold(rule, tol, abstol, bnorm, rn)
        if rn <= thr:
            st.converged, st.resid = True, rn / bnorm
            return x, st
        if st.iters >= maxiters:
            st.resid = rn / bnorm
            x = x + r
            return x, st
        x = x + r


def _left_system(b, V, eps_imag, gp):
    """The left-preconditioned operator v -> gamma A v and gamma b
    (LinearSolve passes Pl = CBSPreconditioner to Krylov.jl, which applies it
    with ldiv!, convergent_born.jl:398-431, 750)."""
    A = lambda v: cbs_precond(cbs_operator(v, V, gp), V, eps_imag)
    return A, cbs_precond(b, V, eps_imag)


def solve_bicgstab(b, x0, V, eps_imag, gp, tol, maxiters, rule="rhs", abstol=0.0, shadow="b"):
    """Left-preconditioned BiCGStab as Krylov.jl ``bicgstab`` runs it for
    LinearSolve's ``KrylovJL_BICGSTAB`` (the route of
    examples/two_beam_interference.jl:58; Krylov.jl is third party, not vendored).

    Krylov.jl's published algorithm (van der Vorst 1992 in its notation): warm
    start ``r0 = M (b - A x0)`` (one operator application outside the iteration
    count), shadow vector ``c = b`` (Krylov.jl's default, the unpreconditioned
    right-hand side; ``shadow='r0'`` gives the textbook choice), ``p1 = r0``;
    iteration k: ``v = M A p``, ``alpha = rho / <c, v>``, ``s = r - alpha v``,
    ``x += alpha p``, ``t = M A s``, ``omega = <t, s> / <t, t>``, ``x += omega s``,
    ``r = s - omega t``, ``rho' = <c, r>``, ``beta = (rho'/rho)(alpha/omega)``,
    ``p = r + beta (p - omega v)``; then stop when ``||r|| <= eps`` (solved),
    ``k >= maxiters`` (tired) or ``alpha`` is 0 or NaN (breakdown).  There is no
    test on the half step ``s``.  Krylov.jl also stops when ``||r|| + 1 <= 1``, an
    absolute test that depends on the units of the source term; it is left out.  ``maxiters`` and
    ``iters`` count iterations (two operator applications each); the iterate
    sequence of the third-party code itself is not imported here.
    """
    A, bh = _left_system(b, V, eps_imag, gp)
    bnorm = _nrm(bh)
    st = SolveStats()
    x = x0.copy(order="F")
    if bnorm == 0.0:
        st.converged, st.resid = True, 0.0
        return np.zeros_like(x), st
    r = bh - A(x)
    st.applications += 1
    rn = _nrm(r)
    st.history.append(rn / bnorm)
    thr = _stop_threshold(rule, tol, abstol, bnorm, rn)
    if rn == 0.0 or rn <= thr:
        st.converged, st.resid = True, rn / bnorm
        return x, st
    c = b.copy(order="F") if shadow == "b" else r.copy(order="F")
    rho = np.vdot(c, r)
    if rho == 0:
        st.breakdown, st.resid = True, rn / bnorm
        return x, st
    p = r.copy(order="F")
    while True:
        st.iters += 1
        v = A(p)
        st.applications += 1
        alpha = rho / np.vdot(c, v)
        s = r - alpha * v
        x = x + alpha * p
        t = A(s)
        st.applications += 1
        tt = np.vdot(t, t).real
        omega = np.vdot(t, s) / tt if tt != 0.0 else 0.0
        x = x + omega * s
        r = s - omega * t
        rho_next = np.vdot(c, r)
        if omega != 0:
            beta = (rho_next / rho) * (alpha / omega)
            p = r + beta * (p - omega * v)
        rho = rho_next
        rn = _nrm(r)
        st.history.append(rn / bnorm)
        solved = rn <= thr
        breakdown = alpha == 0 or np.isnan(alpha) or (omega == 0 and not solved)
        if solved or breakdown or st.iters >= maxiters:
            st.converged, st.breakdown = bool(solved), bool(breakdown and not solved)
            st.resid = rn / bnorm
            return x, st


def _givens(a, b):
    """Rotation (c real, s complex) with [c s; -conj(s) c] [a; b] = [rho; 0]."""
    na = abs(a)
    den = math.sqrt(abs(a) ** 2 + abs(b) ** 2)
    if den == 0.0:
        return 1.0, 0.0j
    if na == 0.0:
        return 0.0, 1.0 + 0.0j
    return na / den, (a / na) * np.conj(b) / den


def solve_gmres(b, x0, V, eps_imag, gp, tol, maxiters, rule="rhs", abstol=0.0, restart=0,
                btol=None):
    """Left-preconditioned GMRES as Krylov.jl ``gmres`` runs it for the reference's
    default ``KrylovJL_GMRES()`` (convergent_born.jl:197, 743-750): warm start
    ``r0 = M (b - A x0)`` (one application outside the iteration count),
    modified Gram-Schmidt Arnoldi, Givens rotations, stop after an Arnoldi step
    when the residual estimate ``|z_{k+1}| <= eps`` (solved; Krylov.jl's absolute
    ``|z| + 1 <= 1`` is left out as for BiCGStab),
    ``h_{k+1,k} <= btol`` (breakdown; Krylov.jl's default ``eps(T)^(3/4)``), or
    ``maxiters`` Arnoldi steps in total.  ``restart = m > 0`` runs GMRES(m): after
    m steps x is updated and the residual recomputed with one more
    application.  ``iters`` counts Arnoldi steps (Krylov.jl ``niter``).
    """
    A, bh = _left_system(b, V, eps_imag, gp)
    if btol is None:
        btol = np.finfo(np.float64).eps ** 0.75
    bnorm = _nrm(bh)
    st = SolveStats()
    x = x0.copy(order="F")
    if bnorm == 0.0:
        st.converged, st.resid = True, 0.0
        return np.zeros_like(x), st
    r = bh - A(x)
    st.applications += 1
    beta = _nrm(r)
    st.history.append(beta / bnorm)
    thr = _stop_threshold(rule, tol, abstol, bnorm, beta)
    if beta == 0.0 or beta <= thr:
        st.converged, st.resid = True, beta / bnorm
        return x, st
    first = True
    while True:
        if not first:  # restart: the true residual of the updated x
            r = bh - A(x)
            st.applications += 1
            beta = _nrm(r)
        first = False
        left = maxiters - st.iters
        m = min(restart, left) if restart > 0 else left
        Vb = [r / beta]
        H = np.zeros((m + 1, m), dtype=np.complex128)
        cs = np.zeros(m, dtype=np.float64)
        sn = np.zeros(m, dtype=np.complex128)
        g = np.zeros(m + 1, dtype=np.complex128)
        g[0] = beta
        k = 0
        solved = breakdown = False
        while True:
            w = A(Vb[k])
            st.applications += 1
            st.iters += 1
            for i in range(k + 1):  # modified Gram-Schmidt
                H[i, k] = np.vdot(Vb[i], w)
                w = w - H[i, k] * Vb[i]
            hnext = _nrm(w)
            H[k + 1, k] = hnext
            for i in range(k):  # the previous rotations on column k
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -np.conj(sn[i]) * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            cs[k], sn[k] = _givens(H[k, k], H[k + 1, k])
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k + 1] = -np.conj(sn[k]) * g[k]
            g[k] = cs[k] * g[k]
            resid = abs(g[k + 1])
            st.history.append(resid / bnorm)
            k += 1
            solved = resid <= thr
            breakdown = hnext <= btol
            if solved or breakdown or k >= m:
                break
            Vb.append(w / hnext)
        y = np.zeros(k, dtype=np.complex128)
        for i in range(k - 1, -1, -1):  # back substitution R y = g
            acc = g[i] - H[i, i + 1:k] @ y[i + 1:k]
            y[i] = 0.0 if abs(H[i, i]) <= btol else acc / H[i, i]
        for i in range(k):
            x = x + y[i] * Vb[i]
        st.resid = abs(g[k]) / bnorm
        if solved or breakdown or st.iters >= maxiters:
            st.converged, st.breakdown = bool(solved), bool(breakdown and not solved)
            return x, st


def true_residual(x, b, V, eps_imag, gp) -> float:
    """FP64 ``||gamma (b - A x)|| / ||gamma b||`` of a returned field."""
    bh = cbs_precond(b, V, eps_imag)
    bn = _nrm(bh)
    if bn == 0:
        return 0.0
    return _nrm(cbs_precond(b - cbs_operator(x, V, gp), V, eps_imag)) / bn


# ---------------------------------------------------------------------------
# solve -- convergent_born.jl:565-600, 703-774
# ---------------------------------------------------------------------------

@dataclass
class Result:
    E: np.ndarray          # total E, cropped to the ROI
    H: np.ndarray          # total H, cropped to the ROI
    E_scat: np.ndarray     # padded scattered E (solution x)
    H_scat: np.ndarray     # padded scattered H
    stats: SolveStats
    init: Init
    rhs: np.ndarray        # G_avg S
    source_term: np.ndarray


def source_term(pot: np.ndarray, E_inc: np.ndarray, eps_imag: float) -> np.ndarray:
    """``V .* E_inc .+ i eps .* E_inc``, convergent_born.jl:585-586."""
    return np.asfortranarray(pot[..., None] * E_inc + complex(0, eps_imag) * E_inc)


def solve(s: Solver, sources, eps_r: np.ndarray, eps_bg: float, method: str = "gmres",
          rule: str = "krylov", maxiters: int | None = None, restart: int = 0) -> Result:
    """solve(solver, sources, medium), convergent_born.jl:565-600.

    ``rule='krylov'`` reproduces the reference's LinearSolve options
    (reltol=tol, abstol=tol*1e-2, maxiters=max(2 Bornmax, 100), x0 = rhs,
    convergent_born.jl:727-750, quirk Q7).
    """
    if tuple(eps_r.shape) != s.grid_size:  # :572-573
        raise ValueError("medium grid_size must match solver grid_size")
    lam = source_wavelength(sources)
    init = initialize_solver(s, eps_r, eps_bg, lam)
    if init.bornmax < 1:  # quirk Q13: the reference leaves Efield undefined
        raise ValueError("iterations_max must be >= 1 (or negative for auto)")
    E_inc, H_inc = incident_field_sum(s, sources, eps_bg)
    S = source_term(init.potential, E_inc, init.eps_imag)
    rhs = green_conv(S, init.green)  # :715-723
    if maxiters is None:
        maxiters = max(init.bornmax * 2, 100)  # :746
    tol = s.tolerance
    abstol = tol * 1e-2  # :745
    solver_fn = {"born": solve_born, "bicgstab": solve_bicgstab, "gmres": solve_gmres}[method]
    kw = dict(rule=rule, abstol=abstol)
    if method == "gmres":
        kw["restart"] = restart
    x, st = solver_fn(rhs, rhs, init.potential, init.eps_imag, init.green, tol, maxiters, **kw)
    Hs = magnetic_field(x, s.resolution, lam)  # :770
    E_tot = x + E_inc  # :596 (quirk Q12)
    H_tot = Hs + H_inc
    return Result(crop_to_roi(E_tot, s), crop_to_roi(H_tot, s), x, Hs, st, init, rhs, S)


# ---------------------------------------------------------------------------
# Field analysis -- src/fields/electromagnetic_field.jl:112-235
# ---------------------------------------------------------------------------

EPS0_FIELD = 8.854187817e-12  # electromagnetic_field.jl:113
MU0_FIELD = 4e-7 * math.pi     # :114


def field_energy(E: np.ndarray, H: np.ndarray, resolution) -> float:
    """(eps0 sum|E|^2 + mu0 sum|H|^2) / 2 times the voxel volume (:112-127)."""
    e = float((E.real ** 2 + E.imag ** 2).sum())
    h = float((H.real ** 2 + H.imag ** 2).sum())
    return (EPS0_FIELD * e + MU0_FIELD * h) / 2 * float(np.prod(resolution))


def poynting_vector(E: np.ndarray, H: np.ndarray) -> np.ndarray:
    """Complex S = E x conj(H), component by component as at :135-154."""
    Hc = np.conj(H)
    S = np.empty_like(E)
    for c in range(3):
        a, b = (c + 1) % 3, (c + 2) % 3
        S[..., c] = E[..., a] * Hc[..., b] - E[..., b] * Hc[..., a]
    return np.asfortranarray(S)


def field_intensity(E: np.ndarray, component: str = "total") -> np.ndarray:
    """|E|^2 summed over the components (``:total``) or of one (:162-180)."""
    a = E.real ** 2 + E.imag ** 2
    if component == "total":
        return np.asfortranarray(a[..., 0] + a[..., 1] + a[..., 2])
    return np.asfortranarray(a[..., "xyz".index(component)])


def extract_plane(F: np.ndarray, plane_axis: int, plane_index: int) -> np.ndarray:
    """The 1-based plane ``plane_index`` normal to ``plane_axis``, kept 4-D (:192-235)."""
    idx = [slice(None)] * 4
    idx[plane_axis - 1] = slice(plane_index - 1, plane_index)
    return np.asfortranarray(F[tuple(idx)])


# ---------------------------------------------------------------------------
# Dense brute force (tiny grids): assemble I - G_avg V column by column
# ---------------------------------------------------------------------------

def dense_operator(V: np.ndarray, gp: GreenParams) -> np.ndarray:
    shape = V.shape + (3,)
    n = int(np.prod(shape))
    A = np.empty((n, n), dtype=np.complex128)
    e = np.zeros(n, dtype=np.complex128)
    for col in range(n):
        e[:] = 0
        e[col] = 1
        A[:, col] = cbs_operator(e.reshape(shape, order="F"), V, gp).reshape(-1, order="F")
    return A


def flat(a: np.ndarray) -> np.ndarray:
    return a.reshape(-1, order="F")


def unflat(v: np.ndarray, shape) -> np.ndarray:
    return v.reshape(shape, order="F")
