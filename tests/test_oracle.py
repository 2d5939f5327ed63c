"""Pins of the FP64 oracle (CPU).  Each check is independent of the formula it
checks: the reference's own unit-test values, direct DFT sums, closed forms,
eigenfunctions, an independent inverse operator, dense brute force and
symmetry.  A plausible slip (dropped term, wrong sign, wrong index, swapped
operand) fails at least one of them.
"""

import math

import numpy as np
import pytest

from cases import N_BG, RES, oracle_setup, random_field, rel


# ---------------------------------------------------------------------------
# boundary conditions -- test/solvers/boundary_conditions.jl:80-136
# ---------------------------------------------------------------------------
def test_bc_reference_values(oracle):
    o = oracle
    ab = o.AbsorbingBC(thickness=2.0e-6, attenuation_thickness=1.5e-6)
    assert o.padding_pixels(o.PeriodicBC(), 50e-9) == 0
    assert o.padding_pixels(o.PeriodicBC(), 100e-9) == 0
    assert o.padding_pixels(ab, 50e-9) == 20
    assert o.padding_pixels(ab, 100e-9) == 10
    assert o.subpixel_shift(o.PeriodicBC()) == 0.0
    assert o.subpixel_shift(ab) == 0.25
    assert not o.requires_padding(o.PeriodicBC()) and o.requires_padding(ab)
    assert not o.requires_averaging(o.PeriodicBC()) and o.requires_averaging(ab)


def test_bc_validation(oracle):
    o = oracle
    for kw in (dict(thickness=-1e-6, attenuation_thickness=1e-6),
               dict(thickness=1e-6, attenuation_thickness=0.0),
               dict(thickness=1e-6, attenuation_thickness=2e-6),
               dict(thickness=2e-6, attenuation_thickness=1e-6, sharpness=1.5)):
        with pytest.raises(ValueError):
            o.AbsorbingBC(**kw)


def test_round_half_even(oracle):
    # Julia round(Int, x) rounds ties to even (0.5 -> 0, 1.5 -> 2, 2.5 -> 2)
    o = oracle
    assert [o.julia_round(v) for v in (0.5, 1.5, 2.5, 3.5, -0.5)] == [0, 2, 2, 4, 0]
    bc = o.AbsorbingBC(thickness=250e-9, attenuation_thickness=250e-9)
    assert o.padding_pixels(bc, 50e-9) == 2  # 2.5 -> 2


def test_tanh_window_closed_form(oracle):
    w = oracle.attenuation_window("tanh", 16)
    t = math.tanh(2.5)
    assert w[0] == pytest.approx((t - 1) / 2, abs=1e-15)  # -0.006693
    assert w[-1] == pytest.approx((1 + t) / 2, abs=1e-15)  # 0.993307
    assert np.all(np.diff(w) > 0)
    # odd symmetry of tanh about the centre: w(x) + w(-x) = tanh(2.5)
    assert np.allclose(w + w[::-1], t, atol=1e-14)
    assert oracle.attenuation_window("tanh", 0).size == 0
    e = oracle.attenuation_window("exponential", 5)
    assert e[-1] == pytest.approx(1.0) and e[0] == pytest.approx(math.exp(-(1 / 0.3) ** 2))
    k = oracle.attenuation_window("tukey", 5)
    assert k[0] == pytest.approx(0.0, abs=1e-16) and k[-1] == pytest.approx(1.0)


def test_attenuated_potential_profile(oracle):
    """V along an absorbing axis: -i eps times the window in the pad, unchanged in the ROI."""
    o = oracle
    bc = (o.PeriodicBC(), o.PeriodicBC(), o.AbsorbingBC(thickness=800e-9, attenuation_thickness=600e-9,
                                                         sharpness=0.8))
    s = o.Solver(resolution=RES, grid_size=(4, 4, 10), boundary_conditions=bc)
    eps = np.full((4, 4, 10), N_BG ** 2, dtype=complex)
    init = o.initialize_solver(s, eps, N_BG ** 2, 500e-9)
    assert init.potential.shape == (4, 4, 10 + 2 * 8)
    col = init.potential[0, 0, :]
    L = 6  # round(600e-9 / 100e-9)
    w = o.attenuation_window("tanh", L) * 0.8 + 0.2
    expect = np.concatenate([w, np.ones(10 + 2 * (8 - L)), w[::-1]]) * (-1j * init.eps_imag)
    assert np.allclose(col, expect, rtol=0, atol=1e-12 * init.eps_imag)
    # eps floor 2^-20 with a zero contrast (quirk Q6)
    assert init.eps_imag == 2.0 ** -20


# ---------------------------------------------------------------------------
# FFT convention -- direct DFT sums
# ---------------------------------------------------------------------------
def test_fft_matches_direct_dft(oracle):
    a = random_field((6, 5, 4), 1)
    n = a.shape[:3]
    F = [np.exp(-2j * np.pi * np.outer(np.arange(m), np.arange(m)) / m) for m in n]
    direct = np.einsum("ia,jb,kc,abcd->ijkd", F[0], F[1], F[2], a)
    assert rel(oracle.fft3(a), direct) < 1e-13
    assert rel(oracle.ifft3(direct), a) < 1e-13


# ---------------------------------------------------------------------------
# Green operator
# ---------------------------------------------------------------------------
def _ghat(k, k2c):
    """Closed form (I - k k^T / kc^2) / (|k|^2 - kc^2) (dyadic_green.jl:185)."""
    return (np.eye(3) - np.outer(k, k) / k2c) / (k @ k - k2c)


@pytest.mark.parametrize("shape", [(6, 5, 4), (8, 8, 4)])
@pytest.mark.parametrize("sign", [+1, -1])
def test_green_bloch_eigenfunctions(oracle, shape, sign):
    """u = p exp(2 pi i n.(m + s)/N) is an eigenfunction of each shifted half with
    eigenvalue Ghat(k = dk (m_mapped + s)): pins the ramp signs, the index map
    (including Nyquist) and the k-space multiply."""
    o = oracle
    dk = tuple(2 * np.pi / (shape[d] * RES[d]) for d in range(3))
    k2c = (1.6713e7) ** 2 + 1j * 7.6e13
    shifts = (0.25 * sign, -0.25 * sign, 0.25 * sign)
    gp = o.GreenParams(k_square=k2c, freq_res=dk, shifts=(0.25, -0.25, 0.25), use_averaging=False)
    rng = np.random.default_rng(3)
    for trial in range(4):
        m = [int(rng.integers(0, shape[d])) for d in range(3)]
        if trial == 0:
            m = [shape[d] // 2 for d in range(3)]  # Nyquist on every even axis
        p = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        n = np.indices(shape)
        ph = sum(n[d] * (m[d] + shifts[d]) / shape[d] for d in range(3))
        u = np.exp(2j * np.pi * ph)[..., None] * p
        u = np.asfortranarray(u)
        got = o.apply_single_convolution(u, gp, shifts)
        # signed frequency of bin m: m for m < N/2, m - N above, -N/2 at Nyquist
        mm = [m[d] - shape[d] if 2 * m[d] >= shape[d] else m[d] for d in range(3)]
        k = np.array([dk[d] * (mm[d] + shifts[d]) for d in range(3)])
        want = u @ _ghat(k, k2c).T
        assert rel(got, want) < 1e-12


def test_green_is_inverse_of_maxwell_operator(oracle):
    """M(k) Ghat(k) = I with M(k) = |k|^2 I - k k^T - kc^2 I (curl curl - kc^2)."""
    rng = np.random.default_rng(4)
    k2c = 2.79e14 + 1j * 7.7e13
    for _ in range(20):
        k = rng.standard_normal(3) * 2e7
        M = (k @ k) * np.eye(3) - np.outer(k, k) - k2c * np.eye(3)
        # the oracle's multiply on a one-voxel grid: its frequency is freq_res * shift,
        # so freq_res = k with shift 1 puts the multiply at wave vector k
        gp = oracle.GreenParams(k_square=k2c, freq_res=tuple(k), shifts=(1.0, 1.0, 1.0),
                                use_averaging=False)
        cols = []
        for c in range(3):
            e = np.zeros((1, 1, 1, 3), dtype=complex)
            e[..., c] = 1
            oracle.multiply_green(e, gp, (1.0, 1.0, 1.0))
            cols.append(e.ravel())
        G = np.array(cols).T
        assert np.allclose(M @ G, np.eye(3), atol=1e-12 * np.abs(M).max() * np.abs(G).max())


def test_green_matches_direct_spectral_sum(oracle):
    """G_avg against an explicit sum over shifted-frequency plane waves (no FFT)."""
    o = oracle
    shape = (4, 3, 4)
    dk = tuple(2 * np.pi / (shape[d] * RES[d]) for d in range(3))
    k2c = (1.6713e7) ** 2 + 1j * 7.6e13
    gp = o.GreenParams(k_square=k2c, freq_res=dk, shifts=(0.25, 0.0, 0.25), use_averaging=True)
    u = random_field(shape, 5)
    got = o.green_conv(u, gp)
    n = np.indices(shape).reshape(3, -1).T  # voxel positions
    N = np.prod(shape)
    want = np.zeros((N, 3), dtype=complex)
    uf = np.array([u[tuple(v)] for v in n])
    for sgn in (+1, -1):
        s = np.array([0.25, 0.0, 0.25]) * sgn
        for m in np.indices(shape).reshape(3, -1).T:
            mm = np.array([((m[d] + shape[d] // 2) % shape[d]) - shape[d] // 2 for d in range(3)])
            k = np.array(dk) * (mm + s)
            phase = np.exp(2j * np.pi * (n @ ((m + s) / np.array(shape))))
            coef = (np.conj(phase)[:, None] * uf).sum(0) / N
            want += 0.5 * phase[:, None] * (_ghat(k, k2c) @ coef)[None, :]
    got_flat = np.array([got[tuple(v)] for v in n])
    assert rel(got_flat, want) < 1e-12


# ---------------------------------------------------------------------------
# curl
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(8, 6, 4), (5, 7, 3)])
def test_curl_of_modes(oracle, shape):
    """curl of p exp(2 pi i n.m/N) = i k x p (...), k from the curl index map
    (i <= div(N,2) ? i-1 : i-N-1): for odd N the mode (N-1)/2 maps to -(N+1)/2."""
    rng = np.random.default_rng(6)
    n = np.indices(shape)
    for trial in range(5):
        m = [int(rng.integers(0, shape[d])) for d in range(3)]
        if trial == 0:
            m = [(shape[d] - 1) // 2 for d in range(3)]
        p = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        wave = np.exp(2j * np.pi * sum(n[d] * m[d] / shape[d] for d in range(3)))
        u = np.asfortranarray(wave[..., None] * p)
        got = oracle.curl_conv(u, RES)
        idx = [m[d] if m[d] + 1 <= shape[d] // 2 else m[d] - shape[d] for d in range(3)]
        k = np.array([2 * np.pi / (shape[d] * RES[d]) * idx[d] for d in range(3)])
        want = wave[..., None] * np.cross(1j * k, p)
        assert rel(got, want) < 1e-12


# ---------------------------------------------------------------------------
# incident plane wave -- test/sources/plane_wave.jl
# ---------------------------------------------------------------------------
def test_incident_field_reference_tests(oracle):
    o = oracle
    s = o.Solver(resolution=(50e-9,) * 3, grid_size=(32, 32, 32), boundary_conditions=o.PeriodicBC())
    E, H = o.incident_field(s, o.PlaneWave(wavelength=633e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1]))
    assert E.shape == (32, 32, 32, 3) and np.isfinite(E).all() and np.isfinite(H).all()
    c = (16, 16, 16)
    assert abs(E[c + (0,)]) / abs(H[c + (1,)]) == pytest.approx(376.730313668, rel=0.1)
    assert abs(E[c + (0,)]) / abs(H[c + (1,)]) == pytest.approx(377.0, rel=1e-14)
    s2 = o.Solver(resolution=(100e-9,) * 3, grid_size=(32, 32, 16), boundary_conditions=o.PeriodicBC())
    Ex, _ = o.incident_field(s2, o.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1]))
    Ey, _ = o.incident_field(s2, o.PlaneWave(wavelength=500e-9, polarization=[0, 1, 0], k_vector=[0, 0, 1]))
    assert np.sum(np.abs(Ex * np.conj(Ey))) == pytest.approx(0.0, abs=1e-12)


def test_incident_field_closed_form(oracle):
    o = oracle
    bc = o.AbsorbingBC(thickness=200e-9, attenuation_thickness=200e-9)
    s = o.Solver(resolution=(50e-9, 40e-9, 30e-9), grid_size=(6, 5, 4), boundary_conditions=bc)
    kv = np.array([0.3, -0.2, math.sqrt(1 - 0.13)])
    pol = np.cross(kv, [0, 0, 1.0]) / np.linalg.norm(np.cross(kv, [0, 0, 1.0]))
    src = o.PlaneWave(wavelength=532e-9, polarization=pol, k_vector=kv, amplitude=2.0, phase=0.3)
    eb = 1.4 ** 2
    E, H = o.incident_field(s, src, eb)
    kb = 2 * np.pi / 532e-9 * 1.4
    pad = s.padding
    for idx in [(0, 0, 0), (3, 2, 1), (E.shape[0] - 1, E.shape[1] - 1, E.shape[2] - 1)]:
        r = np.array([(idx[d] - pad[d]) * s.resolution[d] - (s.grid_size[d] - 1) * s.resolution[d] / 2
                      for d in range(3)])
        a = 2.0 * np.exp(1j * (kb * kv @ r + 0.3))
        assert np.allclose(E[idx], a * pol, rtol=1e-13, atol=1e-15)
        assert np.allclose(H[idx], a * np.cross(kv, pol) * 1.4 / 377.0, rtol=1e-13, atol=1e-17)
    # transversality and coherent superposition (field_generation.jl:64-75)
    assert abs(np.einsum("...c,c->...", E, kv)).max() < 1e-13 * abs(E).max()
    src2 = o.PlaneWave(wavelength=532e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    Es, Hs = o.incident_field_sum(s, [src, src2], eb)
    E2, H2 = o.incident_field(s, src2, eb)
    assert np.allclose(Es, E + E2) and np.allclose(Hs, H + H2)


# ---------------------------------------------------------------------------
# whole solve
# ---------------------------------------------------------------------------
def _tiny(oracle, bc, seed=1, grid=(6, 5, 4)):
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bc, tolerance=1e-11)
    rng = np.random.default_rng(seed)
    eps = np.asfortranarray((N_BG ** 2 + 0.5 * rng.random(grid)).astype(complex))
    return s, eps


@pytest.mark.parametrize("absorb", [False, True])
def test_solvers_match_dense_solve(oracle, absorb):
    o = oracle
    bc = o.AbsorbingBC(thickness=100e-9, attenuation_thickness=100e-9) if absorb else o.PeriodicBC()
    s, eps = _tiny(o, bc)
    src = o.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    res = {}
    # the Born (Richardson) tail on the absorbing grid needs ~2e4 sweeps for
    # 1e-11: there only the Krylov methods run against the dense solve
    methods = ("bicgstab", "gmres") if absorb else ("born", "bicgstab", "gmres")
    for m in methods:
        r = o.solve(s, src, eps, N_BG ** 2, method=m, rule="rhs", maxiters=6000)
        assert r.stats.converged, m
        res[m] = r
    init = res["gmres"].init
    A = o.dense_operator(init.potential, init.green)
    x = np.linalg.solve(A, o.flat(res["gmres"].rhs))
    # residual 1e-11 times the condition of the absorbing operator (~1e3)
    for m, r in res.items():
        assert rel(o.flat(r.E_scat), x) < (5e-8 if absorb else 5e-10), m


def test_spectral_pde_residual(oracle):
    """Periodic: the converged x solves M(k) F[x] = F[V x + S] with the Maxwell
    symbol M(k) = |k|^2 - k k^T - kc^2 (independent of the Green formula)."""
    o = oracle
    s, eps = _tiny(o, o.PeriodicBC(), seed=2, grid=(8, 6, 4))
    r = o.solve(s, o.PlaneWave(wavelength=500e-9, polarization=[0, 1, 0], k_vector=[0, 0, 1]),
                eps, N_BG ** 2, method="gmres", rule="rhs", maxiters=500)
    x, V, S = r.E_scat, r.init.potential, r.source_term
    k2c = r.init.green.k_square
    X = o.fft3(x)
    RHS = o.fft3(V[..., None] * x + S)
    dk = r.init.green.freq_res
    kk = [dk[d] * o.green_freq_index(x.shape[d]) for d in range(3)]
    K = np.stack(np.meshgrid(*kk, indexing="ij"), axis=-1)
    kdotX = np.einsum("...c,...c->...", K, X)
    MX = (np.einsum("...c,...c->...", K, K) - k2c)[..., None] * X - K * kdotX[..., None]
    assert rel(MX, RHS) < 1e-9


def test_homogeneous_medium(oracle):
    o = oracle
    s = o.Solver(resolution=RES, grid_size=(8, 8, 6), boundary_conditions=o.PeriodicBC())
    eps = np.full((8, 8, 6), N_BG ** 2, dtype=complex)
    src = o.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    r = o.solve(s, src, eps, N_BG ** 2, method="born", rule="rhs")
    E_inc, H_inc = o.incident_field(s, src, N_BG ** 2)
    assert r.stats.converged and np.abs(r.E_scat).max() == 0.0
    assert np.array_equal(r.E, E_inc) and np.array_equal(r.H, H_inc)


def test_uniform_medium_on_grid_incidence(oracle):
    """Uniform contrast, periodic box, incident k on the FFT grid: the only
    periodic solution is E_tot = 0 and H_tot = 0 (one Fourier mode, 3x3 algebra)."""
    o = oracle
    g = (4, 4, 8)
    dz = 50e-9
    m = 2
    lam = g[2] * dz * N_BG / m  # k_bg = m * 2 pi / (Nz dz)
    s = o.Solver(resolution=(50e-9, 50e-9, dz), grid_size=g, boundary_conditions=o.PeriodicBC(),
                 tolerance=1e-12)
    eps = np.full(g, 1.6 ** 2, dtype=complex)
    src = o.PlaneWave(wavelength=lam, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    r = o.solve(s, src, eps, N_BG ** 2, method="gmres", rule="rhs", maxiters=200)
    E_inc, H_inc = o.incident_field(s, src, N_BG ** 2)
    assert np.abs(r.E).max() < 1e-10 * np.abs(E_inc).max()
    assert np.abs(r.H).max() < 1e-9 * np.abs(H_inc).max()


def test_mirror_symmetry_centred_bead(oracle):
    """x-polarised normal incidence on a centred sphere: Ex is even in x and y,
    Ey odd in x.  Odd grid lengths along the mirrored axes: an even length has
    a Nyquist bin shared by +k and -k, which the discrete operator maps to -N/2
    only and so is not exactly mirror symmetric.  Absorbing only along x: the
    x mirror maps the (+s,0,0) half onto the (-s,0,0) half.  (With every axis
    absorbing the reference averages only the (+s,+s,+s) and (-s,-s,-s)
    halves, which no single-axis mirror maps onto each other.)"""
    o = oracle
    g = (9, 9, 8)
    from frequencymaxwell import phantom_bead
    med = phantom_bead(g, [1.5 ** 2], 3.0, permittivity_bg=N_BG ** 2)
    ab = o.AbsorbingBC(thickness=200e-9, attenuation_thickness=200e-9)
    for bc in (o.PeriodicBC(), (ab, o.PeriodicBC(), o.PeriodicBC())):
        s = o.Solver(resolution=RES, grid_size=g, boundary_conditions=bc, tolerance=1e-10)
        r = o.solve(s, o.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1]),
                    med.permittivity, N_BG ** 2, method="gmres", rule="rhs", maxiters=400)
        assert r.stats.converged
        Ex, Ey = r.E[..., 0], r.E[..., 1]
        assert rel(Ex[::-1, :, :], Ex) < 1e-8
        assert rel(Ex[:, ::-1, :], Ex) < 1e-8
        assert rel(Ey[::-1, :, :], -Ey) < 1e-8


def test_linearity_in_source(oracle):
    o = oracle
    s, eps = _tiny(o, o.PeriodicBC(), seed=4)
    init = o.initialize_solver(s, eps, N_BG ** 2, 500e-9)
    u, v = random_field(init.potential.shape, 7), random_field(init.potential.shape, 8)
    a, b = 0.3 - 1.2j, 2.0 + 0.5j
    G = lambda f: o.green_conv(f, init.green)
    assert rel(G(a * u + b * v), a * G(u) + b * G(v)) < 1e-13


# ---------------------------------------------------------------------------
# Krylov iterates at fixed budgets: scipy and brute-force characterisations
# ---------------------------------------------------------------------------
def _krylov_case(oracle, absorb, grid=(6, 5, 4), seed=11):
    """Tiny problem with the dense left-preconditioned operator A^ = gamma A,
    b^ = gamma b and the reference's warm start x0 = rhs (convergent_born.jl:727-728)."""
    o = oracle
    bc = o.AbsorbingBC(thickness=100e-9, attenuation_thickness=100e-9) if absorb else o.PeriodicBC()
    s, eps, init = oracle_setup(o, grid, bc, seed)
    V, ei, gp = init.potential, init.eps_imag, init.green
    S = random_field(V.shape, seed + 1)
    rhs = o.green_conv(S, gp)
    # unit-scale right-hand side (G ~ 1/k^2 leaves b ~ 1e-14; scipy's breakdown
    # tests are absolute); the iterations are scale-equivariant
    rhs = np.asfortranarray(rhs * (math.sqrt(rhs.size) / np.linalg.norm(rhs)))
    Ad = o.dense_operator(V, gp)
    gam = (1j / ei) * np.tile(o.flat(V), 3)
    Ah = gam[:, None] * Ad
    bh = gam * o.flat(rhs)
    return o, V, ei, gp, rhs, Ah, bh


def _scipy_op(Ah):
    import scipy.sparse.linalg as sla
    n = Ah.shape[0]
    return sla.LinearOperator((n, n), matvec=lambda v: Ah @ v, dtype=complex)


def _orth_krylov(M, v, k):
    """Orthonormal basis of span{v, M v, ..., M^(k-1) v} (Arnoldi, two passes of
    Gram-Schmidt per vector)."""
    Q = np.zeros((v.size, k), dtype=complex)
    Q[:, 0] = v / np.linalg.norm(v)
    for j in range(1, k):
        w = M @ Q[:, j - 1]
        for _ in range(2):
            w = w - Q[:, :j] @ (Q[:, :j].conj().T @ w)
        Q[:, j] = w / np.linalg.norm(w)
    return Q


@pytest.mark.parametrize("absorb", [False, True])
def test_bicgstab_iterates_match_scipy(oracle, absorb):
    """Shadow r^ = r0 (scipy's choice): the oracle's k-th BiCGStab iterate equals
    scipy.sparse.linalg.bicgstab's after k iterations on the same explicit
    left-preconditioned system, for k = 1..8."""
    import scipy.sparse.linalg as sla
    o, V, ei, gp, rhs, Ah, bh = _krylov_case(oracle, absorb)
    op = _scipy_op(Ah)
    for k in range(1, 9):
        xs, info = sla.bicgstab(op, bh, x0=o.flat(rhs).copy(), rtol=1e-300, atol=0.0, maxiter=k)
        assert info == k
        x, st = o.solve_bicgstab(rhs, rhs, V, ei, gp, 0.0, k, shadow="r0")
        assert st.iters == k and st.applications == 2 * k + 1 and not st.converged
        assert rel(o.flat(x), xs) < 1e-10, k


@pytest.mark.parametrize("absorb", [False, True])
def test_gmres_iterates_match_scipy(oracle, absorb):
    """Full GMRES after k Arnoldi steps (scipy: one cycle of restart k) and
    GMRES(4) over two and three whole cycles equal scipy.sparse.linalg.gmres."""
    import scipy.sparse.linalg as sla
    o, V, ei, gp, rhs, Ah, bh = _krylov_case(oracle, absorb)
    op = _scipy_op(Ah)
    x0 = o.flat(rhs)
    for k in range(1, 9):
        xs, _ = sla.gmres(op, bh, x0=x0.copy(), rtol=1e-300, atol=0.0, restart=k, maxiter=1,
                          callback=lambda r: None, callback_type="pr_norm")
        x, st = o.solve_gmres(rhs, rhs, V, ei, gp, 0.0, k)
        assert st.iters == k and st.applications == k + 1
        assert rel(o.flat(x), xs) < 1e-10, k
    for cycles in (2, 3):
        xs, _ = sla.gmres(op, bh, x0=x0.copy(), rtol=1e-300, atol=0.0, restart=4, maxiter=cycles,
                          callback=lambda r: None, callback_type="pr_norm")
        x, st = o.solve_gmres(rhs, rhs, V, ei, gp, 0.0, 4 * cycles, restart=4)
        # one application per Arnoldi step, one for r0 and one per restart
        assert st.iters == 4 * cycles and st.applications == 4 * cycles + cycles
        assert rel(o.flat(x), xs) < 1e-10, cycles


@pytest.mark.parametrize("absorb", [False, True])
def test_gmres_iterates_minimise_residual(oracle, absorb):
    """The k-th full-GMRES iterate is x0 + W y with y the least-squares minimiser
    of ||r0 - A^ W y|| over an orthonormal basis W of the Krylov space
    K_k(A^, r0); restarted GMRES(3) repeats that from each cycle's residual."""
    o, V, ei, gp, rhs, Ah, bh = _krylov_case(oracle, absorb)
    x0 = o.flat(rhs)
    r0 = bh - Ah @ x0
    for k in range(1, 9):
        W = _orth_krylov(Ah, r0, k)
        y = np.linalg.lstsq(Ah @ W, r0, rcond=None)[0]
        x, _ = o.solve_gmres(rhs, rhs, V, ei, gp, 0.0, k)
        assert rel(o.flat(x), x0 + W @ y) < 1e-9, k
    xb = x0.copy()
    for _ in range(3):
        r = bh - Ah @ xb
        W = _orth_krylov(Ah, r, 3)
        xb = xb + W @ np.linalg.lstsq(Ah @ W, r, rcond=None)[0]
    x, st = o.solve_gmres(rhs, rhs, V, ei, gp, 0.0, 9, restart=3)
    assert st.iters == 9 and rel(o.flat(x), xb) < 1e-9


@pytest.mark.parametrize("absorb", [False, True])
@pytest.mark.parametrize("shadow", ["b", "r0"])
def test_bicgstab_iterates_petrov_galerkin(oracle, absorb, shadow):
    """BiCGStab characterised without its recurrences (van der Vorst 1992):
    r_k = Omega_k(A^) phi_k(A^) r0, where phi_k (degree k, phi_k(0) = 1) is the
    BiCG residual polynomial, fixed by <(A^H)^i c, phi_k(A^) r0> = 0 for i < k,
    and Omega_k = prod_j (1 - omega_j t) with omega_j minimising
    ||(1 - omega A^) s_j||, s_j = Omega_(j-1)(A^) phi_j(A^) r0.  The k-th iterate
    is x0 + A^-1 (r0 - r_k).  Both shadow vectors c: b (Krylov.jl's default) and
    r0."""
    o, V, ei, gp, rhs, Ah, bh = _krylov_case(oracle, absorb)
    x0 = o.flat(rhs)
    r0 = bh - Ah @ x0
    c = x0.copy() if shadow == "b" else r0.copy()
    AhH = Ah.conj().T
    omegas = []
    for k in range(1, 7):
        # phi_k(A^) r0 = r0 + Q y over an orthonormal basis Q of A^ K_k(A^, r0);
        # the conditions are taken against an orthonormal basis P of K_k(A^H, c)
        Q, _ = np.linalg.qr(Ah @ _orth_krylov(Ah, r0, k))
        P = _orth_krylov(AhH, c, k)
        y = np.linalg.solve(P.conj().T @ Q, -(P.conj().T @ r0))
        s = r0 + Q @ y
        for w in omegas:
            s = s - w * (Ah @ s)
        t = Ah @ s
        w = np.vdot(t, s) / np.vdot(t, t)
        omegas.append(w)
        rk = s - w * t
        xk = x0 + np.linalg.solve(Ah, r0 - rk)
        x, st = o.solve_bicgstab(rhs, rhs, V, ei, gp, 0.0, k, shadow=shadow)
        assert st.iters == k
        assert rel(o.flat(x), xk) < 1e-8, k


def test_krylov_counts_and_budget(oracle):
    """Krylov.jl counting: maxiters bounds iterations (BiCGStab: two
    applications each; GMRES: one Arnoldi step each), the warm-start residual
    is an application outside the count, and a converged solve stops at the
    first iteration whose residual meets the threshold."""
    o, V, ei, gp, rhs, Ah, bh = _krylov_case(oracle, True)
    x, st = o.solve_bicgstab(rhs, rhs, V, ei, gp, 1e-9, 500)
    assert st.converged and st.applications == 2 * st.iters + 1
    assert st.history[-1] <= 1e-9 < st.history[-2]
    x, st = o.solve_gmres(rhs, rhs, V, ei, gp, 1e-9, 500)
    assert st.converged and st.applications == st.iters + 1
    assert st.history[-1] <= 1e-9 < st.history[-2]
    # the GMRES estimate is the true residual of the returned iterate
    assert abs(o.true_residual(x, rhs, V, ei, gp) - st.resid) < 1e-6 * st.resid
    x, st = o.solve_gmres(rhs, rhs, V, ei, gp, 1e-9, 500, restart=20)
    cycles = (st.iters + 19) // 20
    assert st.converged and st.applications == st.iters + cycles


# ---------------------------------------------------------------------------
# field analysis (electromagnetic_field.jl:112-235)
# ---------------------------------------------------------------------------
def test_field_analysis_closed_forms(oracle):
    o = oracle
    n = (5, 4, 6)
    res = (30e-9, 40e-9, 50e-9)
    # constant fields: energy = (eps0 N |E|^2 + mu0 N |H|^2) / 2 * voxel volume
    E = np.zeros(n + (3,), complex)
    H = np.zeros(n + (3,), complex)
    E[..., 1] = 2.0 - 1.0j
    H[..., 2] = 0.5j
    N = np.prod(n)
    vol = res[0] * res[1] * res[2]
    want = (8.854187817e-12 * N * 5.0 + 4e-7 * math.pi * N * 0.25) / 2 * vol
    assert o.field_energy(E, H, res) == pytest.approx(want, rel=1e-14)
    # right-hand rule: y x z = x, z x x = y, x x y = z (conj(H) of real unit H)
    for a, b, c in ((1, 2, 0), (2, 0, 1), (0, 1, 2)):
        Eu = np.zeros(n + (3,), complex)
        Hu = np.zeros(n + (3,), complex)
        Eu[..., a] = 1.0
        Hu[..., b] = 1.0
        S = o.poynting_vector(Eu, Hu)
        assert np.all(S[..., c] == 1.0) and np.all(np.delete(S, c, axis=3) == 0.0)
    # plane wave x-polarised along z: S_z = |A|^2 / Z, intensity |A|^2 in x
    z = np.arange(n[2]) * 50e-9
    A = 0.7 * np.exp(2j * np.pi * z / 500e-9)
    Z = 377.0 / 1.33
    E = np.zeros(n + (3,), complex)
    H = np.zeros(n + (3,), complex)
    E[..., 0] = A
    H[..., 1] = A / Z
    S = o.poynting_vector(E, H)
    assert np.allclose(S[..., 2], 0.49 / Z, rtol=1e-14, atol=0)
    assert np.allclose(o.field_intensity(E), 0.49, rtol=1e-14)
    assert np.allclose(o.field_intensity(E, "x"), 0.49, rtol=1e-14)
    assert np.all(o.field_intensity(E, "y") == 0) and np.all(o.field_intensity(E, "z") == 0)
    # planes: 1-based index along each axis
    F = random_field(n, 3)
    assert np.array_equal(o.extract_plane(F, 1, 2)[0], F[1])
    assert np.array_equal(o.extract_plane(F, 2, 4)[:, 0], F[:, 3])
    assert np.array_equal(o.extract_plane(F, 3, 6)[:, :, 0], F[:, :, 5])
    assert o.extract_plane(F, 3, 1).shape == (5, 4, 1, 3)
