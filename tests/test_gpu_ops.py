"""GPU parity of the operator layer: Green apply, CBS operator, preconditioner, curl.

Each test builds one problem on the oracle and sends the same padded
potential and Green parameters through the C ABI.  Bars: FP64 <= 1e-12
relative (the FFT roundoff is ~1e-15); FP32 <= 2e-6 relative.
"""

import numpy as np
import pytest

from cases import N_BG, RES, oracle_setup, random_field, rel, spec_from

pytestmark = pytest.mark.gpu

PERIODIC_GRIDS = [(16, 16, 8), (32, 16, 8), (6, 5, 4), (64, 64, 32), (24, 20, 12), (96, 32, 8)]
ABSORB_GRIDS = [(16, 16, 8), (12, 10, 6)]
FP64_TOL = 1e-12
FP32_TOL = 2e-6


# every compile-time specialised length, plus two runtime-plan lengths
CT_LENGTHS = (8, 16, 32, 64, 96, 128, 160, 192, 224, 256, 288, 320)


def _absorbing(o):
    return o.AbsorbingBC(thickness=400e-9, attenuation_thickness=300e-9)


@pytest.mark.parametrize("grid", PERIODIC_GRIDS)
def test_green_fp64_periodic(fm, oracle, grid):
    print(f"case green fp64 periodic {grid}", flush=True)
    _, _, init = oracle_setup(oracle, grid, oracle.PeriodicBC(), seed=1)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp64")
    u = random_field(init.potential.shape, 2)
    got = prob.green_apply(u)
    want = oracle.green_conv(u, init.green)
    assert rel(got, want) <= FP64_TOL


@pytest.mark.parametrize("grid", ABSORB_GRIDS)
def test_green_fp64_absorbing(fm, oracle, grid):
    print(f"case green fp64 absorbing {grid}", flush=True)
    _, _, init = oracle_setup(oracle, grid, _absorbing(oracle), seed=3)
    assert init.green.use_averaging
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp64")
    u = random_field(init.potential.shape, 4)
    got = prob.green_apply(u)
    want = oracle.green_conv(u, init.green)
    assert rel(got, want) <= FP64_TOL


def test_green_mixed_axes(fm, oracle):
    print("case green fp64 mixed periodic/absorbing", flush=True)
    bc = (oracle.PeriodicBC(), oracle.PeriodicBC(), _absorbing(oracle))
    _, _, init = oracle_setup(oracle, (16, 8, 8), bc, seed=5)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp64")
    u = random_field(init.potential.shape, 6)
    assert rel(prob.green_apply(u), oracle.green_conv(u, init.green)) <= FP64_TOL


@pytest.mark.parametrize("grid,absorb", [((32, 32, 16), False), ((16, 16, 8), True)])
def test_green_fp32(fm, oracle, grid, absorb):
    print(f"case green fp32 {grid} absorb={absorb}", flush=True)
    bc = _absorbing(oracle) if absorb else oracle.PeriodicBC()
    _, _, init = oracle_setup(oracle, grid, bc, seed=7)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp32")
    u = random_field(init.potential.shape, 8)
    assert rel(prob.green_apply(u), oracle.green_conv(u, init.green)) <= FP32_TOL


@pytest.mark.parametrize("absorb", [False, True])
def test_operator_fp64(fm, oracle, absorb):
    print(f"case operator fp64 absorb={absorb}", flush=True)
    bc = _absorbing(oracle) if absorb else oracle.PeriodicBC()
    _, _, init = oracle_setup(oracle, (16, 12, 8), bc, seed=9)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp64")
    x = random_field(init.potential.shape, 10)
    y = random_field(init.potential.shape, 11)
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    got = prob.operator_apply(x, y, a, b)
    want = oracle.cbs_operator(x, init.potential, init.green, alpha=a, beta=b, y=y)
    assert rel(got, want) <= FP64_TOL
    got0 = prob.operator_apply(x)
    assert rel(got0, oracle.cbs_operator(x, init.potential, init.green)) <= FP64_TOL


def test_precond(fm, oracle):
    print("case precond fp64/fp32", flush=True)
    _, _, init = oracle_setup(oracle, (8, 8, 4), oracle.PeriodicBC(), seed=12)
    x = random_field(init.potential.shape, 13)
    want = oracle.cbs_precond(x, init.potential, init.eps_imag)
    p64 = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp64")
    assert rel(p64.precond_apply(x), want) <= 1e-15
    p32 = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp32")
    assert rel(p32.precond_apply(x), want) <= FP32_TOL


@pytest.mark.parametrize("grid", [(16, 16, 8), (7, 5, 6), (32, 8, 16)])
def test_curl_fp64(fm, oracle, grid):
    print(f"case curl fp64 {grid}", flush=True)
    _, _, init = oracle_setup(oracle, grid, oracle.PeriodicBC(), seed=14)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision="fp64")
    e = random_field(init.potential.shape, 15)
    lam = 500e-9
    got = prob.curl_apply(e, oracle.h_scale(lam))
    want = oracle.magnetic_field(e, RES, lam)
    assert rel(got, want) <= FP64_TOL


@pytest.mark.parametrize("L", CT_LENGTHS + (200, 250))
@pytest.mark.parametrize("avg", [False, True])
def test_green_every_length(fm, oracle, L, avg):
    """Every specialised transform length on every axis (the other two axes
    short runtime-plan lengths), FP64 and FP32, with and without the +/-
    subpixel averaging."""
    k_bg = 2 * np.pi * 1.33 / 500e-9
    eps = 0.3 * k_bg ** 2
    for axis in range(3):
        shape = [8, 6, 4]
        shape[axis] = L
        shape = tuple(shape)
        print(f"case green every length {shape} avg={avg}", flush=True)
        gp = oracle.GreenParams(k_square=k_bg ** 2 + 1j * eps,
                                freq_res=tuple(2 * np.pi / (shape[d] * RES[d]) for d in range(3)),
                                shifts=(0.25,) * 3 if avg else (0.0,) * 3, use_averaging=avg)
        spec = fm.GreenSpec(shape=shape, resolution=RES, shifts=gp.shifts, use_averaging=avg,
                            k_square=gp.k_square, eps_imag=eps)
        V = np.asfortranarray(np.full(shape, -1j * eps, dtype=np.complex128))
        u = random_field(shape, 7 + axis)
        want = oracle.green_conv(u, gp)
        for prec, tol in (("fp64", FP64_TOL), ("fp32", FP32_TOL)):
            prob = fm.CBSProblem(V, spec, precision=prec)
            assert rel(prob.green_apply(u), want) <= tol, (shape, prec)


def test_grid_beyond_32bit_tile_offsets_rejected(fm):
    """The y and z tiles address the three spectrum components of a shift
    pipeline with 32-bit element offsets: a grid whose pipeline holds 2^32
    elements or more is refused before anything is read or allocated."""
    import ctypes as C

    from frequencymaxwell import _lib as L
    from frequencymaxwell import engine

    print("case grid 4096x4096x100 rejected", flush=True)
    k2 = (2 * np.pi * 1.33 / 500e-9) ** 2
    spec = engine.GreenSpec(shape=(4096, 4096, 100), resolution=tuple(RES), shifts=(0.25,) * 3,
                            use_averaging=True, k_square=k2 + 0.1j * k2, eps_imag=0.1 * k2)
    d = engine.make_desc(spec, "fp32")
    V = np.zeros(4, dtype=np.complex128)  # never read: validation fails first
    h = C.c_void_p()
    lib = L.load()
    with pytest.raises(ValueError, match="2\\^32"):
        L.check(lib.cbs_problem_create(L.default_context().handle, C.byref(d), L.dptr(V), C.byref(h)))
    assert not h.value


# x lengths whose inverse transform reads the spectra in its first stage
# (two-stage plans with at least 16 butterflies per pencil), with line counts
# (ny * nz of the padded grid) that leave the last x tile partly filled; and
# two-stage z lengths (32, 96) under an x extent of 20, which leaves the last
# z tile (16 or 8 x columns) partly filled while it stages its k values
PARTIAL_TILE_GRIDS = [((64, 7, 5), False), ((96, 5, 3), False), ((160, 3, 7), False),
                      ((56, 5, 3), True), ((20, 6, 32), False), ((20, 4, 96), False)]


@pytest.mark.parametrize("grid, absorb", PARTIAL_TILE_GRIDS)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_x_tiles_partly_filled(fm, oracle, grid, absorb, precision):
    """The x and z passes on tiles holding fewer lines (x) or columns (z) than
    their width: the x inverse reads the spectra of the missing lines from the
    tile's first line and drops their elements in the epilogue; the z tile
    zeroes the k values of its missing columns.  The operator (plain x inverse),
    fused Born sweeps and BiCGStab (preconditioned-matvec partial sums) match
    the oracle at a fixed budget."""
    print(f"case x tiles partly filled {grid} absorb={absorb} {precision}", flush=True)
    bc = _absorbing(oracle) if absorb else oracle.PeriodicBC()
    s, _, init = oracle_setup(oracle, grid, bc, seed=17)
    tol = FP64_TOL if precision == "fp64" else FP32_TOL
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision=precision)
    x = random_field(init.potential.shape, 18)
    if precision == "fp64":
        got = prob.operator_apply(x)
        assert rel(got, oracle.cbs_operator(x, init.potential, init.green)) <= tol
    src = oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    E_inc, _ = oracle.incident_field(s, src, N_BG ** 2)
    S = oracle.source_term(init.potential, E_inc, init.eps_imag)
    rhs = oracle.green_conv(S, init.green)
    solve_tol = 1e-9 if precision == "fp64" else 2e-5
    for method, budget in (("born", 6), ("bicgstab", 3)):
        fn = oracle.solve_born if method == "born" else oracle.solve_bicgstab
        x_o, st_o = fn(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0, budget)
        E, _, st = prob.solve(S, fm.make_opts(method=method, precision=precision, tol=0.0,
                                              maxiters=budget), want_h=False)
        assert st.iters == st_o.iters == budget, method
        assert rel(E, x_o) <= solve_tol, method
