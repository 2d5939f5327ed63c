"""Solve parity at the configured workload sizes (the sizes the benchmark
measures), so the kernels specialised and tuned for those axis lengths are
checked against the oracle where they are credited.

* bench workload: 128x128x64 bead (n 1.5, r 16 px) in water, absorbing 16 px
  every axis -> padded 160x160x96, averaged Green: 12 fixed Born sweeps
  through the whole solve path, E and H against the oracle (FP64 <= 1e-10,
  FP32 <= 2e-5);
* the same workload with GMRES (restarted inside the budget) and BiCGStab at
  a fixed budget, FP64 <= 1e-9 and FP32 <= 2e-5 against the oracle's iterates;
* 64x64x32 bead (r 8 px, phantom_bead, src/medium/phantoms.jl:58-106),
  periodic: 20 fixed FP64 sweeps <= 1e-10; the same bead with 8 px absorbing
  padding converged to 1e-6 (FP32+refine) against the oracle converged to
  1e-10;
* 256x256x128, 20 spheres, absorbing 16 px -> 288x288x160 at 532 nm: 3 fixed
  sweeps in FP64 (<= 1e-10) and FP32 (<= 2e-5);
* 128 right-hand sides (64 directions in the NA 1.2 disk x 2 polarisations)
  of the 128x128x64 cell phantom in one cbs_solve_scene_multi call, short
  budget: every right-hand side equals its single solve bit for bit; the
  first and last in FP64 against the oracle;
* one serving mix holding all five stream grids (64^3 ... 256x256x128, 16 px
  absorbing padding): each request's fields are the same bits alone and in
  the mix; the 64^3 one in FP64 against the oracle.
"""

import numpy as np
import pytest

import bench
from cases import N_BG, RES, rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 2e-5


def _absorbing(fm, oracle, px=16):
    t = 2 * px * 50e-9  # padding_pixels = round(thickness / (2 dx))
    return (fm.AbsorbingBoundaryCondition(thickness=t, attenuation_thickness=t, sharpness=0.9),
            oracle.AbsorbingBC(thickness=t, attenuation_thickness=t, sharpness=0.9))


def _fixed_solve(fm, oracle, med, bcs, sweeps, precision, wavelength=500e-9, pol=(1, 0, 0),
                 k=(0, 0, 1)):
    grid = med.permittivity.shape
    bf, bo = bcs
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bf,
                                     tolerance=0.0, linear_solver="born", stop_rule="rhs",
                                     maxiters=sweeps, precision=precision)
    src = fm.PlaneWaveSource(wavelength=wavelength, polarization=list(pol), k_vector=list(k))
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # a fixed budget ends unconverged
        out = fm.solve(solver, src, med)
    assert solver.iteration_count == sweeps
    return out


def _oracle_fixed(oracle, med, bo, sweeps, wavelength=500e-9, pol=(1, 0, 0), k=(0, 0, 1)):
    grid = med.permittivity.shape
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bo, tolerance=0.0)
    src = oracle.PlaneWave(wavelength=wavelength, polarization=list(pol), k_vector=list(k))
    ref = oracle.solve(s, src, med.permittivity, N_BG ** 2, method="born", rule="rhs",
                       maxiters=sweeps)
    assert ref.stats.iters == sweeps
    return ref


@pytest.mark.timeout(300)
def test_bench_workload_fixed_sweeps(fm, oracle):
    print("case bench workload 160x160x96 averaged, 12 sweeps", flush=True)
    med = fm.phantom_bead((128, 128, 64), [1.5 ** 2], 16.0, permittivity_bg=N_BG ** 2)
    bcs = _absorbing(fm, oracle)
    ref = _oracle_fixed(oracle, med, bcs[1], 12)
    assert ref.init.potential.shape == (160, 160, 96)
    for precision, tol in (("fp64", TOL64), ("fp32", TOL32)):
        out = _fixed_solve(fm, oracle, med, bcs, 12, precision)
        eE, eH = rel(out.E, ref.E), rel(out.H, ref.H)
        print(f"  {precision}: E {eE:.2e} H {eH:.2e}", flush=True)
        assert eE <= tol and eH <= tol, precision


@pytest.mark.timeout(300)
@pytest.mark.parametrize("method, budget, restart", [("gmres", 7, 3), ("bicgstab", 3, 0)])
def test_bench_workload_krylov_fixed_budget(fm, oracle, method, budget, restart):
    """The Krylov solvers on the benchmark workload (160x160x96, averaged) at a
    fixed budget through the whole solve path: GMRES with a restart inside the
    budget (delayed CGS2, two cycles closed by the final reduce and one
    cut short) and BiCGStab, E and H against the oracle's iterates (FP64
    <= 1e-9, FP32 <= 2e-5), Krylov.jl's iteration counts."""
    print(f"case bench workload {method} budget {budget} restart {restart}", flush=True)
    med = fm.phantom_bead((128, 128, 64), [1.5 ** 2], 16.0, permittivity_bg=N_BG ** 2)
    bf, bo = _absorbing(fm, oracle)
    s = oracle.Solver(resolution=RES, grid_size=med.permittivity.shape, boundary_conditions=bo,
                      tolerance=0.0)
    src_o = oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    ref = oracle.solve(s, src_o, med.permittivity, N_BG ** 2, method=method, rule="rhs",
                       maxiters=budget, restart=restart)
    assert ref.stats.iters == budget
    import warnings
    for precision, tol in (("fp64", 1e-9), ("fp32", TOL32)):
        solver = fm.ConvergentBornSolver(resolution=RES, grid_size=med.permittivity.shape,
                                         boundary_conditions=bf, tolerance=0.0,
                                         linear_solver=method, stop_rule="rhs", maxiters=budget,
                                         gmres_restart=restart, precision=precision)
        src = fm.PlaneWaveSource(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")  # a fixed budget ends unconverged
            out = fm.solve(solver, src, med)
        assert solver.iteration_count == budget
        eE, eH = rel(out.E, ref.E), rel(out.H, ref.H)
        print(f"  {precision}: E {eE:.2e} H {eH:.2e}", flush=True)
        assert eE <= tol and eH <= tol, precision


@pytest.mark.timeout(300)
def test_bead_64x64x32(fm, oracle):
    print("case bead 64x64x32 periodic 20 sweeps / absorbing converged", flush=True)
    med = fm.phantom_bead((64, 64, 32), [1.5 ** 2], 8.0, permittivity_bg=N_BG ** 2)
    per = (fm.PeriodicBoundaryCondition(), oracle.PeriodicBC())
    ref = _oracle_fixed(oracle, med, per[1], 20)
    out = _fixed_solve(fm, oracle, med, per, 20, "fp64")
    assert rel(out.E, ref.E) <= TOL64 and rel(out.H, ref.H) <= TOL64
    # absorbing 8 px (padded 80x80x48): converged solves
    bf, bo = _absorbing(fm, oracle, px=8)
    s = oracle.Solver(resolution=RES, grid_size=(64, 64, 32), boundary_conditions=bo,
                      tolerance=1e-10)
    src_o = oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    ref = oracle.solve(s, src_o, med.permittivity, N_BG ** 2, method="bicgstab", rule="rhs",
                       maxiters=2000)
    assert ref.stats.converged
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=(64, 64, 32),
                                     boundary_conditions=bf, tolerance=1e-6,
                                     linear_solver="born", stop_rule="rhs", maxiters=4000,
                                     precision="fp32_refine")
    src = fm.PlaneWaveSource(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    out = fm.solve(solver, src, med)
    st = solver.last_stats
    print(f"  fp32_refine: {st['iters']} iterations, resid {st['resid_true_fp64']:.2e}, "
          f"E err {rel(out.E, ref.E):.2e}", flush=True)
    assert st["converged"] == 1 and st["resid_true_fp64"] <= 1e-6
    # a residual of 1e-6 bounds the field error by the operator's condition
    assert rel(out.E, ref.E) <= 2e-5 and rel(out.H, ref.H) <= 2e-5


@pytest.mark.timeout(600)
def test_large_phantom_fixed_sweeps(fm, oracle):
    print("case large phantom 288x288x160, 3 sweeps", flush=True)
    med = bench.large_phantom(fm)
    bcs = _absorbing(fm, oracle)
    ref = _oracle_fixed(oracle, med, bcs[1], 3, wavelength=532e-9)
    assert ref.init.potential.shape == (288, 288, 160)
    for precision, tol in (("fp64", TOL64), ("fp32", TOL32)):
        out = _fixed_solve(fm, oracle, med, bcs, 3, precision, wavelength=532e-9)
        eE, eH = rel(out.E, ref.E), rel(out.H, ref.H)
        print(f"  {precision}: E {eE:.2e} H {eH:.2e}", flush=True)
        assert eE <= tol and eH <= tol, precision


@pytest.mark.timeout(600)
def test_batched_illuminations_128_rhs(fm, oracle):
    print("case 128 rhs on the cell phantom", flush=True)
    from frequencymaxwell import solver as S
    from frequencymaxwell.engine import solve_scene_multi

    med = bench.cell_phantom(fm)
    srcs = bench.illuminations(fm, 64)
    assert len(srcs) == 128
    bf, bo = _absorbing(fm, oracle)
    sol = fm.ConvergentBornSolver(resolution=RES, grid_size=(128, 128, 64), boundary_conditions=bf,
                                  tolerance=0.0, linear_solver="born", precision="fp32_refine")
    d, eps, waves, keep = S.scene(sol, srcs, med)
    o = fm.make_opts(method="born", precision="fp32_refine", tol=0.0, stop_rule="rhs",
                     maxiters=8, use_graph=True, check_every=8, compute_h=False)
    E_all, _, sts, info = solve_scene_multi(d, eps, [[w] for w in waves], o, want_h=False)
    assert len(E_all) == 128
    for i, w in enumerate(waves):
        E1, _, s1, _ = solve_scene_multi(d, eps, [[w]], o, want_h=False)
        assert sts[i].iters == s1[0].iters and sts[i].applications == s1[0].applications
        assert np.array_equal(E_all[i], E1[0]), f"right-hand side {i} differs in the batch"
    del E_all
    # FP64 parity of the c4 inputs (GRF phantom, oblique incidence) for two
    # right-hand sides: 8 Born sweeps against the oracle
    o64 = fm.make_opts(method="born", precision="fp64", tol=0.0, stop_rule="rhs", maxiters=8,
                       compute_h=True)
    d64, eps64, waves64, keep64 = S.scene(sol, srcs, med, precision="fp64")
    s = oracle.Solver(resolution=RES, grid_size=(128, 128, 64), boundary_conditions=bo,
                      tolerance=0.0)
    for i in (0, 127):
        E, H, st, _ = solve_scene_multi(d64, eps64, [[waves64[i]]], o64)
        src = srcs[i]
        ref = oracle.solve(s, oracle.PlaneWave(wavelength=500e-9,
                                               polarization=list(src.polarization),
                                               k_vector=list(src.k_vector)),
                           med.permittivity, N_BG ** 2, method="born", rule="rhs", maxiters=8)
        assert rel(E[0], ref.E) <= TOL64 and rel(H[0], ref.H) <= TOL64, i


@pytest.mark.timeout(600)
def test_serving_mix_of_all_stream_grids(fm, oracle):
    print("case serving mix of the five stream grids", flush=True)
    import math

    from frequencymaxwell import solver as S

    rng = np.random.default_rng(17)
    reqs, keep, meta = [], [], []
    for g in bench.SERVE_GRIDS:
        med = bench.random_spheres(fm, g, rng)
        sin_t = math.sqrt(rng.uniform(0.0, 1.0)) / N_BG
        phi = rng.uniform(0.0, 2 * math.pi)
        k = [sin_t * math.cos(phi), sin_t * math.sin(phi), math.sqrt(1 - sin_t ** 2)]
        pol = [-math.sin(phi), math.cos(phi), 0.0]
        src = fm.PlaneWaveSource(wavelength=500e-9, polarization=pol, k_vector=k)
        sol = fm.ConvergentBornSolver(resolution=RES, grid_size=g,
                                      boundary_conditions=bench.boundary(fm))
        d, eps, waves, kp = S.scene(sol, src, med)
        reqs.append((g, d, eps, waves))
        keep.append(kp)  # the window arrays the descriptor points to
        meta.append((med, pol, k))
    opts = [fm.make_opts(method="born", precision="fp32_refine", tol=0.0, maxiters=12,
                         compute_h=True),
            fm.make_opts(method="bicgstab", precision="fp32_refine", tol=0.0, maxiters=6,
                         compute_h=True)]
    res = {}
    for active in (1, 8):
        srv = fm.CBSServer(max_active=active)
        try:
            ts = [srv.submit_scene(d, e, w, opts[i % 2]) for i, (g, d, e, w) in enumerate(reqs)]
            res[active] = [srv.wait(t) for t in ts]
        finally:
            srv.stop()
    for i, (g, d, e, w) in enumerate(reqs):
        (E1, H1, s1), (E2, H2, s2) = res[1][i], res[8][i]
        print(f"  grid {g}: iters {s1.iters}", flush=True)
        assert E1.shape == g + (3,)
        assert s1.iters == s2.iters
        assert np.array_equal(E1, E2) and np.array_equal(H1, H2), f"grid {g} differs in the mix"
    # the 64^3 request's inputs in FP64 against the oracle (fixed budget)
    g, d, e, w = reqs[0]
    med, pol, k = meta[0]
    from frequencymaxwell import engine
    o64 = fm.make_opts(method="born", precision="fp64", tol=0.0, maxiters=10, compute_h=True)
    E, H, st, info = engine.solve_scene(d, e, w, o64)
    bf, bo = _absorbing(fm, oracle)
    s = oracle.Solver(resolution=RES, grid_size=g, boundary_conditions=bo, tolerance=0.0)
    ref = oracle.solve(s, oracle.PlaneWave(wavelength=500e-9, polarization=pol, k_vector=k),
                       med.permittivity, N_BG ** 2, method="born", rule="rhs", maxiters=10)
    assert rel(E, ref.E) <= TOL64 and rel(H, ref.H) <= TOL64
