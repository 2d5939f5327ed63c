"""GPU parity of the on-device Krylov solvers (BiCGStab, GMRES) with the oracle.

* FP64, same algorithm and a fixed budget of Krylov.jl iterations (BiCGStab:
  two operator applications each, GMRES: one Arnoldi step each; the warm
  start's residual outside the count): BiCGStab iterates match to 1e-9; GMRES
  (device CGS2 vs oracle MGS, equal in exact arithmetic) to 1e-9; iteration
  and application counts match exactly;
* converged FP64 solves (tol 1e-10) match the oracle's converged solution
  within the error the residual allows;
* FP32 + FP64 refinement with each Krylov inner solver meets an FP64-verified
  1e-6;
* the reference's default route: solve(...) with GMRES and the Krylov.jl
  stopping rule (reltol on ||r0||, abstol = tol*1e-2, x0 = rhs).
"""

import numpy as np
import pytest

from cases import N_BG, RES, oracle_setup, rel, spec_from

pytestmark = pytest.mark.gpu


def _setup(fm, oracle, grid, bc, seed, precision="fp64"):
    s, eps, init = oracle_setup(oracle, grid, bc, seed)
    src = oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    E_inc, _ = oracle.incident_field(s, src, N_BG ** 2)
    S = oracle.source_term(init.potential, E_inc, init.eps_imag)
    rhs = oracle.green_conv(S, init.green)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision=precision)
    return init, S, rhs, prob


def _bc(oracle, absorb):
    return oracle.AbsorbingBC(thickness=300e-9, attenuation_thickness=300e-9) if absorb else oracle.PeriodicBC()


@pytest.mark.parametrize("absorb", [False, True])
@pytest.mark.parametrize("graph", [False, True])
def test_bicgstab_fixed_budget_fp64(fm, oracle, absorb, graph):
    print(f"case bicgstab fixed absorb={absorb} graph={graph}", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, (16, 12, 8), _bc(oracle, absorb), seed=31)
    budget = 7  # iterations: 2 * 7 + 1 operator applications
    x_o, st_o = oracle.solve_bicgstab(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0, budget)
    E, _, st = prob.solve(S, fm.make_opts(method="bicgstab", tol=0.0, maxiters=budget,
                                          use_graph=graph, check_every=3), want_h=False)
    assert st.iters == st_o.iters == budget
    assert st.applications == st_o.applications == 2 * budget + 1
    assert st.voxel_iters == st.applications * int(np.prod(init.potential.shape))
    assert rel(E, x_o) <= 1e-9
    assert st.resid == pytest.approx(st_o.resid, rel=1e-8)


@pytest.mark.parametrize("absorb", [False, True])
@pytest.mark.parametrize("restart", [0, 5])
def test_gmres_fixed_budget_fp64(fm, oracle, absorb, restart):
    print(f"case gmres fixed absorb={absorb} restart={restart}", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, (16, 12, 8), _bc(oracle, absorb), seed=32)
    budget = 14
    x_o, st_o = oracle.solve_gmres(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0,
                                   budget, restart=restart)
    E, _, st = prob.solve(S, fm.make_opts(method="gmres", tol=0.0, maxiters=budget,
                                          gmres_restart=restart), want_h=False)
    assert st.iters == st_o.iters == budget
    # one application per Arnoldi step and one per cycle start (5 + 5 + 4 steps)
    assert st.applications == st_o.applications == budget + (1 if restart == 0 else 3)
    assert rel(E, x_o) <= 1e-9
    assert st.resid == pytest.approx(st_o.resid, rel=1e-7)


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
def test_krylov_converged_fp64(fm, oracle, method):
    print(f"case {method} converged fp64", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, (32, 32, 16), oracle.PeriodicBC(), seed=33)
    fn = oracle.solve_bicgstab if method == "bicgstab" else oracle.solve_gmres
    budget = 1000 if method == "bicgstab" else 400  # full GMRES: at most 400 basis vectors
    x_o, st_o = fn(rhs, rhs, init.potential, init.eps_imag, init.green, 1e-10, budget)
    assert st_o.converged
    E, _, st = prob.solve(S, fm.make_opts(method=method, tol=1e-10, maxiters=budget, use_graph=True,
                                          check_every=8), want_h=False)
    assert st.converged == 1
    r_true = oracle.true_residual(E, rhs, init.potential, init.eps_imag, init.green)
    assert r_true <= 1.5e-10
    assert rel(E, x_o) <= 1e-7
    assert abs(st.iters - st_o.iters) <= max(4, st_o.iters // 10)


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
def test_krylov_fp32_refine(fm, oracle, method):
    print(f"case {method} fp32_refine", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, (32, 32, 16), oracle.PeriodicBC(), seed=34,
                                precision="fp32")
    E, _, st = prob.solve(S, fm.make_opts(method=method, precision="fp32_refine", tol=1e-6,
                                          maxiters=3000 if method == "bicgstab" else 400,
                                          use_graph=True, check_every=4),
                          want_h=False)
    r_true = oracle.true_residual(E, rhs, init.potential, init.eps_imag, init.green)
    assert st.converged == 1
    assert r_true <= 1.05e-6
    assert st.resid_true_fp64 == pytest.approx(r_true, rel=1e-6)


def test_reference_default_route_gmres(fm, oracle):
    """solve(solver, source, medium) with the reference defaults: GMRES, Krylov.jl
    stopping (reltol on the initial residual, abstol = tol*1e-2), x0 = G S."""
    print("case default route gmres", flush=True)
    grid = (32, 32, 16)
    med = fm.phantom_bead(grid, [1.5 ** 2], 5.0, permittivity_bg=N_BG ** 2)
    src = fm.PlaneWaveSource(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid,
                                     boundary_conditions=fm.PeriodicBoundaryCondition(),
                                     tolerance=1e-9, maxiters=400)
    out = fm.solve(solver, src, med)
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=oracle.PeriodicBC(),
                      tolerance=1e-9)
    ref = oracle.solve(s, oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0],
                                           k_vector=[0, 0, 1]),
                       med.permittivity, N_BG ** 2, method="gmres", rule="krylov", maxiters=400)
    assert solver.last_stats["converged"] == 1 and ref.stats.converged
    assert abs(solver.iteration_count - ref.stats.iters) <= 2
    assert rel(out.E, ref.E) <= 1e-7
    assert rel(out.H, ref.H) <= 1e-7


def test_multi_rhs_equals_single(fm, oracle):
    """Right-hand sides swept together give the single-solve results bit for bit."""
    print("case multi rhs", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, (16, 16, 8), oracle.PeriodicBC(), seed=35)
    rng = np.random.default_rng(0)
    Ss = [S, S * np.exp(1j * 0.3), S * rng.standard_normal(S.shape)]
    for method in ("born", "bicgstab", "gmres"):
        o = fm.make_opts(method=method, tol=1e-8, maxiters=600 if method != "gmres" else 400)
        Es, _, sts = prob.solve_multi(Ss, o, want_h=False)
        for k, Sk in enumerate(Ss):
            E1, _, st1 = prob.solve(Sk, o, want_h=False)
            assert sts[k].iters == st1.iters, method
            assert np.array_equal(Es[k], E1), method


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_krylov_batch_of_grids_many_partial_chunks(fm, oracle, method, precision):
    """Problems of different grids in one Krylov set, each with hundreds of
    x-inverse tiles and vector-kernel blocks: the BiCGStab finalize first sums
    the tiles' partials in 128-row chunks on many CUs, the GMRES dot finalize
    sums each coefficient over the blocks in a block of its own, with a
    different number of chunks and blocks per problem.  Every result equals
    its single solve bit for bit, and in FP64 the largest matches the
    oracle's iterates at a fixed budget."""
    print(f"case {method} batch of grids {precision}", flush=True)
    grids = [(64, 48, 40), (32, 40, 24)]
    budget = 5 if method == "bicgstab" else 9
    probs, Ss, refs = [], [], []
    for g, seed in zip(grids, (41, 42)):
        init, S, rhs, prob = _setup(fm, oracle, g, oracle.PeriodicBC(), seed, precision=precision)
        probs.append(prob)
        Ss.append(S)
        refs.append((init, rhs))
    o = fm.make_opts(method=method, precision=precision, tol=0.0, maxiters=budget,
                     gmres_restart=4 if method == "gmres" else 0)
    Es, _, sts = fm.solve_problems(probs, Ss, o, want_h=False)
    for k, (p, S) in enumerate(zip(probs, Ss)):
        E1, _, st1 = p.solve(S, o, want_h=False)
        assert sts[k].iters == st1.iters == budget
        assert np.array_equal(Es[k], E1), grids[k]
    if precision == "fp64":
        init, rhs = refs[0]
        if method == "bicgstab":
            x_o, st_o = oracle.solve_bicgstab(rhs, rhs, init.potential, init.eps_imag,
                                              init.green, 0.0, budget)
        else:
            x_o, st_o = oracle.solve_gmres(rhs, rhs, init.potential, init.eps_imag, init.green,
                                           0.0, budget, restart=4)
        assert rel(Es[0], x_o) <= 1e-9
        assert sts[0].resid == pytest.approx(st_o.resid, rel=1e-8 if method == "bicgstab" else 1e-7)


def test_gmres_basis_limit(fm, oracle):
    """Full GMRES keeps one basis vector per iteration, at most 400: a larger
    maxiters without a restart length is refused (CBS_EINVAL), as is a restart
    length above 400; GMRES(400) with more iterations runs."""
    print("case gmres basis limit", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, (8, 8, 8), oracle.PeriodicBC(), seed=36)
    for kw in (dict(maxiters=401), dict(maxiters=20, gmres_restart=401)):
        with pytest.raises(ValueError, match="400 basis vectors"):  # CBS_EINVAL
            prob.solve(S, fm.make_opts(method="gmres", tol=1e-6, **kw), want_h=False)
    _, _, st = prob.solve(S, fm.make_opts(method="gmres", tol=1e-6, maxiters=1000,
                                          gmres_restart=400), want_h=False)
    assert st.converged == 1


@pytest.mark.parametrize("grid", [(16, 12, 8), (15, 13, 9)])
def test_gmres_fp32_even_and_odd_voxel_counts(fm, oracle, grid):
    """Complex-FP32 GMRES moves two values per 16-byte access when a grid's voxel
    count is even and one at a time when it is odd: both layouts follow the
    oracle's iterates at a fixed budget over restarted cycles to FP32 accuracy,
    with Krylov.jl's iteration and application counts."""
    nvox = int(np.prod(grid))
    print(f"case gmres fp32 grid {grid} ({'odd' if nvox % 2 else 'even'} voxel count)", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, grid, oracle.PeriodicBC(), seed=37, precision="fp32")
    budget = 11
    x_o, st_o = oracle.solve_gmres(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0,
                                   budget, restart=4)
    E, _, st = prob.solve(S, fm.make_opts(method="gmres", precision="fp32", tol=0.0,
                                          maxiters=budget, gmres_restart=4), want_h=False)
    assert st.iters == st_o.iters == budget
    assert st.applications == st_o.applications == budget + 3  # cycles of 4, 4, 3 steps
    assert rel(E, x_o) <= 2e-5


@pytest.mark.parametrize("budget, restart", [(1, 0), (2, 0), (3, 1), (5, 2)])
def test_gmres_smallest_budgets_and_cycles(fm, oracle, budget, restart):
    """The delayed re-orthogonalisation at its edges: a budget of one or two
    Arnoldi steps (the cycle's only columns are closed by the final reduce)
    and cycles of one or two steps, against the oracle's iterates (FP64)."""
    print(f"case gmres budget {budget} restart {restart}", flush=True)
    init, S, rhs, prob = _setup(fm, oracle, (12, 10, 8), oracle.PeriodicBC(), seed=38)
    x_o, st_o = oracle.solve_gmres(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0,
                                   budget, restart=restart)
    E, _, st = prob.solve(S, fm.make_opts(method="gmres", tol=0.0, maxiters=budget,
                                          gmres_restart=restart), want_h=False)
    cycles = 1 if restart == 0 else -(-budget // restart)
    assert st.iters == st_o.iters == budget
    assert st.applications == st_o.applications == budget + cycles
    assert rel(E, x_o) <= 1e-9
    assert st.resid == pytest.approx(st_o.resid, rel=1e-7)
