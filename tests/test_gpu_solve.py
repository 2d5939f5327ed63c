"""GPU parity of the iterative solve against the FP64 oracle.

* same algorithm, same fixed iteration count: relative field error <= 1e-10
  (FP64 mode), including the reference's whole `solve` path (E and H);
* a converged FP64 solve (tol 1e-11) stops within one sweep of the oracle
  and matches it to 1e-10;
* FP32 and FP32+FP64-refinement: FP64-verified residual meets the tolerance
  and the field error stays within the bound the residual implies.
"""

import numpy as np
import pytest

from cases import N_BG, RES, oracle_setup, random_field, rel, spec_from

pytestmark = pytest.mark.gpu


def _problem_and_rhs(fm, oracle, grid, bc, seed, precision="fp64"):
    s, eps, init = oracle_setup(oracle, grid, bc, seed)
    src = oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    E_inc, _ = oracle.incident_field(s, src, N_BG ** 2)
    S = oracle.source_term(init.potential, E_inc, init.eps_imag)
    rhs = oracle.green_conv(S, init.green)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision=precision)
    return init, S, rhs, prob


@pytest.mark.parametrize("absorb", [False, True])
@pytest.mark.parametrize("graph", [False, True])
def test_born_fixed_iterations_fp64(fm, oracle, absorb, graph):
    print(f"case born fixed iters absorb={absorb} graph={graph}", flush=True)
    bc = oracle.AbsorbingBC(thickness=400e-9, attenuation_thickness=300e-9) if absorb else oracle.PeriodicBC()
    init, S, rhs, prob = _problem_and_rhs(fm, oracle, (16, 16, 12), bc, seed=21)
    k = 9
    x_o, st_o = oracle.solve_born(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0, k)
    opts = fm.make_opts(method="born", precision="fp64", tol=0.0, maxiters=k, use_graph=graph,
                        check_every=4 if graph else 1)
    E, H, st = prob.solve(S, opts, want_h=False)
    assert st.iters == k == st_o.iters
    assert st.converged == 0
    assert rel(E, x_o) <= 1e-10
    assert abs(st.resid - st_o.resid) <= 1e-10 * st_o.resid


@pytest.mark.parametrize("grid", [(16, 12, 10), (12, 16, 10)])  # x: specialised, runtime plan
@pytest.mark.parametrize("absorb", [False, True])
def test_born_fused_sweeps_same_as_separate(fm, oracle, grid, absorb):
    """Fusing each sweep's x inverse with the next sweep's x forward changes
    no bit of the iterates, whatever the chunking; FP64 matches the oracle."""
    print(f"case born fused sweeps grid={grid} absorb={absorb}", flush=True)
    bc = oracle.AbsorbingBC(thickness=300e-9, attenuation_thickness=200e-9) if absorb else oracle.PeriodicBC()
    init, S, rhs, prob = _problem_and_rhs(fm, oracle, grid, bc, seed=23)
    k = 11
    x_o, _ = oracle.solve_born(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0, k)
    for precision in ("fp64", "fp32"):
        p = prob if precision == "fp64" else fm.CBSProblem(init.potential, spec_from(fm, init),
                                                           precision="fp32")
        res = []
        for fuse, every, graph in ((False, 1, False), (True, 1, False), (True, 4, True),
                                   (True, 3, False)):
            E, _, st = p.solve(S, fm.make_opts(method="born", precision=precision, tol=0.0,
                                               maxiters=k, check_every=every, use_graph=graph,
                                               fuse_sweeps=fuse), want_h=False)
            assert st.iters == k
            res.append(E)
        for i, E in enumerate(res[1:], 1):
            assert np.array_equal(E, res[0]), f"{precision} variant {i}: rel {rel(E, res[0]):.3e}"
        if precision == "fp64":
            assert rel(res[0], x_o) <= 1e-10
        else:
            assert rel(res[0], x_o) <= 1e-5
            p.close()


def test_born_converged_fp64_rhs_rule(fm, oracle):
    print("case born converged fp64", flush=True)
    init, S, rhs, prob = _problem_and_rhs(fm, oracle, (16, 12, 8), oracle.PeriodicBC(), seed=22)
    x_o, st_o = oracle.solve_born(rhs, rhs, init.potential, init.eps_imag, init.green, 1e-11, 5000)
    assert st_o.converged
    E, _, st = prob.solve(S, fm.make_opts(method="born", tol=1e-11, maxiters=5000), want_h=False)
    assert st.converged == 1
    assert abs(st.iters - st_o.iters) <= 1
    assert rel(E, x_o) <= 1e-10
    assert oracle.true_residual(E, rhs, init.potential, init.eps_imag, init.green) <= 1.01e-11


@pytest.mark.parametrize("device_setup", [True, False])
def test_solve_api_bead_fp64(fm, oracle, device_setup):
    """The reference's solve(solver, source, medium) path end to end: potential,
    incident field, source term, G S, 300 Born sweeps, curl, crop -- with the
    setup and teardown on the GPU (cbs_solve_scene) and on the host."""
    print(f"case solve api bead fp64 device_setup={device_setup}", flush=True)
    grid = (32, 32, 16)
    med = fm.phantom_bead(grid, [1.5 ** 2], 5.0, permittivity_bg=N_BG ** 2)
    src = fm.PlaneWaveSource(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid,
                                     boundary_conditions=fm.PeriodicBoundaryCondition(),
                                     tolerance=0.0, linear_solver="born", stop_rule="rhs",
                                     maxiters=300, device_setup=device_setup)
    with pytest.warns(UserWarning):
        out = fm.solve(solver, src, med)
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=oracle.PeriodicBC(),
                      tolerance=0.0)
    ref = oracle.solve(s, oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0],
                                           k_vector=[0, 0, 1]),
                       med.permittivity, N_BG ** 2, method="born", rule="rhs", maxiters=300)
    assert solver.iteration_count == ref.stats.iters == 300
    assert rel(out.E, ref.E) <= 1e-10
    assert rel(out.H, ref.H) <= 1e-10


@pytest.mark.parametrize("device_setup", [True, False])
def test_solve_api_absorbing_fp64(fm, oracle, device_setup):
    print(f"case solve api absorbing fp64 device_setup={device_setup}", flush=True)
    grid = (16, 16, 12)
    med = fm.phantom_bead(grid, [1.45 ** 2], 3.0, permittivity_bg=N_BG ** 2)
    bc = fm.AbsorbingBoundaryCondition(thickness=400e-9, attenuation_thickness=300e-9)
    src = fm.PlaneWaveSource(wavelength=500e-9, polarization=[0, 1, 0], k_vector=[0, 0, 1])
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bc,
                                     tolerance=0.0, linear_solver="born", stop_rule="rhs",
                                     maxiters=300, device_setup=device_setup)
    with pytest.warns(UserWarning):
        out = fm.solve(solver, src, med)
    s = oracle.Solver(resolution=RES, grid_size=grid,
                      boundary_conditions=oracle.AbsorbingBC(thickness=400e-9, attenuation_thickness=300e-9),
                      tolerance=0.0)
    ref = oracle.solve(s, oracle.PlaneWave(wavelength=500e-9, polarization=[0, 1, 0],
                                           k_vector=[0, 0, 1]),
                       med.permittivity, N_BG ** 2, method="born", rule="rhs", maxiters=300)
    assert solver.iteration_count == ref.stats.iters == 300
    assert rel(out.E, ref.E) <= 1e-10
    assert rel(out.H, ref.H) <= 1e-10


@pytest.mark.parametrize("precision", ["fp32", "fp32_refine"])
def test_born_mixed_precision(fm, oracle, precision):
    print(f"case born {precision}", flush=True)
    init, S, rhs, prob = _problem_and_rhs(fm, oracle, (32, 32, 16), oracle.PeriodicBC(), seed=23,
                                          precision="fp32")
    x_o, st_o = oracle.solve_born(rhs, rhs, init.potential, init.eps_imag, init.green, 1e-12, 8000)
    tol = 1e-6 if precision == "fp32_refine" else 1e-4
    E, _, st = prob.solve(S, fm.make_opts(method="born", precision=precision, tol=tol,
                                          maxiters=8000), want_h=False)
    r_true = oracle.true_residual(E, rhs, init.potential, init.eps_imag, init.green)
    assert r_true <= tol * 1.05
    assert st.converged == 1
    if precision == "fp32_refine":
        assert abs(st.resid_true_fp64 - r_true) <= 1e-3 * r_true + 1e-14
        assert st.iters_fp64 >= 2
    assert rel(E, x_o) <= 50 * tol


@pytest.mark.parametrize("method", ["born", "bicgstab"])
def test_fp32_refine_tolerance_zero_runs_its_budget(fm, oracle, method):
    """Tolerance 0 with FP32+refine is a fixed budget: the correction solves
    run at the FP32 floor until maxiters (inner iterations plus FP64 residual
    evaluations) ends the solve, and the FP64 residual keeps falling."""
    print(f"case fp32_refine tol 0 {method}", flush=True)
    init, S, rhs, prob = _problem_and_rhs(fm, oracle, (16, 16, 12), oracle.PeriodicBC(), seed=24,
                                          precision="fp32")
    budget = 60
    E, _, st = prob.solve(S, fm.make_opts(method=method, precision="fp32_refine", tol=0.0,
                                          maxiters=budget), want_h=False)
    assert st.converged == 0
    # an inner solve that stops at the floor one short of the budget leaves no
    # room for another correction and the final FP64 evaluation
    assert budget - 1 <= st.iters <= budget
    assert st.refine_steps >= 1 and st.iters_fp64 == st.refine_steps + 1
    per = 2 if method == "bicgstab" else 1
    assert st.applications == st.iters_fp64 + per * (st.iters - st.iters_fp64)
    r_true = oracle.true_residual(E, rhs, init.potential, init.eps_imag, init.green)
    assert st.resid_true_fp64 == pytest.approx(r_true, rel=1e-6)
    # 60 iterations of a contraction reduce the residual well below the start
    x1, st1 = oracle.solve_born(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0, 1)
    assert r_true < 0.25 * st1.resid
