"""GPU tests of the stepwise solver sets behind cbs_solve_multi / cbs_solve.

A Born set launches the next chunk of sweeps before it checks the previous
one (two chunks in flight); members that converge leave the set between
chunks while the chunk launched ahead still runs (their sweeps are skipped on
the device).  Whatever the chunk length and launch mode, every right-hand side
must come out bit for bit as it does alone with one sweep per check.
"""

import numpy as np
import pytest

from cases import N_BG, RES, oracle_setup, rel, spec_from

pytestmark = pytest.mark.gpu


def _problem(fm, oracle, grid, seed, precision):
    s, eps, init = oracle_setup(oracle, grid, oracle.PeriodicBC(), seed)
    src = oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    E_inc, _ = oracle.incident_field(s, src, N_BG ** 2)
    S = oracle.source_term(init.potential, E_inc, init.eps_imag)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision=precision)
    return init, S, prob


@pytest.mark.parametrize("every,graph", [(2, False), (8, False), (8, True), (6, True)])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_born_set_members_leave_at_different_chunks(fm, oracle, every, graph, precision):
    print(f"case born set every={every} graph={graph} {precision}", flush=True)
    init, S, prob = _problem(fm, oracle, (16, 16, 8), 37,
                             "fp64" if precision == "fp64" else "fp32")
    rng = np.random.default_rng(7)
    # sources of different spectral content converge after different sweep counts
    Ss = [S, S * rng.standard_normal(S.shape), 1e-3 * S + S.mean() * np.exp(0.4j),
          np.zeros_like(S)]
    # Born on this grid stagnates near 1e-7: a reachable tolerance for both precisions
    o = fm.make_opts(method="born", precision=precision, tol=2e-6, maxiters=3000,
                     check_every=every, use_graph=graph)
    Es, _, sts = prob.solve_multi(Ss, o, want_h=False)
    ref = fm.make_opts(method="born", precision=precision, tol=2e-6, maxiters=3000)
    iters = []
    for k, Sk in enumerate(Ss):
        E1, _, st1 = prob.solve(Sk, ref, want_h=False)
        print(f"  rhs {k}: iters {sts[k].iters} resid {sts[k].resid:.2e}", flush=True)
        assert sts[k].iters == st1.iters
        assert sts[k].converged == st1.converged == 1
        assert np.array_equal(Es[k], E1), f"rhs {k}: rel {rel(Es[k], E1):.3e}"
        iters.append(st1.iters)
    # the zero source is converged at once (x = 0); the others leave at different chunks
    assert np.all(Es[3] == 0)
    assert len(set(iters[:3])) == 3
    prob.close()
