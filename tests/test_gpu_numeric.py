"""Non-finite values end a solve with CBS_ENUMERIC, forced from the host.

* Inputs: NaN in the potential, the source term or the permittivity is
  refused before anything runs on the GPU.
* Iterates: a problem description whose preconditioner scale eps is far below
  max|V| (the reference sets eps >= 1.01 max|V0|, convergent_born.jl:641)
  makes the Born iteration diverge: the residual grows by about max|V|/eps per
  sweep until its FP64 squared norm overflows; the device finalize flags it
  and the call returns CBS_ENUMERIC instead of running to maxiters.  Every
  kernel still ends normally (the diverging values are ordinary floats, then
  Inf and NaN).
"""

import math

import numpy as np
import pytest

from cases import N_BG, RES, oracle_setup, random_field, spec_from

pytestmark = pytest.mark.gpu


def _diverging(fm, oracle, precision):
    s, eps, init = oracle_setup(oracle, (16, 16, 8), oracle.PeriodicBC(), 41)
    spec = spec_from(fm, init)
    # eps 1000x below its floor: gamma = i V / eps makes Born sweeps grow ~1000x
    spec = fm.GreenSpec(shape=spec.shape, resolution=spec.resolution, shifts=spec.shifts,
                        use_averaging=spec.use_averaging, k_square=spec.k_square,
                        eps_imag=init.eps_imag * 1e-3)
    prob = fm.CBSProblem(init.potential, spec, precision=precision)
    S = np.asfortranarray(random_field(init.potential.shape, 42) * 1e14)
    return prob, S


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32_refine"])
def test_diverging_born_raises_enumeric(fm, oracle, precision):
    print(f"case diverging born {precision}", flush=True)
    prob, S = _diverging(fm, oracle, "fp64" if precision == "fp64" else "fp32")
    with pytest.raises(fm.CBSError) as ei:
        prob.solve(S, fm.make_opts(method="born", precision=precision, tol=1e-6, maxiters=3000),
                   want_h=False)
    assert ei.value.code == fm._lib.CBS_ENUMERIC
    assert "NaN or Inf" in str(ei.value)
    prob.close()


def test_diverging_born_served_raises_enumeric(fm, oracle):
    """A served request whose residual overflows ends CBS_ENUMERIC; the server
    keeps serving the other requests."""
    print("case diverging born served", flush=True)
    s, eps, init = oracle_setup(oracle, (16, 16, 8), oracle.PeriodicBC(), 43)
    spec = spec_from(fm, init)
    bad = fm.GreenSpec(shape=spec.shape, resolution=spec.resolution, shifts=spec.shifts,
                       use_averaging=spec.use_averaging, k_square=spec.k_square,
                       eps_imag=init.eps_imag * 1e-3)
    S = np.asfortranarray(random_field(init.potential.shape, 44) * 1e14)
    srv = fm.CBSServer(max_active=4)
    try:
        o = fm.make_opts(method="born", precision="fp64", tol=1e-6, maxiters=3000)
        t_bad = srv.submit(init.potential, bad, S, o)
        t_ok = srv.submit(init.potential, spec, S, fm.make_opts(method="born", tol=0.0,
                                                                 maxiters=10))
        with pytest.raises(fm.CBSError) as ei:
            srv.wait(t_bad)
        assert ei.value.code == fm._lib.CBS_ENUMERIC
        _, _, st = srv.wait(t_ok)
        assert st.iters == 10
    finally:
        srv.stop()


def test_nan_inputs_refused(fm, oracle):
    print("case nan inputs", flush=True)
    s, eps, init = oracle_setup(oracle, (8, 8, 8), oracle.PeriodicBC(), 45)
    spec = spec_from(fm, init)
    V = init.potential.copy(order="F")
    V[1, 2, 3] = complex(math.nan, 0.0)
    with pytest.raises(fm.CBSError) as ei:
        fm.CBSProblem(V, spec, precision="fp64")
    assert ei.value.code == fm._lib.CBS_ENUMERIC
    prob = fm.CBSProblem(init.potential, spec, precision="fp64")
    S = random_field(init.potential.shape, 46)
    S[0, 0, 0, 1] = complex(0.0, math.inf)
    with pytest.raises(fm.CBSError) as ei:
        prob.solve(S, fm.make_opts(method="born", tol=1e-6, maxiters=10), want_h=False)
    assert ei.value.code == fm._lib.CBS_ENUMERIC
    prob.close()
    # permittivity of a scene (device-side setup)
    med = fm.phantom_bead((8, 8, 8), [1.5 ** 2], 2.0, permittivity_bg=N_BG ** 2)
    med.permittivity[4, 4, 4] = complex(math.nan, 0.0)
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=(8, 8, 8),
                                     boundary_conditions=fm.PeriodicBoundaryCondition(),
                                     tolerance=1e-6, linear_solver="born")
    src = fm.PlaneWaveSource(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    with pytest.raises(fm.CBSError) as ei:
        fm.solve(solver, src, med)
    assert ei.value.code == fm._lib.CBS_ENUMERIC
