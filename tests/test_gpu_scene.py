"""GPU parity of the device-side setup and teardown (cbs_scene_build,
cbs_solve_scene) against the oracle's host restatement of the reference:

* the padded potential V (k_bg^2 (eps/eps_bg - 1), zero padding, eps_imag,
  attenuation windows of every profile and sharpness) and eps_imag / Bornmax;
* the source term S = V E_inc + i eps E_inc of one or several coherent plane
  waves at oblique incidence with amplitude and phase;
* the cropped total fields of a fixed-iteration solve (E and H).
Bars: V and the scalars to 1e-14 relative (same operations, hypot and sincos
differ by an ulp), S to 1e-13, the fields to 1e-10 (as the solve tests).
"""

import numpy as np
import pytest

from cases import N_BG, RES, rel, scene_bcs as _bcs, scene_medium as _medium, scene_sources as _sources

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["periodic", "absorbing", "z_exponential"])
@pytest.mark.parametrize("many", [False, True])
def test_scene_build_matches_oracle(fm, oracle, kind, many):
    print(f"case scene build {kind} many={many}", flush=True)
    from frequencymaxwell import engine
    from frequencymaxwell import solver as S

    grid = (12, 10, 8)
    bf, bo = _bcs(fm, oracle, kind)
    srcf, srco = _sources(fm, oracle, many)
    med = _medium(fm, grid, 3)
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bf)
    d, eps, waves, keep = S.scene(solver, srcf, med)
    V, Sg, info = engine.scene_build(d, eps, waves)
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bo)
    init = oracle.initialize_solver(s, med.permittivity, N_BG ** 2, 500e-9)
    E_inc, _ = oracle.incident_field_sum(s, srco, N_BG ** 2)
    S_o = oracle.source_term(init.potential, E_inc, init.eps_imag)
    assert tuple(info.n) == init.potential.shape
    assert abs(info.eps_imag - init.eps_imag) <= 1e-15 * init.eps_imag
    assert info.bornmax == init.bornmax
    assert info.maxiters == max(2 * init.bornmax, 100)
    assert rel(V, init.potential) <= 1e-14
    assert rel(Sg, S_o) <= 1e-13


@pytest.mark.parametrize("kind", ["periodic", "absorbing"])
def test_solve_scene_fields_match_oracle(fm, oracle, kind):
    """Whole solve on the GPU (two oblique coherent sources): E and H total
    fields, cropped, after a fixed Born budget."""
    print(f"case solve scene {kind}", flush=True)
    grid = (16, 12, 10)
    bf, bo = _bcs(fm, oracle, kind)
    srcf, srco = _sources(fm, oracle, True)
    med = _medium(fm, grid, 4)
    k = 60
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bf,
                                     tolerance=0.0, linear_solver="born", stop_rule="rhs",
                                     maxiters=k)
    with pytest.warns(UserWarning):
        out = fm.solve(solver, srcf, med)
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bo, tolerance=0.0)
    ref = oracle.solve(s, srco, med.permittivity, N_BG ** 2, method="born", rule="rhs",
                       maxiters=k)
    assert solver.iteration_count == ref.stats.iters == k
    assert solver.Bornmax == ref.init.bornmax
    assert out.E.shape == grid + (3,)
    assert rel(out.E, ref.E) <= 1e-10
    assert rel(out.H, ref.H) <= 1e-10


def test_solve_scene_default_maxiters(fm, oracle):
    """Without maxiters the device applies the reference budget
    max(2 Bornmax, 100) (convergent_born.jl:746) from its own eps_imag."""
    print("case solve scene default maxiters", flush=True)
    grid = (16, 16, 12)
    bf, bo = _bcs(fm, oracle, "absorbing")
    srcf, srco = _sources(fm, oracle, False)
    med = fm.phantom_bead(grid, [1.45 ** 2], 3.0, permittivity_bg=N_BG ** 2)
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bf,
                                     tolerance=1e-15, linear_solver="born", stop_rule="rhs")
    with pytest.warns(UserWarning):
        fm.solve(solver, srcf, med)
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bo)
    init = oracle.initialize_solver(s, med.permittivity, N_BG ** 2, 500e-9)
    assert solver.Bornmax == init.bornmax
    assert solver.iteration_count == max(2 * init.bornmax, 100)


def test_solve_scene_fp32_refine(fm, oracle):
    """An FP32+refine BiCGStab solve through the device setup meets its
    FP64-verified tolerance and the converged oracle field."""
    print("case solve scene fp32 refine", flush=True)
    grid = (16, 16, 12)
    bf, bo = _bcs(fm, oracle, "absorbing")
    srcf, srco = _sources(fm, oracle, False)
    med = fm.phantom_bead(grid, [1.45 ** 2], 3.0, permittivity_bg=N_BG ** 2)
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bf,
                                     tolerance=1e-6, linear_solver="bicgstab",
                                     precision="fp32_refine", stop_rule="rhs", maxiters=3000)
    out = fm.solve(solver, srcf, med)
    assert solver.last_stats["converged"] == 1
    assert solver.last_stats["resid_true_fp64"] <= 1e-6
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bo, tolerance=1e-9)
    ref = oracle.solve(s, srco, med.permittivity, N_BG ** 2, method="gmres", rule="rhs",
                       maxiters=2000)
    assert ref.stats.converged
    assert rel(out.E, ref.E) <= 1e-5


def test_solve_batch_device_equals_single(fm, oracle):
    """solve_batch on the device (cbs_solve_scene_multi): every right-hand side
    (two single plane waves and one coherent pair) equals its own solve bit for
    bit, and the oracle's whole solve to 1e-10."""
    print("case solve batch device", flush=True)
    grid = (16, 12, 10)
    bf, bo = _bcs(fm, oracle, "absorbing")
    srcf, srco = _sources(fm, oracle, True)
    single = fm.PlaneWaveSource(wavelength=500e-9, polarization=[0, 1, 0], k_vector=[0, 0, 1])
    med = _medium(fm, grid, 9)
    kw = dict(resolution=RES, grid_size=grid, boundary_conditions=bf, tolerance=0.0,
              linear_solver="born", stop_rule="rhs", maxiters=25)
    batch = [srcf[0], single, srcf]
    solver = fm.ConvergentBornSolver(**kw)
    with pytest.warns(UserWarning):
        outs = fm.solve_batch(solver, batch, med)
    for src, out in zip(batch, outs):
        one = fm.ConvergentBornSolver(**kw)
        with pytest.warns(UserWarning):
            ref = fm.solve(one, src, med)
        assert np.array_equal(out.E, ref.E) and np.array_equal(out.H, ref.H)
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bo, tolerance=0.0)
    ref = oracle.solve(s, srco, med.permittivity, N_BG ** 2, method="born", rule="rhs", maxiters=25)
    assert rel(outs[2].E, ref.E) <= 1e-10 and rel(outs[2].H, ref.H) <= 1e-10
