"""GPU field analysis on resident fields (cbs_field_*, cbs_solve_scene_field)
against the oracle's restatement of src/fields/electromagnetic_field.jl:112-235:
intensity and planes bit for bit (same operations), Poynting vector to
1e-14 (fused complex products), energy to 1e-12 (summation order); a scene
solve left on the device equals the downloaded one bit for bit.
"""

import numpy as np
import pytest

from cases import RES, rel, scene_bcs, scene_medium, scene_sources

pytestmark = pytest.mark.gpu


def _random_field(fm, shape, seed, with_h=True):
    rng = np.random.default_rng(seed)
    E = np.asfortranarray(rng.standard_normal(shape + (3,)) + 1j * rng.standard_normal(shape + (3,)))
    H = np.asfortranarray(rng.standard_normal(shape + (3,)) + 1j * rng.standard_normal(shape + (3,)))
    return fm.ElectromagneticField(E, H if with_h else np.zeros_like(E), (30e-9, 40e-9, 50e-9), 532e-9)


@pytest.mark.parametrize("shape", [(7, 5, 3), (32, 24, 16)])
def test_device_field_analysis_matches_oracle(fm, oracle, shape):
    print(f"case device field analysis {shape}", flush=True)
    f = _random_field(fm, shape, 5)
    d = fm.DeviceField.from_fields(f)
    try:
        e_ref = oracle.field_energy(f.E, f.H, f.resolution)
        assert abs(d.energy() - e_ref) <= 1e-12 * e_ref
        for comp in ("total", "x", "y", "z"):
            assert np.array_equal(d.intensity(comp), oracle.field_intensity(f.E, comp)), comp
        assert rel(d.poynting(), oracle.poynting_vector(f.E, f.H)) <= 1e-14
        for axis in (1, 2, 3):
            for idx in (1, shape[axis - 1]):
                p_d = d.extract_plane(axis, idx)
                assert np.array_equal(p_d.E, oracle.extract_plane(f.E, axis, idx))
                assert np.array_equal(p_d.H, oracle.extract_plane(f.H, axis, idx))
        back = d.fields()
        assert np.array_equal(back.E, f.E) and np.array_equal(back.H, f.H)
        with pytest.raises(ValueError):
            d.extract_plane(2, shape[1] + 1)
    finally:
        d.close()


def test_device_field_without_h(fm, oracle):
    f = _random_field(fm, (6, 6, 4), 6)
    E_only = fm.ElectromagneticField(f.E, f.H, f.resolution, f.wavelength)
    E_only.H = None
    d = fm.DeviceField.from_fields(E_only)
    try:
        assert np.array_equal(d.intensity(), oracle.field_intensity(f.E))
        with pytest.raises(ValueError):
            d.poynting()
    finally:
        d.close()


def test_solve_scene_field_resident(fm, oracle):
    """A scene solved with its totals kept on the device: the same fields as
    cbs_solve_scene, and the analysis of them equals the oracle's."""
    print("case solve scene field", flush=True)
    from frequencymaxwell import engine
    from frequencymaxwell import solver as S

    grid = (16, 12, 10)
    bf, _ = scene_bcs(fm, oracle, "absorbing")
    srcf, _ = scene_sources(fm, oracle, True)
    med = scene_medium(fm, grid, 7)
    solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bf)
    d, eps, waves, keep = S.scene(solver, srcf, med)
    o = fm.make_opts(method="born", tol=0.0, maxiters=30)
    E, H, st, info = engine.solve_scene(d, eps, waves, o)
    fd, st2, info2 = engine.solve_scene_field(d, eps, waves, o)
    try:
        assert st2.iters == st.iters == 30
        back = fd.fields()
        assert np.array_equal(back.E, E) and np.array_equal(back.H, H)
        e_ref = oracle.field_energy(E, H, RES)
        assert abs(fd.energy() - e_ref) <= 1e-12 * e_ref
        assert np.array_equal(fd.intensity(), oracle.field_intensity(E))
        assert np.array_equal(fd.extract_plane(3, 5).E, oracle.extract_plane(E, 3, 5))
    finally:
        fd.close()
