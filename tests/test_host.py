"""CPU tests of the product's host side and of the C-ABI boundary (no GPU calls).

* the library loads and exports every entry point include/cbs.h declares;
* the ctypes structures match the C layout (offsets printed by a gcc-built probe);
* the Python mirror of the reference's host code (boundary conditions, potential
  build, incident field, phantom) agrees with the oracle and with the
  reference's own unit-test values;
* argument validation raises ValueError (Julia ArgumentError) where the
  reference throws;
* without a GPU the product fails loudly (no CPU fallback).
"""

import ctypes
import math
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cases import N_BG, RES, rel
from conftest import ROOT, gpu_available

HEADER = ROOT / "include" / "cbs.h"


def _declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:cbs_status|const char \*|void)\s*(cbs_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(fm):
    lib = ctypes.CDLL(str(fm._lib.LIB_PATH))
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(fm._lib.EXPORTED)


_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "cbs.h"
#define F(S, f) printf(#S "." #f " %zu\n", offsetof(S, f));
int main(void) {
  printf("cbs_problem_desc %zu\ncbs_solve_opts %zu\ncbs_solve_stats %zu\ncbs_request %zu\n",
         sizeof(cbs_problem_desc), sizeof(cbs_solve_opts), sizeof(cbs_solve_stats), sizeof(cbs_request));
  F(cbs_problem_desc, n) F(cbs_problem_desc, res) F(cbs_problem_desc, shift) F(cbs_problem_desc, averaging)
  F(cbs_problem_desc, k2_re) F(cbs_problem_desc, k2_im) F(cbs_problem_desc, eps_imag) F(cbs_problem_desc, precision)
  F(cbs_solve_opts, method) F(cbs_solve_opts, precision) F(cbs_solve_opts, stop_rule) F(cbs_solve_opts, maxiters)
  F(cbs_solve_opts, tol) F(cbs_solve_opts, abstol) F(cbs_solve_opts, gmres_restart) F(cbs_solve_opts, max_refine)
  F(cbs_solve_opts, inner_tol) F(cbs_solve_opts, check_every) F(cbs_solve_opts, use_graph) F(cbs_solve_opts, x0_zero)
  F(cbs_solve_opts, compute_h) F(cbs_solve_opts, h_scale_re) F(cbs_solve_opts, h_scale_im)
  F(cbs_solve_stats, iters) F(cbs_solve_stats, iters_fp64) F(cbs_solve_stats, refine_steps) F(cbs_solve_stats, converged)
  F(cbs_solve_stats, resid) F(cbs_solve_stats, resid_true_fp64) F(cbs_solve_stats, time_ms) F(cbs_solve_stats, bytes_moved)
  F(cbs_solve_stats, voxel_iters) F(cbs_solve_stats, applications)
  F(cbs_request, desc) F(cbs_request, V) F(cbs_request, S) F(cbs_request, opts)
  printf("cbs_scene_desc %zu\ncbs_plane_wave %zu\ncbs_scene_info %zu\n", sizeof(cbs_scene_desc),
         sizeof(cbs_plane_wave), sizeof(cbs_scene_info));
  F(cbs_scene_desc, grid) F(cbs_scene_desc, res) F(cbs_scene_desc, pad) F(cbs_scene_desc, att_len)
  F(cbs_scene_desc, window) F(cbs_scene_desc, shift) F(cbs_scene_desc, averaging) F(cbs_scene_desc, eps_bg)
  F(cbs_scene_desc, wavelength) F(cbs_scene_desc, iterations_max) F(cbs_scene_desc, precision)
  F(cbs_plane_wave, k_vec) F(cbs_plane_wave, e_perp) F(cbs_plane_wave, h_perp) F(cbs_plane_wave, amplitude)
  F(cbs_plane_wave, phase)
  F(cbs_scene_info, n) F(cbs_scene_info, eps_imag) F(cbs_scene_info, k2_re) F(cbs_scene_info, k2_im)
  F(cbs_scene_info, bornmax) F(cbs_scene_info, maxiters)
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_ctypes_layout_matches_c(fm, tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    L = fm._lib
    structs = {"cbs_problem_desc": L.ProblemDesc, "cbs_solve_opts": L.SolveOpts,
               "cbs_solve_stats": L.SolveStats, "cbs_request": L.Request,
               "cbs_scene_desc": L.SceneDesc, "cbs_plane_wave": L.PlaneWave,
               "cbs_scene_info": L.SceneInfo}
    for line in filter(None, out):
        key, val = line.split()
        if "." in key:
            s, f = key.split(".")
            assert getattr(structs[s], f).offset == int(val), key
        else:
            assert ctypes.sizeof(structs[key]) == int(val), key


# ---------------------------------------------------------------------------
# host mirror of the reference (src/solvers, src/sources, src/medium)
# ---------------------------------------------------------------------------
def test_boundary_conditions_reference_values(fm):
    """test/solvers/boundary_conditions.jl:80-136."""
    ab = fm.AbsorbingBoundaryCondition(thickness=2.0e-6, attenuation_thickness=1.5e-6)
    per = fm.PeriodicBoundaryCondition()
    assert fm.padding_pixels(per, 50e-9) == 0 and fm.padding_pixels(per, 100e-9) == 0
    assert fm.padding_pixels(ab, 50e-9) == 20 and fm.padding_pixels(ab, 100e-9) == 10
    assert fm.subpixel_shift(per) == 0.0 and fm.subpixel_shift(ab) == 0.25
    assert not fm.requires_padding(per) and fm.requires_padding(ab)
    assert not fm.requires_averaging(per) and fm.requires_averaging(ab)
    assert ab.sharpness == 0.9 and ab.profile is fm.TanhProfile
    for kw in (dict(thickness=-1e-6, attenuation_thickness=1e-6),
               dict(thickness=1e-6, attenuation_thickness=0.0),
               dict(thickness=1e-6, attenuation_thickness=2e-6),
               dict(thickness=2e-6, attenuation_thickness=1e-6, sharpness=1.5)):
        with pytest.raises(ValueError):
            fm.AbsorbingBoundaryCondition(**kw)


@pytest.mark.parametrize("bcname", ["periodic", "absorbing", "mixed"])
def test_potential_build_matches_oracle(fm, oracle, bcname):
    from frequencymaxwell import solver as S
    grid = (12, 10, 8)
    rng = np.random.default_rng(2)
    eps = (N_BG ** 2 + 0.4 * rng.random(grid)).astype(complex)
    if bcname == "periodic":
        bc_p, bc_o = fm.PeriodicBoundaryCondition(), oracle.PeriodicBC()
    elif bcname == "absorbing":
        bc_p = fm.AbsorbingBoundaryCondition(thickness=600e-9, attenuation_thickness=400e-9, sharpness=0.7,
                                             profile=fm.TukeyProfile)
        bc_o = oracle.AbsorbingBC(thickness=600e-9, attenuation_thickness=400e-9, sharpness=0.7,
                                  profile="tukey")
    else:
        a_p = fm.AbsorbingBoundaryCondition(thickness=500e-9, attenuation_thickness=500e-9)
        bc_p = (fm.PeriodicBoundaryCondition(), a_p, fm.PeriodicBoundaryCondition())
        bc_o = (oracle.PeriodicBC(), oracle.AbsorbingBC(thickness=500e-9, attenuation_thickness=500e-9),
                oracle.PeriodicBC())
    sol = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bc_p)
    med = fm.Medium(eps, N_BG ** 2)
    pot, eps_imag, bornmax, k_bg, spec = S._initialize_solver(sol, med, 532e-9)
    so = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bc_o)
    init = oracle.initialize_solver(so, eps, N_BG ** 2, 532e-9)
    assert pot.shape == init.potential.shape
    assert rel(pot, init.potential) < 1e-15
    assert eps_imag == pytest.approx(init.eps_imag, rel=1e-15)
    assert bornmax == init.bornmax
    assert spec.shifts == init.green.shifts and spec.use_averaging == init.green.use_averaging
    assert complex(spec.k_square) == pytest.approx(init.green.k_square, rel=1e-15)
    assert sol.compute_ROI() == oracle.compute_roi(so)


def test_incident_field_matches_oracle(fm, oracle):
    bc_p = fm.AbsorbingBoundaryCondition(thickness=300e-9, attenuation_thickness=200e-9)
    bc_o = oracle.AbsorbingBC(thickness=300e-9, attenuation_thickness=200e-9)
    grid = (10, 9, 8)
    sol = fm.ConvergentBornSolver(resolution=(50e-9, 40e-9, 60e-9), grid_size=grid, boundary_conditions=bc_p)
    so = oracle.Solver(resolution=(50e-9, 40e-9, 60e-9), grid_size=grid, boundary_conditions=bc_o)
    kv = np.array([0.2, 0.1, np.sqrt(1 - 0.05)])
    p1 = np.cross(kv, [0, 1.0, 0])
    p1 /= np.linalg.norm(p1)
    srcs_p = [fm.PlaneWaveSource(wavelength=532e-9, polarization=p1, k_vector=kv, amplitude=1.5, phase=0.2),
              fm.PlaneWaveSource(wavelength=532e-9, polarization=[1, 1j, 0], k_vector=[0, 0, 1])]
    srcs_o = [oracle.PlaneWave(wavelength=532e-9, polarization=p1, k_vector=kv, amplitude=1.5, phase=0.2),
              oracle.PlaneWave(wavelength=532e-9, polarization=[1, 1j, 0], k_vector=[0, 0, 1])]
    f = fm.generate_incident_field(srcs_p, sol, permittivity_bg=1.4 ** 2)
    E, H = oracle.incident_field_sum(so, srcs_o, 1.4 ** 2)
    assert rel(f.E, E) < 1e-14 and rel(f.H, H) < 1e-14
    with pytest.raises(ValueError):
        fm.generate_incident_field([], sol)
    with pytest.raises(ValueError):
        fm.PlaneWaveSource(wavelength=500e-9, polarization=[0, 0, 1], k_vector=[0, 0, 1])
    with pytest.raises(ValueError):
        fm.PlaneWaveSource(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 2])


def test_plane_wave_reference_tests(fm):
    """test/sources/plane_wave.jl: sizes, finiteness, impedance, orthogonality, superposition."""
    sol = fm.ConvergentBornSolver(resolution=(50e-9,) * 3, grid_size=(32, 32, 32),
                                  boundary_conditions=fm.PeriodicBoundaryCondition())
    f = fm.generate_incident_field(fm.PlaneWaveSource(wavelength=633e-9, polarization=[1, 0, 0],
                                                      k_vector=[0, 0, 1]), sol)
    assert f.E.shape == (32, 32, 32, 3) and np.isfinite(f.E).all() and np.isfinite(f.H).all()
    c = (16, 16, 16)
    assert abs(f.E[c + (0,)]) / abs(f.H[c + (1,)]) == pytest.approx(376.730313668, rel=0.1)
    s2 = fm.ConvergentBornSolver(resolution=(100e-9,) * 3, grid_size=(32, 32, 16),
                                 boundary_conditions=fm.PeriodicBoundaryCondition())
    ex = fm.generate_incident_field(fm.PlaneWaveSource(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1]), s2)
    ey = fm.generate_incident_field(fm.PlaneWaveSource(wavelength=500e-9, polarization=[0, 1, 0], k_vector=[0, 0, 1]), s2)
    assert np.sum(np.abs(ex.E * np.conj(ey.E))) == pytest.approx(0.0, abs=1e-12)
    both = ex + ey
    assert np.allclose(both.E, ex.E + ey.E) and np.allclose(both.H, ex.H + ey.H)


def test_phantom_bead_reference_tests(fm):
    """test/geometry/phantoms.jl:35-42: centre value and sphere volume within 30 %."""
    g = (64, 64, 32)
    med = fm.phantom_bead(g, [1.5 ** 2], 8.0, permittivity_bg=1.33 ** 2)
    eps = med.permittivity
    c = tuple(int(round((s + 1) / 2)) - 1 for s in g)
    assert eps[c] == pytest.approx(1.5 ** 2)
    vol = np.sum(np.abs(eps - 1.5 ** 2) < 1e-12)
    assert vol == pytest.approx(4 / 3 * np.pi * 8.0 ** 3, rel=0.3)
    assert fm.grid_size(med) == g and fm.permittivity_bg(med) == pytest.approx(1.33 ** 2)
    with pytest.raises(ValueError):
        fm.Medium(np.ones(g, complex), -1.0)


def test_solver_validation(fm):
    kw = dict(resolution=RES, grid_size=(8, 8, 8), boundary_conditions=fm.PeriodicBoundaryCondition())
    s = fm.ConvergentBornSolver(**kw)
    assert s.tolerance == pytest.approx(1e-6) and s.iterations_max == -1 and s.Bornmax == 0
    assert s.domain_size() == pytest.approx((400e-9,) * 3)
    for bad in (dict(resolution=(0, 1e-9, 1e-9)), dict(grid_size=(0, 8, 8)), dict(tolerance=2.0),
                dict(device="cpu"), dict(linear_solver="cg")):
        with pytest.raises(ValueError):
            fm.ConvergentBornSolver(**{**kw, **bad})
    s.iteration_count = 5
    s.residual_history.append(1.0)
    fm.reset(s)
    assert s.iteration_count == 0 and s.residual_history == []


def test_request_footprint_without_gpu(fm):
    """Admission arithmetic is host-only: it grows with the grid, FP64 > FP32."""
    L = fm._lib
    lib = L.load()
    d = L.ProblemDesc()
    for k in range(3):
        d.n[k] = 64
        d.res[k] = 50e-9
    d.k2_re, d.k2_im, d.eps_imag = 2.79e14, 7.7e13, 7.7e13
    out = ctypes.c_uint64()
    sizes = {}
    for prec in (L.CBS_FP64, L.CBS_FP32_REFINE):
        o = fm.make_opts(precision=prec)
        L.check(lib.cbs_request_footprint(ctypes.byref(d), ctypes.byref(o), ctypes.byref(out)))
        sizes[prec] = out.value
    assert sizes[L.CBS_FP32_REFINE] > 0 and sizes[L.CBS_FP64] > 64 ** 3 * 16 * 4
    d.n[0] = 0
    with pytest.raises(ValueError):
        L.check(lib.cbs_request_footprint(ctypes.byref(d), ctypes.byref(fm.make_opts()), ctypes.byref(out)))


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly(fm):
    with pytest.raises((fm.CBSError, ValueError)):
        fm.Context(0)


# ---------------------------------------------------------------------------
# field analysis (src/fields/electromagnetic_field.jl:97-237), host mirror
# ---------------------------------------------------------------------------
def test_field_analysis_reference_cases(fm):
    """test/fields/electromagnetic_field.jl "Field Utilities" and "Field Domain
    Properties", with the closed forms of a uniform field."""
    E = np.ones((16, 16, 8, 3), dtype=np.complex128) * 0.1
    H = np.ones((16, 16, 8, 3), dtype=np.complex128) * 0.01
    res = (50e-9, 50e-9, 100e-9)
    f = fm.ElectromagneticField(E, H, res, 633e-9)
    n = 16 * 16 * 8
    want = (8.854187817e-12 * 3 * n * 0.01 + 4e-7 * math.pi * 3 * n * 1e-4) / 2 * (res[0] * res[1] * res[2])
    assert abs(fm.field_energy(f) - want) <= 1e-14 * want
    inten = fm.field_intensity(f)
    assert inten.shape == (16, 16, 8) and np.allclose(inten, 0.03, rtol=1e-15, atol=0)
    assert np.allclose(fm.field_intensity(f, "y"), 0.01, rtol=1e-15, atol=0)
    assert fm.poynting_vector(f).shape == (16, 16, 8, 3)
    assert np.all(fm.poynting_vector(f) == 0)  # parallel E and H
    g = fm.ElectromagneticField(np.zeros((20, 30, 10, 3), complex), np.zeros((20, 30, 10, 3), complex),
                                (25e-9, 40e-9, 80e-9), 500e-9)
    assert np.allclose(fm.domain_size(g), (20 * 25e-9, 30 * 40e-9, 10 * 80e-9), rtol=1e-15)
    with pytest.raises(ValueError):
        fm.field_intensity(f, "w")
    with pytest.raises(ValueError):
        fm.extract_plane(f, 4, 1)
    with pytest.raises(ValueError):
        fm.extract_plane(f, 3, 9)


def test_field_analysis_plane_wave(fm):
    """x-polarised plane wave along z: S = z_hat |A|^2 / Z, |E|^2 = |A|^2 in x only,
    and a plane keeps the field's values."""
    n = (6, 5, 8)
    z = np.arange(n[2]) * 50e-9
    A = 0.7 * np.exp(1j * 2 * np.pi * z / 500e-9)
    Z = 377.0 / 1.33
    E = np.zeros(n + (3,), complex)
    H = np.zeros(n + (3,), complex)
    E[..., 0] = A[None, None, :]
    H[..., 1] = A[None, None, :] / Z
    f = fm.ElectromagneticField(E, H, (50e-9,) * 3, 500e-9)
    S = fm.poynting_vector(f)
    assert np.allclose(S[..., 2], 0.49 / Z, rtol=1e-14, atol=0)
    assert np.all(S[..., 0] == 0) and np.all(S[..., 1] == 0)
    assert np.allclose(fm.field_intensity(f, "x"), 0.49, rtol=1e-14)
    assert np.all(fm.field_intensity(f, "z") == 0)
    p = fm.extract_plane(f, 3, 4)
    assert p.E.shape == (6, 5, 1, 3) and np.array_equal(p.E[:, :, 0], E[:, :, 3])
    assert np.array_equal(fm.extract_plane(f, 1, 2).H[0], H[1])


def test_refine_inner_tol(tmp_path):
    """The inner tolerance rule of FP32+refine (csrc/cbs_refine.h), compiled
    on its own with g++: equal inner reductions above the FP32 floor 9e-4, and
    the floor for a need of zero (tolerance 0) or a need that is not finite."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <cstdio>
#include <limits>
#include "cbs_refine.h"
int main() {
  const double inf = std::numeric_limits<double>::infinity();
  const double nan = std::numeric_limits<double>::quiet_NaN();
  const double v[] = {0.0, 1e-1, 1e-3, 1e-6, 1e-12, 1e-300, 5e-324, inf, nan, -1.0};
  for (double x : v) std::printf("%.17g\n", cbs::refine_inner_tol(x, 0.0));
  std::printf("%.17g\n", cbs::refine_inner_tol(0.0, 0.25));
  return 0;
}
''')
    exe = tmp_path / "t"
    inc = ROOT / "frequencymaxwell.jl_amd" / "csrc"
    subprocess.run([gxx, "-std=c++17", "-O2", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = [float(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                            text=True).stdout.split()]
    floor = 9e-4
    assert out[0] == floor                               # tolerance 0
    assert out[1] == pytest.approx(0.09)                 # one inner solve, 0.9 need
    assert out[2] == pytest.approx(0.9e-3)               # at the floor: one inner solve
    for need, eta in zip((1e-6, 1e-12, 1e-300, 5e-324), out[3:7]):
        t = 0.9 * need
        m = math.ceil(math.log(t) / math.log(floor))
        assert m >= 2 and eta == pytest.approx(t ** (1.0 / m), rel=1e-12)
        assert floor <= eta < 1.0                       # m equal reductions above the floor
    assert out[7] == floor and out[8] == floor and out[9] == floor
    assert out[10] == 0.25                               # a caller's inner_tol wins
