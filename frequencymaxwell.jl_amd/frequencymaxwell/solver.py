"""ConvergentBornSolver and ``solve`` -- the reference's API over the HIP engine.

Mirrors src/solvers/convergent_born.jl: the solver object (:87-233), ROI
(:250-260), the preconditioner and linear operator (:379-518), ``solve``
(:565-600), ``_initialize_solver!`` (:619-675), ``_solve_cbs_scattering``
(:703-774), the magnetic field (:784-800), padding/attenuation/crop
(:807-960), ``reset!`` and the size helpers (:967-1021).

Host side (as in the reference): potential build, incident field, source
term, crop.  Device side (``libcbs_mi355x.so``): Green operator, CBS operator,
preconditioner, curl and the whole iterative solve.
"""

from __future__ import annotations

import math
import warnings

import numpy as np

from . import _lib as L
from .boundary_conditions import (AbsorbingBoundaryCondition, AbstractBoundaryCondition,
                                  convert_boundary, create_attenuation_window, padding_pixels,
                                  requires_averaging, requires_padding, subpixel_shift)
from .engine import CBSProblem, GreenSpec, make_opts, method_code
from .fields import ElectromagneticField, crop_array
from .medium import Medium, grid_size, permittivity, permittivity_bg
from .sources import generate_incident_field, source_wavelength

Z0 = 377.0
DEVICES = ("mi355x",)


class ConvergentBornSolver:
    """ConvergentBornSolver(; resolution, grid_size, boundary_conditions, iterations_max=-1,
    tolerance=1e-6, linear_solver="gmres", preconditioner="none", device="mi355x").

    Extra keywords of this engine (defaults keep the reference behaviour):
    ``precision`` ("fp64" | "fp32" | "fp32_refine"), ``stop_rule``
    ("krylov": reltol on the initial residual plus abstol = tol*1e-2 as
    LinearSolve passes it; "rhs": tol * ||gamma b||), ``maxiters`` (default
    max(2 Bornmax, 100)), ``gmres_restart``, ``use_graph``, ``check_every``,
    ``device_setup`` (potential, incident field, source term and the final
    incident add + crop on the GPU; False keeps them on the host).
    """

    def __init__(self, *, resolution, grid_size, boundary_conditions, iterations_max: int = -1,
                 tolerance: float = 1e-6, linear_solver="gmres", preconditioner: str = "none",
                 device: str = "mi355x", precision: str = "fp64", stop_rule: str = "krylov",
                 maxiters: int | None = None, gmres_restart: int = 0, use_graph: bool = False,
                 check_every: int = 1, device_setup: bool = True):
        resolution = tuple(float(r) for r in resolution)
        grid_size = tuple(int(g) for g in grid_size)
        if len(resolution) != 3 or len(grid_size) != 3:
            raise ValueError("resolution and grid_size need three components")
        if not all(r > 0 for r in resolution):  # convergent_born.jl:117
            raise ValueError("all resolution components must be positive")
        if not all(g > 0 for g in grid_size):  # :119
            raise ValueError("all grid_size components must be positive")
        if not (0 <= tolerance <= 1):  # :121
            raise ValueError("tolerance must be in [0, 1]")
        if isinstance(boundary_conditions, AbstractBoundaryCondition):  # :207-217
            bcs = (boundary_conditions,) * 3
        else:
            bcs = tuple(convert_boundary(b) for b in boundary_conditions)
            if len(bcs) != 3:
                raise ValueError("boundary_conditions needs one or three entries")
        dev = str(device).lower().lstrip(":")
        if dev not in DEVICES:  # create_backend_from_symbol, :335-361
            raise ValueError(f"device :{dev} is not served by this engine; it runs on :mi355x only")
        method_code(linear_solver)
        self.resolution = resolution
        self.grid_size = grid_size
        self.boundary_conditions = bcs
        self.iterations_max = int(iterations_max)
        self.tolerance = float(tolerance)
        self.linear_solver = linear_solver
        self.preconditioner = preconditioner  # stored, unused (as in the reference)
        self.device = dev
        self.precision = precision
        self.stop_rule = stop_rule
        self.maxiters = maxiters
        self.gmres_restart = int(gmres_restart)
        self.use_graph = bool(use_graph)
        self.check_every = int(check_every)
        self.device_setup = bool(device_setup)
        # state
        self.Green_function: DyadicGreen | None = None
        self.Bornmax = 0
        self.iteration_count = 0
        self.residual_history: list = []
        self.last_stats: dict | None = None

    # convergent_born.jl:250-260
    def compute_ROI(self):
        p = self.padding
        g = self.grid_size
        return (p[0] + 1, p[0] + g[0], p[1] + 1, p[1] + g[1], p[2] + 1, p[2] + g[2])

    @property
    def padding(self):
        return tuple(padding_pixels(self.boundary_conditions[d], self.resolution[d]) for d in range(3))

    @property
    def padded_size(self):
        return tuple(self.grid_size[d] + 2 * self.padding[d] for d in range(3))

    def reset(self):  # reset!, :967-972
        self.iteration_count = 0
        self.residual_history = []

    def domain_size(self):  # :1015-1021
        return tuple(self.grid_size[i] * self.resolution[i] for i in range(3))

    def grid_spacing(self):  # :1012
        return self.resolution

    def __repr__(self):
        return (f"ConvergentBornSolver{{Float64}}:\n  resolution: {self.resolution}\n"
                f"  grid_size: {self.grid_size}\n  device: :{self.device}\n"
                f"  iterations: {self.iteration_count}")


def compute_ROI(solver: ConvergentBornSolver):
    return solver.compute_ROI()


def domain_size(obj):
    """domain_size of a solver (:1015-1021) or of an ElectromagneticField
    (electromagnetic_field.jl:97-100)."""
    if isinstance(obj, ElectromagneticField):
        from .fields import field_domain_size

        return field_domain_size(obj)
    return _solver_domain_size(obj)


def _solver_domain_size(solver):
    return solver.domain_size()


def grid_spacing(solver):
    return solver.grid_spacing()


def reset(solver):
    solver.reset()


# ---------------------------------------------------------------------------
# operators (the reference's device objects, backed by the HIP library)
# ---------------------------------------------------------------------------
class DyadicGreen:
    """DyadicGreen (dyadic_green.jl:45-99); ``conv`` = ``conv!`` on a copy."""

    def __init__(self, problem: CBSProblem):
        self.problem = problem
        s = problem.spec
        self.k_square = s.k_square
        self.freq_res = tuple(2 * math.pi / (s.shape[d] * s.resolution[d]) for d in range(3))
        self.shifts = s.shifts
        self.use_averaging = s.use_averaging

    def conv(self, src) -> np.ndarray:
        return self.problem.green_apply(src)


class CBSLinearOperator:
    """(I - G_avg V) as a LinearSolve operator (convergent_born.jl:452-518)."""

    def __init__(self, problem: CBSProblem):
        self.problem = problem
        n = int(np.prod(problem.field_shape))
        self.shape = (n, n)

    def mul(self, x, alpha=1.0, beta=0.0, y=None) -> np.ndarray:
        return self.problem.operator_apply(x, y, alpha, beta)


class CBSPreconditioner:
    """Left preconditioner (i/eps) V (convergent_born.jl:379-431)."""

    def __init__(self, problem: CBSProblem):
        self.problem = problem

    def ldiv(self, x) -> np.ndarray:
        return self.problem.precond_apply(x)


def conv(green: DyadicGreen, src) -> np.ndarray:
    return green.conv(src)


# ---------------------------------------------------------------------------
# potential build -- convergent_born.jl:619-675, 807-944
# ---------------------------------------------------------------------------
def _pad_array(arr: np.ndarray, pad) -> np.ndarray:
    """:807-858 -- ``:replicate`` fills zeros (the replicate helpers are no-ops)."""
    if all(p == 0 for p in pad):
        return np.array(arr, dtype=np.complex128, order="F", copy=True)
    shape = tuple(arr.shape[d] + 2 * pad[d] for d in range(3))
    out = np.zeros(shape, dtype=np.complex128, order="F")
    out[pad[0]:pad[0] + arr.shape[0], pad[1]:pad[1] + arr.shape[1], pad[2]:pad[2] + arr.shape[2]] = arr
    return out


def attenuation_windows(solver: ConvergentBornSolver):
    """Per axis, the sharpness-blended attenuation window of
    _apply_attenuation_to_potential! (:893-944): length L = min(pad,
    round(attenuation_thickness / (2 dx))); empty where the axis does not attenuate."""
    out = []
    for dim in range(3):
        bc = solver.boundary_conditions[dim]
        if not requires_padding(bc) or not isinstance(bc, AbsorbingBoundaryCondition):
            out.append(np.zeros(0))
            continue
        max_L = padding_pixels(bc, solver.resolution[dim])
        if max_L == 0:
            out.append(np.zeros(0))
            continue
        att = int(np.rint(bc.attenuation_thickness / (solver.resolution[dim] * 2)))
        L_ = min(max_L, att)
        out.append(create_attenuation_window(bc.profile, L_) * bc.sharpness + (1 - bc.sharpness))
    return out


def _apply_attenuation_to_potential(solver: ConvergentBornSolver, potential: np.ndarray) -> None:
    """:893-944 -- multiply V by a 1-D window along every absorbing axis:
    filter = [w; ones; reverse(w)] over the padded axis."""
    for dim, w in enumerate(attenuation_windows(solver)):
        bc = solver.boundary_conditions[dim]
        if not requires_padding(bc) or not isinstance(bc, AbsorbingBoundaryCondition):
            continue
        if padding_pixels(bc, solver.resolution[dim]) == 0:
            continue
        n = potential.shape[dim]
        filt = np.concatenate([w, np.ones(n - 2 * w.size), w[::-1]])
        shape = [1, 1, 1]
        shape[dim] = filt.size
        potential *= filt.reshape(shape)


def _initialize_solver(solver: ConvergentBornSolver, medium: Medium, wavelength: float):
    """_initialize_solver! -- returns (potential, eps_imag, Bornmax, k_bg, GreenSpec)."""
    eps = permittivity(medium)
    eps_bg = permittivity_bg(medium)
    k_bg = 2 * math.pi * math.sqrt(eps_bg) / wavelength
    v0 = (k_bg ** 2 * (eps / eps_bg - 1)).astype(np.complex128)
    pot = _pad_array(v0, solver.padding)
    eps_imag = max(2.0 ** -20, float(np.max(np.abs(pot))) * 1.01)
    pot -= 1j * eps_imag
    _apply_attenuation_to_potential(solver, pot)
    if solver.iterations_max < 0:
        steps = abs(2 * k_bg / eps_imag)
        extent = math.sqrt(sum((pot.shape[d] * solver.resolution[d]) ** 2 for d in range(3)))
        bornmax = math.ceil(extent / steps / 2 + 1) * 2 * abs(solver.iterations_max)
    else:
        bornmax = solver.iterations_max
    spec = GreenSpec(shape=pot.shape, resolution=solver.resolution,
                     shifts=tuple(subpixel_shift(b) for b in solver.boundary_conditions),
                     use_averaging=any(requires_averaging(b) for b in solver.boundary_conditions),
                     k_square=k_bg ** 2 + 1j * eps_imag, eps_imag=eps_imag)
    return pot, eps_imag, bornmax, k_bg, spec


def _solve_opts(solver: ConvergentBornSolver, bornmax: int, wavelength: float):
    maxiters = solver.maxiters if solver.maxiters is not None else max(2 * bornmax, 100)
    return make_opts(method=solver.linear_solver, precision=solver.precision,
                     tol=solver.tolerance, abstol=solver.tolerance * 1e-2,
                     stop_rule=solver.stop_rule, maxiters=maxiters,
                     gmres_restart=solver.gmres_restart, use_graph=solver.use_graph,
                     check_every=solver.check_every, h_scale=h_scale(wavelength))


def h_scale(wavelength: float) -> complex:
    """``-i lambda / (2 pi 377)`` (convergent_born.jl:797)."""
    return complex(0, -1) * wavelength / (2 * math.pi * Z0)


def _prepare(solver, sources, medium):
    if grid_size(medium) != solver.grid_size:  # :572-573
        raise ValueError(f"medium grid_size {grid_size(medium)} must match solver grid_size {solver.grid_size}")
    lam = source_wavelength(sources)
    pot, eps_imag, bornmax, k_bg, spec = _initialize_solver(solver, medium, lam)
    if bornmax < 1:  # the reference would leave Efield undefined (:717-773)
        raise ValueError("iterations_max must be >= 1 (or negative for the automatic count)")
    problem = CBSProblem(pot, spec, precision="fp64" if solver.precision == "fp64" else "fp32")
    solver.Green_function = DyadicGreen(problem)
    solver.Bornmax = bornmax
    solver.iteration_count = 0
    solver.residual_history = []
    return lam, pot, eps_imag, bornmax, problem


def scene(solver: ConvergentBornSolver, sources, medium: Medium, precision=None):
    """The scene description of solve(solver, sources, medium) for the device
    setup: ROI permittivity, padding and attenuation windows, Green shifts and
    the plane waves (unit vectors from plane_wave.jl:218-232).  Returns
    (SceneDesc, eps, [PlaneWave], keep) -- keep holds the window arrays the
    description points to."""
    from .engine import precision_code
    from .sources import AbstractCurrentSource, plane_wave_vectors

    if grid_size(medium) != solver.grid_size:  # :572-573
        raise ValueError(f"medium grid_size {grid_size(medium)} must match solver grid_size {solver.grid_size}")
    lam = source_wavelength(sources)
    eps_bg = permittivity_bg(medium)
    srcs = [sources] if isinstance(sources, AbstractCurrentSource) else list(sources)
    if not srcs:
        raise ValueError("sources vector cannot be empty")
    for s_ in srcs[1:]:
        if not math.isclose(s_.wavelength, lam, rel_tol=1e-10, abs_tol=0.0):
            raise ValueError("All sources must have the same wavelength for coherent interference")
    d = L.SceneDesc()
    wins = [np.ascontiguousarray(w, dtype=np.float64) for w in attenuation_windows(solver)]
    pad = solver.padding
    for k in range(3):
        d.grid[k] = solver.grid_size[k]
        d.res[k] = solver.resolution[k]
        d.pad[k] = pad[k]
        d.att_len[k] = wins[k].size
        d.window[k] = wins[k].ctypes.data_as(L._D) if wins[k].size else None
        d.shift[k] = subpixel_shift(solver.boundary_conditions[k])
    d.averaging = int(any(requires_averaging(b) for b in solver.boundary_conditions))
    d.eps_bg = float(eps_bg)
    d.wavelength = float(lam)
    d.iterations_max = int(solver.iterations_max)
    prec = precision if precision is not None else solver.precision
    d.precision = L.CBS_FP64 if precision_code(prec) == L.CBS_FP64 else L.CBS_FP32
    waves = []
    for src in srcs:
        k_vec, e_perp, h_perp = plane_wave_vectors(src, eps_bg)
        w = L.PlaneWave()
        for c in range(3):
            w.k_vec[c] = float(k_vec[c])
            w.e_perp[2 * c], w.e_perp[2 * c + 1] = complex(e_perp[c]).real, complex(e_perp[c]).imag
            w.h_perp[2 * c], w.h_perp[2 * c + 1] = complex(h_perp[c]).real, complex(h_perp[c]).imag
        w.amplitude = float(src.amplitude)
        w.phase = float(src.phase)
        waves.append(w)
    eps = np.asfortranarray(permittivity(medium), dtype=np.complex128)
    return d, eps, waves, wins


def _record(solver, st: L.SolveStats):
    solver.iteration_count = int(st.iters)
    solver.residual_history.append(float(st.resid))
    solver.last_stats = st.as_dict()
    if not st.converged:  # :753-755
        warnings.warn("LinearSolve.jl did not converge.")


def solve(solver: ConvergentBornSolver, sources, medium: Medium) -> ElectromagneticField:
    """solve(solver, sources, medium) -> total field cropped to the ROI (:565-600).

    With ``solver.device_setup`` (the default) the potential, the incident
    field, the source term, H and the incident add + crop run on the GPU
    (cbs_solve_scene); otherwise they run on the host as in the reference and
    only the operator layer and the solve go to the GPU."""
    if solver.device_setup:
        from .engine import solve_scene

        d, eps, waves, keep = scene(solver, sources, medium)
        lam = d.wavelength
        opts = _solve_opts(solver, 0, lam)
        if solver.maxiters is None:
            opts.maxiters = 0  # the library applies max(2 Bornmax, 100)
        E, H, st, info = solve_scene(d, eps, waves, opts)
        del keep
        solver.Green_function = None
        solver.Bornmax = int(info.bornmax)
        solver.iteration_count = 0
        solver.residual_history = []
        _record(solver, st)
        return ElectromagneticField(E, H, solver.resolution, lam)
    lam, pot, eps_imag, bornmax, problem = _prepare(solver, sources, medium)
    inc = generate_incident_field(sources, solver, permittivity_bg=permittivity_bg(medium))
    S = pot[..., None] * inc.E + complex(0, eps_imag) * inc.E  # :585-586
    E, H, st = problem.solve(S, _solve_opts(solver, bornmax, lam))
    _record(solver, st)
    roi = solver.compute_ROI()
    return ElectromagneticField(crop_array(E + inc.E, roi), crop_array(H + inc.H, roi),
                                solver.resolution, lam)


def solve_batch(solver: ConvergentBornSolver, sources_list, medium: Medium):
    """One ``solve`` per entry of ``sources_list`` (each entry a source or a
    vector of coherent sources), all right-hand sides swept together."""
    lam = source_wavelength(sources_list[0])
    for s in sources_list[1:]:
        if source_wavelength(s) != lam:
            raise ValueError("solve_batch needs one wavelength for all right-hand sides")
    if solver.device_setup:
        from .engine import solve_scene_multi
        from .sources import AbstractCurrentSource

        groups = [[s] if isinstance(s, AbstractCurrentSource) else list(s) for s in sources_list]
        flat = [s for g in groups for s in g]
        d, eps, waves, keep = scene(solver, flat, medium)
        per, i = [], 0
        for g in groups:
            per.append(waves[i:i + len(g)])
            i += len(g)
        opts = _solve_opts(solver, 0, lam)
        if solver.maxiters is None:
            opts.maxiters = 0  # the library applies max(2 Bornmax, 100)
        Es, Hs, sts, info = solve_scene_multi(d, eps, per, opts)
        del keep
        solver.Green_function = None
        solver.Bornmax = int(info.bornmax)
        solver.iteration_count = 0
        solver.residual_history = []
        out = []
        for E, H, st in zip(Es, Hs, sts):
            _record(solver, st)
            out.append(ElectromagneticField(E, H, solver.resolution, lam))
        return out
    lam, pot, eps_imag, bornmax, problem = _prepare(solver, sources_list[0], medium)
    incs = [generate_incident_field(s, solver, permittivity_bg=permittivity_bg(medium))
            for s in sources_list]
    Ss = [pot[..., None] * inc.E + complex(0, eps_imag) * inc.E for inc in incs]
    Es, Hs, sts = problem.solve_multi(Ss, _solve_opts(solver, bornmax, lam))
    roi = solver.compute_ROI()
    out = []
    for E, H, inc, st in zip(Es, Hs, incs, sts):
        _record(solver, st)
        out.append(ElectromagneticField(crop_array(E + inc.E, roi), crop_array(H + inc.H, roi),
                                        solver.resolution, lam))
    return out


def compute_magnetic_field(problem: CBSProblem, E, wavelength: float) -> np.ndarray:
    """_compute_magnetic_field (:784-800)."""
    return problem.curl_apply(E, h_scale(wavelength))
