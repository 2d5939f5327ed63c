"""Thin object layer over the C ABI: one ``CBSProblem`` per padded potential.

Every method calls ``libcbs_mi355x.so``; there is no other compute path.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

_PREC = {"fp64": L.CBS_FP64, "fp32": L.CBS_FP32, "fp32_refine": L.CBS_FP32_REFINE}
_METHOD = {"born": L.CBS_BORN, "bicgstab": L.CBS_BICGSTAB, "gmres": L.CBS_GMRES}
_RULE = {"rhs": L.CBS_STOP_RHS, "krylov": L.CBS_STOP_KRYLOV}


def precision_code(p) -> int:
    if isinstance(p, int):
        return p
    try:
        return _PREC[str(p).lower()]
    except KeyError:
        raise ValueError(f"unknown precision {p!r} (fp64, fp32, fp32_refine)") from None


def method_code(m) -> int:
    if isinstance(m, int):
        return m
    key = str(m).lower().replace("krylovjl_", "").replace("()", "")
    try:
        return _METHOD[key]
    except KeyError:
        raise ValueError(f"unknown linear solver {m!r} (gmres, bicgstab, born)") from None


def make_opts(*, method="born", precision="fp64", tol=1e-6, abstol=0.0, stop_rule="rhs",
              maxiters=1000, gmres_restart=0, max_refine=20, inner_tol=0.0, check_every=1,
              use_graph=False, x0_zero=False, h_scale=0j, compute_h=True,
              fuse_sweeps=True) -> L.SolveOpts:
    o = L.default_opts()
    o.method = method_code(method)
    o.precision = precision_code(precision)
    o.stop_rule = _RULE[stop_rule] if isinstance(stop_rule, str) else int(stop_rule)
    o.maxiters = int(maxiters)
    o.tol = float(tol)
    o.abstol = float(abstol)
    o.gmres_restart = int(gmres_restart)
    o.max_refine = int(max_refine)
    o.inner_tol = float(inner_tol)
    o.check_every = int(check_every)
    o.use_graph = int(bool(use_graph))
    o.x0_zero = int(bool(x0_zero))
    o.compute_h = int(bool(compute_h))
    o.h_scale_re = float(complex(h_scale).real)
    o.h_scale_im = float(complex(h_scale).imag)
    o.fuse_sweeps = int(bool(fuse_sweeps))
    return o


@dataclass
class GreenSpec:
    """Parameters of the dyadic Green operator (src/core/dyadic_green.jl:45-92)."""

    shape: tuple          # padded grid (Nx, Ny, Nz)
    resolution: tuple
    shifts: tuple         # per-axis subpixel shift
    use_averaging: bool
    k_square: complex     # k_bg^2 + i eps
    eps_imag: float


def make_desc(spec: GreenSpec, precision="fp64") -> L.ProblemDesc:
    """``cbs_problem_desc`` of a Green operator."""
    d = L.ProblemDesc()
    for k in range(3):
        d.n[k] = int(spec.shape[k])
        d.res[k] = float(spec.resolution[k])
        d.shift[k] = float(spec.shifts[k])
    d.averaging = int(bool(spec.use_averaging))
    d.k2_re = float(complex(spec.k_square).real)
    d.k2_im = float(complex(spec.k_square).imag)
    d.eps_imag = float(spec.eps_imag)
    d.precision = precision_code(precision)
    return d


class CBSProblem:
    """A padded potential V uploaded to the GPU plus its Green operator."""

    def __init__(self, V: np.ndarray, spec: GreenSpec, precision="fp64", ctx: L.Context | None = None):
        self.ctx = ctx or L.default_context()
        self._lib = L.load()
        self.spec = spec
        self.shape = tuple(int(s) for s in spec.shape)
        V = L.as_field(V, self.shape)
        d = make_desc(spec, precision)
        self.desc = d
        h = C.c_void_p()
        L.check(self._lib.cbs_problem_create(self.ctx.handle, C.byref(d), L.dptr(V), C.byref(h)))
        self.handle = h

    @property
    def field_shape(self):
        return self.shape + (3,)

    def close(self):
        if self.handle:
            L.check(self._lib.cbs_problem_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self._lib.cbs_problem_destroy(self.handle)
        except Exception:
            pass

    # -- operator layer --------------------------------------------------
    def green_apply(self, field) -> np.ndarray:
        """``conv!(G, src)`` on a copy: returns G_avg field."""
        f = np.array(L.as_field(field, self.field_shape), order="F", copy=True)
        L.check(self._lib.cbs_green_apply(self.handle, L.dptr(f)))
        return f

    def operator_apply(self, x, y=None, alpha=1.0, beta=0.0) -> np.ndarray:
        """``mul!(y, op, x, alpha, beta)``: returns alpha (x - G_avg V x) + beta y."""
        x = L.as_field(x, self.field_shape)
        if y is None:
            y = np.zeros(self.field_shape, dtype=np.complex128, order="F")
        else:
            y = np.array(L.as_field(y, self.field_shape), order="F", copy=True)
        a, b = complex(alpha), complex(beta)
        L.check(self._lib.cbs_operator_apply(self.handle, L.dptr(y), L.dptr(x), a.real, a.imag,
                                             b.real, b.imag))
        return y

    def precond_apply(self, x) -> np.ndarray:
        """``ldiv!(y, Pl, x)``: returns (i/eps) V x."""
        x = L.as_field(x, self.field_shape)
        y = np.empty(self.field_shape, dtype=np.complex128, order="F")
        L.check(self._lib.cbs_precond_apply(self.handle, L.dptr(y), L.dptr(x)))
        return y

    def curl_apply(self, field, scale=1.0 + 0j) -> np.ndarray:
        """``scale * conv(Curl, field)``."""
        f = np.array(L.as_field(field, self.field_shape), order="F", copy=True)
        s = complex(scale)
        L.check(self._lib.cbs_curl_apply(self.handle, L.dptr(f), s.real, s.imag))
        return f

    # -- solves ------------------------------------------------------------
    def solve(self, S, opts: L.SolveOpts, want_h: bool = True):
        S = L.as_field(S, self.field_shape)
        E = np.empty(self.field_shape, dtype=np.complex128, order="F")
        H = np.empty(self.field_shape, dtype=np.complex128, order="F") if want_h else None
        st = L.SolveStats()
        L.check(self._lib.cbs_solve(self.handle, L.dptr(S), L.dptr(E),
                                    L.dptr(H) if H is not None else None, C.byref(opts),
                                    C.byref(st)))
        return E, H, st

    def set_source(self, S):
        S = L.as_field(S, self.field_shape)
        L.check(self._lib.cbs_problem_set_source(self.handle, L.dptr(S)))

    def run(self, opts: L.SolveOpts) -> L.SolveStats:
        st = L.SolveStats()
        L.check(self._lib.cbs_problem_run(self.handle, C.byref(opts), C.byref(st)))
        return st

    def get_fields(self, h_scale=None):
        E = np.empty(self.field_shape, dtype=np.complex128, order="F")
        H = None
        hs = 0j
        if h_scale is not None:
            H = np.empty(self.field_shape, dtype=np.complex128, order="F")
            hs = complex(h_scale)
        L.check(self._lib.cbs_problem_get_fields(self.handle, L.dptr(E),
                                                 L.dptr(H) if H is not None else None,
                                                 hs.real, hs.imag))
        return E, H

    def time_sweeps(self, precision="fp32", reps: int = 20):
        """Average duration (ms) and algorithmic bytes of each sweep kernel,
        timed with HIP events on the library stream: x forward, y forward,
        z + Green, y inverse, x inverse + Born update, and the fused x pass
        (x inverse + Born update + next x forward) of the Born sweeps."""
        ms = (C.c_double * 6)()
        by = (C.c_double * 6)()
        L.check(self._lib.cbs_time_sweeps(self.handle, precision_code(precision), int(reps), ms, by))
        return list(ms), list(by)

    def solve_multi(self, S_list, opts: L.SolveOpts, want_h: bool = True):
        """Many right-hand sides of this problem in one set of launches per sweep."""
        n = len(S_list)
        Ss = [L.as_field(s, self.field_shape) for s in S_list]
        Es = [np.empty(self.field_shape, dtype=np.complex128, order="F") for _ in range(n)]
        Hs = [np.empty(self.field_shape, dtype=np.complex128, order="F") for _ in range(n)] if want_h else None
        ArrD = L._D * n
        sts = (L.SolveStats * n)()
        L.check(self._lib.cbs_solve_multi(
            self.handle, n, ArrD(*[L.dptr(a) for a in Ss]), ArrD(*[L.dptr(a) for a in Es]),
            ArrD(*[L.dptr(a) for a in Hs]) if want_h else None, C.byref(opts), sts))
        return Es, Hs, list(sts)


def stream_copy_ms(nbytes: int, reps: int = 20, ctx=None) -> float:
    """Average duration (ms) of a plain device copy of ``nbytes`` bytes."""
    ctx = ctx or L.default_context()
    ms = C.c_double()
    L.check(L.load().cbs_stream_copy(ctx.handle, int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def solve_batch(problems, S_list, opts: L.SolveOpts, want_h: bool = True, ctx=None):
    """Problems of possibly different grids, one RHS each, swept together."""
    ctx = ctx or problems[0].ctx
    lib = L.load()
    n = len(problems)
    Ss = [L.as_field(s, p.field_shape) for p, s in zip(problems, S_list)]
    Es = [np.empty(p.field_shape, dtype=np.complex128, order="F") for p in problems]
    Hs = [np.empty(p.field_shape, dtype=np.complex128, order="F") for p in problems] if want_h else None
    ArrD = L._D * n
    ArrP = C.c_void_p * n
    sts = (L.SolveStats * n)()
    L.check(lib.cbs_solve_batch(ctx.handle, n, ArrP(*[p.handle for p in problems]),
                                ArrD(*[L.dptr(a) for a in Ss]), ArrD(*[L.dptr(a) for a in Es]),
                                ArrD(*[L.dptr(a) for a in Hs]) if want_h else None,
                                C.byref(opts), sts))
    return Es, Hs, list(sts)


def request_footprint(spec: GreenSpec, opts: L.SolveOpts) -> int:
    """Bytes of HBM a served request holds while it is resident."""
    out = C.c_uint64()
    d = make_desc(spec, opts.precision)
    L.check(L.load().cbs_request_footprint(C.byref(d), C.byref(opts), C.byref(out)))
    return out.value


class CBSServer:
    """Request loop on one GPU (``cbs_server_*``): submitted solves are admitted
    by HBM footprint, share every sweep launch with the other resident solves
    (grids may differ) and leave as soon as their own stopping rule is met."""

    def __init__(self, max_active: int = 8, ctx: L.Context | None = None, policy: str = "lead",
                 kappa: float = 0.2):
        """``max_active``: requests being loaded or solved at once (a solved
        request's download does not hold a slot).
        ``policy``: "lead" (one request per chunk, first come first served with
        requests of estimated work <= kappa x the mean passing) or "share" (equal
        shares of every chunk in shared launches); see cbs_server_policy."""
        self.ctx = ctx or L.default_context()
        self._lib = L.load()
        self._shapes = {}
        codes = {"share": L.CBS_SERVE_SHARE, "lead": L.CBS_SERVE_LEAD}
        if policy not in codes:
            raise ValueError(f"unknown serving policy {policy!r} (lead, share)")
        L.check(self._lib.cbs_server_start(self.ctx.handle, int(max_active)))
        self.running = True
        L.check(self._lib.cbs_server_policy(self.ctx.handle, codes[policy], float(kappa)))

    def submit(self, V, spec: GreenSpec, S, opts: L.SolveOpts) -> int:
        """Queue one solve (V and S are copied); returns its ticket."""
        shape = tuple(int(x) for x in spec.shape)
        V = L.as_field(V, shape)
        S = L.as_field(S, shape + (3,))
        req = L.Request()
        req.desc = make_desc(spec, opts.precision)
        req.V = L.dptr(V)
        req.S = L.dptr(S)
        req.opts = opts
        t = C.c_uint64()
        L.check(self._lib.cbs_submit(self.ctx.handle, C.byref(req), C.byref(t)))
        self._shapes[t.value] = (shape + (3,), bool(opts.compute_h))
        return t.value

    def submit_scene(self, desc: L.SceneDesc, eps, waves, opts: L.SolveOpts) -> int:
        """Queue a scene (permittivity + plane waves; potential, source term,
        H and the incident add + crop on the GPU); wait() returns the total
        fields cropped to the ROI."""
        g = tuple(int(x) for x in desc.grid)
        eps = L.as_field(eps, g)
        arr = (L.PlaneWave * len(waves))(*waves)
        t = C.c_uint64()
        L.check(self._lib.cbs_submit_scene(self.ctx.handle, C.byref(desc), L.dptr(eps), len(waves),
                                           arr, C.byref(opts), C.byref(t)))
        self._shapes[t.value] = (g + (3,), bool(opts.compute_h))
        return t.value

    def poll(self, ticket: int):
        """(done, stats or None)."""
        done = C.c_int32()
        st = L.SolveStats()
        L.check(self._lib.cbs_poll(self.ctx.handle, int(ticket), C.byref(done), C.byref(st)))
        return bool(done.value), (st if done.value else None)

    def wait(self, ticket: int):
        """Block until the solve ends: (E_scat, H_scat or None, stats)."""
        info = self._shapes.pop(int(ticket), None)
        if info is None:  # not submitted here: the library reports the ticket
            L.check(self._lib.cbs_wait(self.ctx.handle, int(ticket), None, None, None))
            return None, None, None
        shape, want_h = info
        E = np.empty(shape, dtype=np.complex128, order="F")
        H = np.empty(shape, dtype=np.complex128, order="F") if want_h else None
        st = L.SolveStats()
        L.check(self._lib.cbs_wait(self.ctx.handle, int(ticket), L.dptr(E),
                                   L.dptr(H) if H is not None else None, C.byref(st)))
        return E, H, st

    def cancel(self, ticket: int):
        L.check(self._lib.cbs_cancel(self.ctx.handle, int(ticket)))

    def stop(self):
        if self.running:
            L.check(self._lib.cbs_server_stop(self.ctx.handle))
            self.running = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()


# -- device-side setup and teardown (cbs_scene_build / cbs_solve_scene) ------
def scene_build(desc: L.SceneDesc, eps, waves, ctx=None):
    """The padded potential V and source term S as the GPU builds them from the
    ROI permittivity; returns (V, S, SceneInfo)."""
    ctx = ctx or L.default_context()
    eps = L.as_field(eps, tuple(desc.grid))
    n = _padded(desc)
    V = np.empty(n, dtype=np.complex128, order="F")
    S = np.empty(n + (3,), dtype=np.complex128, order="F")
    info = L.SceneInfo()
    arr = (L.PlaneWave * len(waves))(*waves)
    L.check(L.load().cbs_scene_build(ctx.handle, C.byref(desc), L.dptr(eps), len(waves), arr,
                                     L.dptr(V), L.dptr(S), C.byref(info)))
    return V, S, info


def solve_scene(desc: L.SceneDesc, eps, waves, opts: L.SolveOpts, want_h: bool = True, ctx=None):
    """The whole solve on the GPU: returns the total fields cropped to the ROI
    (E, H or None), the solve statistics and the SceneInfo."""
    ctx = ctx or L.default_context()
    g = tuple(desc.grid)
    eps = L.as_field(eps, g)
    E = np.empty(g + (3,), dtype=np.complex128, order="F")
    H = np.empty(g + (3,), dtype=np.complex128, order="F") if want_h else None
    st = L.SolveStats()
    info = L.SceneInfo()
    arr = (L.PlaneWave * len(waves))(*waves)
    L.check(L.load().cbs_solve_scene(ctx.handle, C.byref(desc), L.dptr(eps), len(waves), arr,
                                     C.byref(opts), L.dptr(E), L.dptr(H) if H is not None else None,
                                     C.byref(st), C.byref(info)))
    return E, H, st, info


def _padded(desc) -> tuple:
    return tuple(int(desc.grid[d]) + 2 * int(desc.pad[d]) for d in range(3))


# -- fields resident on the device (cbs_field_*) ------------------------------
class DeviceField:
    """E and H kept in HBM, with the reference's field analysis run on the GPU
    (src/fields/electromagnetic_field.jl:97-237): field_energy,
    field_intensity, poynting_vector, extract_plane."""

    _COMP = {"total": 0, "x": 1, "y": 2, "z": 3}

    def __init__(self, handle, shape, resolution, wavelength, has_h, ctx):
        self.handle = handle
        self.shape = tuple(int(x) for x in shape)
        self.resolution = tuple(float(r) for r in resolution)
        self.wavelength = float(wavelength)
        self.has_h = bool(has_h)
        self.ctx = ctx
        self._lib = L.load()

    @classmethod
    def from_fields(cls, field, ctx=None) -> "DeviceField":
        """Upload an ElectromagneticField (H may be absent)."""
        ctx = ctx or L.default_context()
        E = L.as_field(field.E)
        H = L.as_field(field.H) if field.H is not None else None
        n = (C.c_int64 * 3)(*E.shape[:3])
        res = (C.c_double * 3)(*field.resolution)
        h = C.c_void_p()
        L.check(L.load().cbs_field_create(ctx.handle, n, res, float(field.wavelength), L.dptr(E),
                                          L.dptr(H) if H is not None else None, C.byref(h)))
        return cls(h, E.shape[:3], field.resolution, field.wavelength, H is not None, ctx)

    def close(self):
        if self.handle:
            L.check(self._lib.cbs_field_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self._lib.cbs_field_destroy(self.handle)
        except Exception:
            pass

    def fields(self):
        """Copy back: ElectromagneticField(E, H)."""
        from .fields import ElectromagneticField

        E = np.empty(self.shape + (3,), dtype=np.complex128, order="F")
        H = np.empty(self.shape + (3,), dtype=np.complex128, order="F") if self.has_h else None
        L.check(self._lib.cbs_field_get(self.handle, L.dptr(E), L.dptr(H) if H is not None else None))
        return ElectromagneticField(E, H if H is not None else np.zeros_like(E), self.resolution,
                                    self.wavelength)

    def energy(self) -> float:
        out = C.c_double()
        L.check(self._lib.cbs_field_energy(self.handle, C.byref(out)))
        return out.value

    def intensity(self, component="total") -> np.ndarray:
        c = str(component).lstrip(":")
        if c not in self._COMP:
            raise ValueError("component must be :total, :x, :y, or :z")
        out = np.empty(self.shape, dtype=np.float64, order="F")
        L.check(self._lib.cbs_field_intensity(self.handle, self._COMP[c],
                                              out.ctypes.data_as(L._D)))
        return out

    def poynting(self) -> np.ndarray:
        S = np.empty(self.shape + (3,), dtype=np.complex128, order="F")
        L.check(self._lib.cbs_field_poynting(self.handle, L.dptr(S)))
        return S

    def extract_plane(self, plane_axis: int, plane_index: int):
        """ElectromagneticField of the plane (1-based index) normal to plane_axis (1..3)."""
        from .fields import ElectromagneticField

        shp = list(self.shape)
        if not 1 <= int(plane_axis) <= 3:
            raise ValueError("plane_axis must be 1, 2, or 3")
        shp[int(plane_axis) - 1] = 1
        E = np.empty(tuple(shp) + (3,), dtype=np.complex128, order="F")
        H = np.empty(tuple(shp) + (3,), dtype=np.complex128, order="F") if self.has_h else None
        L.check(self._lib.cbs_field_extract_plane(self.handle, int(plane_axis), int(plane_index),
                                                  L.dptr(E), L.dptr(H) if H is not None else None))
        return ElectromagneticField(E, H if H is not None else np.zeros_like(E), self.resolution,
                                    self.wavelength)


def solve_scene_field(desc: L.SceneDesc, eps, waves, opts: L.SolveOpts, ctx=None):
    """cbs_solve_scene with the cropped totals left on the device: returns
    (DeviceField, stats, SceneInfo)."""
    ctx = ctx or L.default_context()
    g = tuple(int(x) for x in desc.grid)
    eps = L.as_field(eps, g)
    st = L.SolveStats()
    info = L.SceneInfo()
    h = C.c_void_p()
    arr = (L.PlaneWave * len(waves))(*waves)
    L.check(L.load().cbs_solve_scene_field(ctx.handle, C.byref(desc), L.dptr(eps), len(waves), arr,
                                           C.byref(opts), C.byref(h), C.byref(st), C.byref(info)))
    f = DeviceField(h, g, tuple(desc.res), desc.wavelength, bool(opts.compute_h), ctx)
    return f, st, info


def solve_scene_multi(desc: L.SceneDesc, eps, waves_per_rhs, opts: L.SolveOpts, want_e=True,
                      want_h=True, ctx=None):
    """Many right-hand sides of one scene in shared launches: RHS i is the
    coherent sum of the plane waves waves_per_rhs[i].  Returns (Es, Hs, stats,
    info); Es / Hs are lists of cropped totals (None entries when not wanted)."""
    ctx = ctx or L.default_context()
    g = tuple(int(x) for x in desc.grid)
    eps = L.as_field(eps, g)
    n = len(waves_per_rhs)
    flat = [w for ws in waves_per_rhs for w in ws]
    counts = (C.c_int32 * n)(*[len(ws) for ws in waves_per_rhs])
    arr = (L.PlaneWave * len(flat))(*flat)
    Es = [np.empty(g + (3,), dtype=np.complex128, order="F") if want_e else None for _ in range(n)]
    Hs = [np.empty(g + (3,), dtype=np.complex128, order="F") if want_h else None for _ in range(n)]
    ArrD = L._D * n
    pe = ArrD(*[L.dptr(a) if a is not None else None for a in Es])
    ph = ArrD(*[L.dptr(a) if a is not None else None for a in Hs])
    sts = (L.SolveStats * n)()
    info = L.SceneInfo()
    L.check(L.load().cbs_solve_scene_multi(ctx.handle, C.byref(desc), L.dptr(eps), n, counts, arr,
                                           C.byref(opts), pe, ph, sts, C.byref(info)))
    return Es, Hs, list(sts), info
