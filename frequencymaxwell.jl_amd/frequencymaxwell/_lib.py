"""ctypes binding of ``libcbs_mi355x.so`` (the C ABI declared in include/cbs.h).

The library is the only compute path: if it is missing, or no GPU is
visible, every call raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "lib" / "libcbs_mi355x.so"

# status codes
CBS_OK, CBS_EINVAL, CBS_ENOMEM, CBS_EDEVICE = 0, -1, -2, -3
CBS_ENOTFOUND, CBS_EBUSY, CBS_ECANCELED, CBS_ENUMERIC = -4, -5, -6, -7
# precision
CBS_FP64, CBS_FP32, CBS_FP32_REFINE = 0, 1, 2
# method
CBS_BORN, CBS_BICGSTAB, CBS_GMRES = 0, 1, 2
# stop rule
CBS_STOP_RHS, CBS_STOP_KRYLOV = 0, 1
# serving policy
CBS_SERVE_SHARE, CBS_SERVE_LEAD = 0, 1

EXPORTED = (
    "cbs_ctx_create", "cbs_ctx_destroy", "cbs_last_error", "cbs_version", "cbs_default_opts",
    "cbs_problem_create", "cbs_problem_destroy", "cbs_green_apply", "cbs_operator_apply",
    "cbs_precond_apply", "cbs_curl_apply", "cbs_solve", "cbs_problem_set_source",
    "cbs_problem_run", "cbs_problem_get_fields", "cbs_solve_multi", "cbs_solve_batch",
    "cbs_request_footprint", "cbs_server_start", "cbs_server_stop", "cbs_submit", "cbs_poll",
    "cbs_wait", "cbs_cancel", "cbs_time_sweeps", "cbs_stream_copy", "cbs_scene_build",
    "cbs_solve_scene", "cbs_submit_scene", "cbs_field_create", "cbs_field_destroy", "cbs_field_get",
    "cbs_field_energy", "cbs_field_intensity", "cbs_field_poynting", "cbs_field_extract_plane",
    "cbs_solve_scene_field", "cbs_solve_scene_multi", "cbs_operator_apply_noalloc",
    "cbs_server_policy",
)


class ProblemDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("res", C.c_double * 3),
        ("shift", C.c_double * 3),
        ("averaging", C.c_int32),
        ("k2_re", C.c_double),
        ("k2_im", C.c_double),
        ("eps_imag", C.c_double),
        ("precision", C.c_int32),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("precision", C.c_int32),
        ("stop_rule", C.c_int32),
        ("maxiters", C.c_int32),
        ("tol", C.c_double),
        ("abstol", C.c_double),
        ("gmres_restart", C.c_int32),
        ("max_refine", C.c_int32),
        ("inner_tol", C.c_double),
        ("check_every", C.c_int32),
        ("use_graph", C.c_int32),
        ("x0_zero", C.c_int32),
        ("compute_h", C.c_int32),
        ("h_scale_re", C.c_double),
        ("h_scale_im", C.c_double),
        ("fuse_sweeps", C.c_int32),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int32),
        ("iters_fp64", C.c_int32),
        ("refine_steps", C.c_int32),
        ("converged", C.c_int32),
        ("resid", C.c_double),
        ("resid_true_fp64", C.c_double),
        ("time_ms", C.c_double),
        ("bytes_moved", C.c_double),
        ("voxel_iters", C.c_int64),
        ("applications", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(C.Structure):
    _fields_ = [
        ("desc", ProblemDesc),
        ("V", C.POINTER(C.c_double)),
        ("S", C.POINTER(C.c_double)),
        ("opts", SolveOpts),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("grid", C.c_int64 * 3),
        ("res", C.c_double * 3),
        ("pad", C.c_int32 * 3),
        ("att_len", C.c_int32 * 3),
        ("window", C.POINTER(C.c_double) * 3),
        ("shift", C.c_double * 3),
        ("averaging", C.c_int32),
        ("eps_bg", C.c_double),
        ("wavelength", C.c_double),
        ("iterations_max", C.c_int32),
        ("precision", C.c_int32),
    ]


class PlaneWave(C.Structure):
    _fields_ = [
        ("k_vec", C.c_double * 3),
        ("e_perp", C.c_double * 6),
        ("h_perp", C.c_double * 6),
        ("amplitude", C.c_double),
        ("phase", C.c_double),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("eps_imag", C.c_double),
        ("k2_re", C.c_double),
        ("k2_im", C.c_double),
        ("bornmax", C.c_int32),
        ("maxiters", C.c_int32),
    ]


class CBSError(RuntimeError):
    """A HIP-library failure (status code in ``.code``)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[cbs status {code}] {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()

_P = C.c_void_p
_D = C.POINTER(C.c_double)


def _declare(lib):
    s = C.c_int32
    sig = {
        "cbs_ctx_create": (s, [C.POINTER(_P), C.c_int32, C.c_uint64]),
        "cbs_ctx_destroy": (s, [_P]),
        "cbs_last_error": (C.c_char_p, []),
        "cbs_version": (C.c_char_p, []),
        "cbs_default_opts": (None, [C.POINTER(SolveOpts)]),
        "cbs_problem_create": (s, [_P, C.POINTER(ProblemDesc), _D, C.POINTER(_P)]),
        "cbs_problem_destroy": (s, [_P]),
        "cbs_green_apply": (s, [_P, _D]),
        "cbs_operator_apply": (s, [_P, _D, _D, C.c_double, C.c_double, C.c_double, C.c_double]),
        "cbs_operator_apply_noalloc": (s, [_P, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                           C.c_double, C.c_double]),
        "cbs_precond_apply": (s, [_P, _D, _D]),
        "cbs_curl_apply": (s, [_P, _D, C.c_double, C.c_double]),
        "cbs_solve": (s, [_P, _D, _D, _D, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
        "cbs_problem_set_source": (s, [_P, _D]),
        "cbs_problem_run": (s, [_P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
        "cbs_problem_get_fields": (s, [_P, _D, _D, C.c_double, C.c_double]),
        "cbs_solve_multi": (s, [_P, C.c_int32, C.POINTER(_D), C.POINTER(_D), C.POINTER(_D),
                                C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
        "cbs_solve_batch": (s, [_P, C.c_int32, C.POINTER(_P), C.POINTER(_D), C.POINTER(_D),
                                C.POINTER(_D), C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
        "cbs_request_footprint": (s, [C.POINTER(ProblemDesc), C.POINTER(SolveOpts),
                                      C.POINTER(C.c_uint64)]),
        "cbs_server_start": (s, [_P, C.c_int32]),
        "cbs_server_stop": (s, [_P]),
        "cbs_server_policy": (s, [_P, C.c_int32, C.c_double]),
        "cbs_submit": (s, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
        "cbs_poll": (s, [_P, C.c_uint64, C.POINTER(C.c_int32), C.POINTER(SolveStats)]),
        "cbs_wait": (s, [_P, C.c_uint64, _D, _D, C.POINTER(SolveStats)]),
        "cbs_cancel": (s, [_P, C.c_uint64]),
        "cbs_time_sweeps": (s, [_P, C.c_int32, C.c_int32, _D, _D]),
        "cbs_stream_copy": (s, [_P, C.c_uint64, C.c_int32, _D]),
        "cbs_scene_build": (s, [_P, C.POINTER(SceneDesc), _D, C.c_int32, C.POINTER(PlaneWave),
                                _D, _D, C.POINTER(SceneInfo)]),
        "cbs_solve_scene": (s, [_P, C.POINTER(SceneDesc), _D, C.c_int32, C.POINTER(PlaneWave),
                                C.POINTER(SolveOpts), _D, _D, C.POINTER(SolveStats),
                                C.POINTER(SceneInfo)]),
        "cbs_submit_scene": (s, [_P, C.POINTER(SceneDesc), _D, C.c_int32, C.POINTER(PlaneWave),
                                 C.POINTER(SolveOpts), C.POINTER(C.c_uint64)]),
        "cbs_field_create": (s, [_P, C.POINTER(C.c_int64), _D, C.c_double, _D, _D, C.POINTER(_P)]),
        "cbs_field_destroy": (s, [_P]),
        "cbs_field_get": (s, [_P, _D, _D]),
        "cbs_field_energy": (s, [_P, _D]),
        "cbs_field_intensity": (s, [_P, C.c_int32, _D]),
        "cbs_field_poynting": (s, [_P, _D]),
        "cbs_field_extract_plane": (s, [_P, C.c_int32, C.c_int64, _D, _D]),
        "cbs_solve_scene_multi": (s, [_P, C.POINTER(SceneDesc), _D, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(PlaneWave), C.POINTER(SolveOpts), C.POINTER(_D),
                                      C.POINTER(_D), C.POINTER(SolveStats), C.POINTER(SceneInfo)]),
        "cbs_solve_scene_field": (s, [_P, C.POINTER(SceneDesc), _D, C.c_int32, C.POINTER(PlaneWave),
                                      C.POINTER(SolveOpts), C.POINTER(_P), C.POINTER(SolveStats),
                                      C.POINTER(SceneInfo)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def load(path: os.PathLike | None = None):
    """Load the library (once).  Raises if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            env = os.environ.get("CBS_LIBRARY")
            p = Path(path) if path else (Path(env) if env else LIB_PATH)
            if not p.exists():
                raise CBSError(CBS_EDEVICE, f"HIP library not built: {p} (run __graft_entry__.build())")
            lib = C.CDLL(str(p))
            _declare(lib)
            _lib = lib
        return _lib


def check(status: int):
    if status != CBS_OK:
        msg = load().cbs_last_error().decode(errors="replace")
        if status == CBS_EINVAL:
            raise ValueError(msg)
        if status == CBS_ENOMEM:
            raise MemoryError(msg)
        raise CBSError(status, msg)


def dptr(a: np.ndarray):
    """``double*`` of a complex128 array (must be contiguous)."""
    return a.ctypes.data_as(_D)


def as_field(a, shape=None) -> np.ndarray:
    """Complex128, Fortran-contiguous copy/view of a field array."""
    out = np.asfortranarray(a, dtype=np.complex128)
    if shape is not None and out.shape != tuple(shape):
        raise ValueError(f"field shape {out.shape} != {tuple(shape)}")
    return out


def default_opts() -> SolveOpts:
    o = SolveOpts()
    load().cbs_default_opts(C.byref(o))
    return o


class Context:
    """One GPU (``cbs_ctx``): a HIP stream and an optional HBM budget."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._lib = load()
        h = _P()
        check(self._lib.cbs_ctx_create(C.byref(h), int(device), int(hbm_budget)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            check(self._lib.cbs_ctx_destroy(self.handle))
            self.handle = _P()

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self._lib.cbs_ctx_destroy(self.handle)
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    """The process-wide context on GPU 0 (created on first use)."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
