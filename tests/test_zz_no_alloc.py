"""cbs_operator_apply_noalloc: the allocation-free operator apply on device
arrays gives the same bits as cbs_operator_apply on the same inputs, on calls
in a row whose grids shrink (the scratch array keeps the previous call's
contents) and in both working precisions; a grid whose workspace does not fit
the scratch array is refused."""

import ctypes

import numpy as np
import pytest

from cases import oracle_setup, random_field, spec_from

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Dev:
    """Device buffers through the HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def put(self, a):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.held.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held = []


def _case(fm, oracle, grid, absorb, seed, precision):
    bc = (oracle.AbsorbingBC(thickness=300e-9, attenuation_thickness=200e-9) if absorb
          else oracle.PeriodicBC())
    s, eps, init = oracle_setup(oracle, grid, bc, seed)
    prob = fm.CBSProblem(init.potential, spec_from(fm, init), precision=precision)
    x = random_field(init.potential.shape, seed + 1)
    y = random_field(init.potential.shape, seed + 2)
    return prob, x, y


def _noalloc(fm, dev, prob, x, y, alpha, beta, precision):
    ct = np.complex128 if precision == "fp64" else np.complex64
    xd = np.asfortranarray(x.astype(ct))
    yd = np.asfortranarray(y.astype(ct))
    px, py = dev.put(xd), dev.put(yd)
    a, b = complex(alpha), complex(beta)
    fm._lib.check(fm._lib.load().cbs_operator_apply_noalloc(prob.handle, py, px, a.real, a.imag,
                                                            b.real, b.imag))
    return dev.get(py, yd).astype(np.complex128)


def test_noalloc_matches_operator_apply(fm, oracle):
    dev = _Dev()
    try:
        # three calls in a row; the second grid is smaller than the first, the
        # third smaller still and in FP32
        for k, (grid, absorb, prec) in enumerate((((48, 40, 32), True, "fp64"),
                                                  ((16, 16, 8), False, "fp64"),
                                                  ((12, 16, 8), True, "fp32"))):
            print(f"case noalloc call {k}: grid {grid} absorb={absorb} {prec}", flush=True)
            prob, x, y = _case(fm, oracle, grid, absorb, 70 + k, prec)
            alpha, beta = 0.7 - 0.2j, -0.3 + 1.1j
            want = prob.operator_apply(x, y.copy(order="F"), alpha, beta)
            got = _noalloc(fm, dev, prob, x, y, alpha, beta, prec)
            assert np.array_equal(got, want), f"call {k} differs from cbs_operator_apply"
            # beta = 0: y's old contents are not read
            want0 = prob.operator_apply(x, None, 1.0, 0.0)
            got0 = _noalloc(fm, dev, prob, x, np.full_like(y, np.nan), 1.0, 0.0, prec)
            assert np.array_equal(got0, want0), f"call {k} (beta 0) differs"
            prob.close()
    finally:
        dev.free()


def test_noalloc_refuses_grids_beyond_the_scratch(fm, oracle):
    print("case noalloc oversized grid", flush=True)
    dev = _Dev()
    try:
        # FP64 with averaging: 2 x 3 x 16 bytes per voxel of workspace, over 256 MiB
        prob, x, y = _case(fm, oracle, (192, 192, 96), True, 80, "fp64")
        with pytest.raises(ValueError, match="too large for cbs_operator_apply_noalloc"):
            _noalloc(fm, dev, prob, x, y, 1.0, 0.0, "fp64")
        prob.close()
    finally:
        dev.free()
