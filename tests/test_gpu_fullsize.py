"""Operator parity at the workload sizes (not only on small grids):

* 128x128x64 periodic (the mixed-precision workload grid): plane waves
  p exp(2 pi i n.m/N) are eigenfunctions of G with the closed-form eigenvalue
  Ghat(k) = (I - k k^T / k_c^2) / (|k|^2 - k_c^2) at the mapped frequency,
  Nyquist bins included -- FP64 to 1e-12, FP32 to 2e-6;
* 288x288x160 with +/- subpixel averaging (the large-phantom workload grid):
  one Green application against the oracle, FP64 to 1e-12, FP32 to 2e-6.
"""

import numpy as np
import pytest

from cases import RES, rel

pytestmark = pytest.mark.gpu


def _ghat(k, k2c):
    return (np.eye(3) - np.outer(k, k) / k2c) / (k @ k - k2c)


def _spec(fm, shape, avg, k2c, eps):
    return fm.GreenSpec(shape=shape, resolution=RES, shifts=(0.25,) * 3 if avg else (0.0,) * 3,
                        use_averaging=avg, k_square=k2c, eps_imag=eps)


def test_green_plane_waves_128x128x64(fm):
    shape = (128, 128, 64)
    print(f"case full-size plane waves {shape}", flush=True)
    k_bg = 2 * np.pi * 1.33 / 500e-9
    eps = 7.673e13
    k2c = k_bg ** 2 + 1j * eps
    V = np.asfortranarray(np.full(shape, -1j * eps, dtype=np.complex128))
    probs = {p: fm.CBSProblem(V, _spec(fm, shape, False, k2c, eps), precision=p)
             for p in ("fp64", "fp32")}
    dk = [2 * np.pi / (shape[d] * RES[d]) for d in range(3)]
    rng = np.random.default_rng(11)
    n = [np.arange(shape[d]) for d in range(3)]
    for trial in range(3):
        m = [shape[d] // 2 for d in range(3)] if trial == 0 else \
            [int(rng.integers(0, shape[d])) for d in range(3)]
        p = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        ph = [np.exp(2j * np.pi * n[d] * m[d] / shape[d]) for d in range(3)]
        s = ph[0][:, None, None] * ph[1][None, :, None] * ph[2][None, None, :]
        u = np.asfortranarray(s[..., None] * p)
        mm = [m[d] - shape[d] if 2 * m[d] >= shape[d] else m[d] for d in range(3)]
        k = np.array([dk[d] * mm[d] for d in range(3)])
        want = np.asfortranarray(s[..., None] * (_ghat(k, k2c) @ p))
        for prec, tol in (("fp64", 1e-12), ("fp32", 2e-6)):
            assert rel(probs[prec].green_apply(u), want) <= tol, (m, prec)


def test_green_averaged_288x288x160_vs_oracle(fm, oracle):
    shape = (288, 288, 160)
    print(f"case full-size averaged Green {shape}", flush=True)
    k_bg = 2 * np.pi * 1.33 / 532e-9
    eps = 1.1145e14
    k2c = k_bg ** 2 + 1j * eps
    rng = np.random.default_rng(12)
    u = np.asfortranarray(rng.standard_normal(shape + (3,)) + 1j * rng.standard_normal(shape + (3,)))
    gp = oracle.GreenParams(k_square=k2c,
                            freq_res=tuple(2 * np.pi / (shape[d] * RES[d]) for d in range(3)),
                            shifts=(0.25,) * 3, use_averaging=True)
    want = oracle.green_conv(u, gp)
    print("  oracle done", flush=True)
    V = np.asfortranarray(np.full(shape, -1j * eps, dtype=np.complex128))
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-6)):
        prob = fm.CBSProblem(V, _spec(fm, shape, True, k2c, eps), precision=prec)
        assert rel(prob.green_apply(u), want) <= tol, prec
        prob.close()
