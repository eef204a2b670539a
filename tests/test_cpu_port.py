"""The oracle's C restatement of the evaluation pass (oracle/cpu_eval.c, the CPU baseline of
bench.py): its value against the direct formula in NumPy and its adjoint against central
differences of that value (dvalue/dRe c_k = 2 Re G_k, dvalue/dIm c_k = -2 Im G_k), on random
partial waves with ragged block counts, tiny f (the clip_log Taylor branch), negative and zero
weights, in data and MC mode."""

import ctypes

import numpy as np
import pytest

from oracle import build_cpu


@pytest.fixture(scope="module")
def port():
    build_cpu.build()
    return build_cpu.load()


def _value(h, w, c, mode):
    A = np.einsum("k,kel->el", c, h)
    f = (np.abs(A) ** 2).sum(1)
    if mode == 1:
        return float((w * f).sum())
    eps = 1e-6
    t = (f - eps) / eps
    cl = np.where(f > eps, np.log(np.where(f > eps, f, 1.0)), (np.log(eps) + t) - t * t / 2)
    return float((w * cl).sum())


def _run(L, h, w, c, mode):
    P, N, LG = h.shape
    hh = np.ascontiguousarray(h).view(np.float64)
    cc = np.ascontiguousarray(c).view(np.float64)
    out = np.zeros(2 + 2 * P)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.cpu_eval(hh.ctypes.data_as(dp), np.ascontiguousarray(w).ctypes.data_as(dp), N, P, LG,
                    cc.ctypes.data_as(dp), mode, out.ctypes.data_as(dp))
    assert rc == 0
    return out[0], out[2:].view(np.complex128)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("P,N,LG", [(5, 700, 4), (3, 257, 2), (7, 1, 1)])
def test_cpu_port_value_and_adjoint(port, mode, P, N, LG):
    rng = np.random.default_rng(P * 1000 + N + LG)
    h = rng.normal(size=(P, N, LG)) + 1j * rng.normal(size=(P, N, LG))
    h[:, : N // 10] *= 1e-5  # some events below the clip_log threshold
    w = rng.uniform(-0.5, 1.5, N)
    w[::7] = 0.0
    c = rng.normal(size=P) + 1j * rng.normal(size=P)
    v, G = _run(port, h, w, c, mode)
    ref = _value(h, w, c, mode)
    assert abs(v - ref) <= 1e-11 * max(1.0, np.abs(w).sum())
    for k in range(P):
        for d, part in ((1.0, lambda g: 2 * g.real), (1j, lambda g: -2 * g.imag)):
            step = 1e-6 * max(1.0, abs(c[k]))
            cp, cm = c.copy(), c.copy()
            cp[k] += d * step
            cm[k] -= d * step
            fd = (_value(h, w, cp, mode) - _value(h, w, cm, mode)) / (2 * step)
            assert abs(part(G[k]) - fd) <= 1e-5 * max(1.0, abs(fd)), (k, d)
