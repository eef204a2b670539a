"""Host logic of the serving runtime (tpa_serve.cpp) on the CPU: the server is linked against
a CPU stand-in of the device pool (tests/cpu_pool/mock_pool.cpp) that evaluates the same
per-request data sums from a synthetic partial-wave cache, so the coefficient map, the NLL and
gradient assembly, the BFGS, the batcher and the thread lifecycle run without a GPU.
"""

import ctypes
import os
import subprocess
import time

import numpy as np
import pytest

from conftest import ROOT

P, LG = 6, 2


@pytest.fixture(scope="module")
def mocklib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mock") / "libmockserve.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "tf-pwa_amd", "csrc"), "-o", out,
                    os.path.join(ROOT, "tf-pwa_amd", "csrc", "tpa_serve.cpp"),
                    os.path.join(ROOT, "tests", "cpu_pool", "mock_pool.cpp")], check=True)
    from tf_pwa_amd import _lib

    L = ctypes.CDLL(out)
    vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "mock_pool_create": [ctypes.c_int, ctypes.c_int, i64, ctypes.POINTER(vp)],
        "mock_pool_destroy": [vp],
        "mock_pool_add": [vp, dp, i64, dp, ctypes.POINTER(i64)],
        "tpa_server_create": [vp, ctypes.POINTER(_lib.TpaCouplingMap), dp, ctypes.c_double,
                              ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                              ctypes.POINTER(vp)],
        "tpa_server_destroy": [vp],
        "tpa_server_nll_grad": [vp, i64, i64, ctypes.c_double, dp, dp, dp],
        "tpa_server_nll_grad_hess": [vp, i64, i64, ctypes.c_double, dp, dp, dp, dp],
        "tpa_server_submit": [vp, i64, i64, ctypes.c_double, dp, ip],
        "tpa_server_poll": [vp, ip, ctypes.c_int, ip],
        "tpa_server_result": [vp, ctypes.c_int, dp, dp, ip, ctypes.POINTER(i64),
                              ctypes.POINTER(i64)],
        "tpa_server_counters": [vp, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                ctypes.POINTER(i64), dp],
    }
    for name, args in sig.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ctypes.c_int
    L.tpa_last_error.restype = ctypes.c_char_p
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Problem:
    """Synthetic partial waves: 3 chains x 2 waves, c_k = total_chain * g_k * const_k with
    polar factors; chain 0's total is fixed at 1. Variable 0 (a magnitude) is bounded."""

    def __init__(self, L, sizes, seed=0):
        rng = np.random.default_rng(seed)
        self.L = L
        n_all = sum(sizes)
        self.h = (rng.normal(size=(P, n_all, LG)) + 1j * rng.normal(size=(P, n_all, LG)))
        self.h[1] *= 0.4
        hm = rng.normal(size=(P, 20000, LG)) + 1j * rng.normal(size=(P, 20000, LG))
        self.gram = np.einsum("kel,mel->km", hm.conj(), hm) / 20000
        # factors: 0..2 chain totals, 3..8 wave couplings g_k, 9 a fixed cartesian factor,
        # 10 the fixed unit factor of the unused fourth slot
        self.nf = 11
        self.wave_factor = np.array([[k // 2, 3 + k, 9, 10] for k in range(P)], dtype=np.int32)
        self.wave_const = np.linspace(0.8, 1.3, P)
        self.factor_polar = np.array([1] * 9 + [0, 0], dtype=np.int32)
        fv = -np.ones((self.nf, 2), dtype=np.int32)
        nv = 0
        for f in (1, 2, 4, 5, 6, 7, 8):  # factor 0 (chain 0 total) and 3 (g_0) fixed
            fv[f] = (nv, nv + 1)
            nv += 2
        self.factor_var = fv
        self.factor_fixed = np.zeros((self.nf, 2))
        self.factor_fixed[:, 0] = 1.0
        self.factor_fixed[9] = (0.9, 0.2)
        self.factor_fixed[10] = (1.0, 0.0)
        self.nv = nv
        self.bound_kind = np.zeros(nv, dtype=np.int32)
        self.bound_ab = np.zeros((nv, 2))
        self.bound_kind[0] = 1
        self.bound_ab[0] = (0.1, 3.0)
        self.x_true = np.array([1.2, 0.4, 0.8, -1.0, 0.7, 0.3, 1.1, -0.5, 0.6, 2.0, 0.9, 1.4,
                                1.3, -2.2])
        from tf_pwa_amd import _lib

        i32p = ctypes.POINTER(ctypes.c_int32)
        self.map = _lib.TpaCouplingMap(
            P=P, n_factor=self.nf, n_var=nv, reserved=0,
            wave_factor=self.wave_factor.ctypes.data_as(i32p), wave_const=_dp(self.wave_const),
            factor_polar=self.factor_polar.ctypes.data_as(i32p),
            factor_var=self.factor_var.ctypes.data_as(i32p), factor_fixed=_dp(self.factor_fixed),
            bound_kind=self.bound_kind.ctypes.data_as(i32p), bound_ab=_dp(self.bound_ab))
        self.pool = ctypes.c_void_p()
        assert L.mock_pool_create(P, LG, n_all, ctypes.byref(self.pool)) == 0
        self.slots, a = [], 0
        for n in sizes:
            hs = np.ascontiguousarray(self.h[:, a:a + n, :]).view(np.float64).ravel()
            w = np.ones(n)
            off = ctypes.c_int64()
            assert L.mock_pool_add(self.pool, _dp(hs), n, _dp(w), ctypes.byref(off)) == 0
            self.slots.append((off.value, n))
            a += n

    def server(self, **kw):
        g = np.ascontiguousarray(self.gram).view(np.float64).ravel()
        s = ctypes.c_void_p()
        opts = dict(gtol=1e-5, maxiter=0, max_batch=64, window_us=200.0, hess_start=0)
        opts.update(kw)
        assert self.L.tpa_server_create(self.pool, ctypes.byref(self.map), _dp(g), opts["gtol"],
                                        opts["maxiter"], opts["max_batch"], opts["window_us"],
                                        opts["hess_start"], ctypes.byref(s)) == 0, \
            self.L.tpa_last_error()
        return s

    def coefs(self, y):
        fv = self.factor_var
        r = np.where(fv[:, 0] >= 0, y[np.maximum(fv[:, 0], 0)], self.factor_fixed[:, 0])
        i = np.where(fv[:, 1] >= 0, y[np.maximum(fv[:, 1], 0)], self.factor_fixed[:, 1])
        v = np.where(self.factor_polar == 1, r * np.exp(1j * i), r + 1j * i)
        wf = self.wave_factor
        return self.wave_const * v[wf[:, 0]] * v[wf[:, 1]] * v[wf[:, 2]] * v[wf[:, 3]]

    def nll(self, slot, y):
        """numpy restatement: -sum log f + n log(c^dagger I c) (unit weights, f >> 1e-6)."""
        off, n = slot
        c = self.coefs(y)
        A = np.einsum("k,kel->el", c, self.h[:, off:off + n, :])
        f = (np.abs(A) ** 2).sum(1)
        return -np.log(f).sum() + n * np.log(np.real(np.vdot(c, self.gram @ c)))


def _nll_grad(L, s, slot, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    g = np.zeros(y.size)
    f = ctypes.c_double()
    rc = L.tpa_server_nll_grad(s, slot[0], slot[1], float(slot[1]), _dp(y), ctypes.byref(f),
                               _dp(g))
    assert rc == 0, L.tpa_last_error()
    return f.value, g


def _wait(L, s, n, timeout=60):
    ids = (ctypes.c_int * 64)()
    nd = ctypes.c_int()
    got, t0 = [], time.time()
    while len(got) < n:
        assert L.tpa_server_poll(s, ids, 64, ctypes.byref(nd)) == 0
        got += list(ids[:nd.value])
        assert time.time() - t0 < timeout
        time.sleep(0.001)
    return got


def _result(L, s, fid, nv):
    y = np.zeros(nv)
    f = ctypes.c_double()
    st = ctypes.c_int()
    nfev, nit = ctypes.c_int64(), ctypes.c_int64()
    assert L.tpa_server_result(s, fid, _dp(y), ctypes.byref(f), ctypes.byref(st),
                               ctypes.byref(nfev), ctypes.byref(nit)) == 0
    return y, f.value, st.value, nfev.value, nit.value


def test_server_nll_and_gradient(mocklib):
    pb = Problem(mocklib, [1500, 900])
    s = pb.server()
    rng = np.random.default_rng(1)
    for _ in range(3):
        y = pb.x_true + rng.normal(scale=0.2, size=pb.nv)
        for slot in pb.slots:
            f, g = _nll_grad(mocklib, s, slot, y)
            ref = pb.nll(slot, y)
            assert abs(f - ref) <= 1e-10 * abs(ref)
            for j in range(pb.nv):
                e = np.zeros(pb.nv)
                e[j] = 1e-6
                fd = (pb.nll(slot, y + e) - pb.nll(slot, y - e)) / 2e-6
                assert abs(fd - g[j]) <= 1e-5 * max(1.0, np.abs(g).max()), (j, fd, g[j])
    mocklib.tpa_server_destroy(s)


def test_server_exact_hessian(mocklib):
    """The server's Hessian (second moments + chain rule in C++) against central differences of
    its own gradient, and its NLL/gradient against the batched path."""
    pb = Problem(mocklib, [1200])
    s = pb.server()
    slot = pb.slots[0]
    rng = np.random.default_rng(5)
    y = np.ascontiguousarray(pb.x_true + rng.normal(scale=0.2, size=pb.nv))
    g = np.zeros(pb.nv)
    H = np.zeros((pb.nv, pb.nv))
    f = ctypes.c_double()
    assert mocklib.tpa_server_nll_grad_hess(s, slot[0], slot[1], float(slot[1]), _dp(y),
                                            ctypes.byref(f), _dp(g), _dp(H)) == 0
    f2, g2 = _nll_grad(mocklib, s, slot, y)
    assert abs(f.value - f2) <= 1e-12 * abs(f2)
    np.testing.assert_allclose(g, g2, rtol=1e-10, atol=1e-10 * np.abs(g2).max())
    h = 1e-6
    for j in range(pb.nv):
        e = np.zeros(pb.nv)
        e[j] = h
        col = (_nll_grad(mocklib, s, slot, y + e)[1] - _nll_grad(mocklib, s, slot, y - e)[1]) / (
            2 * h)
        np.testing.assert_allclose(H[:, j], col, rtol=0, atol=1e-5 * np.abs(H).max())
    assert np.array_equal(H, H.T)
    mocklib.tpa_server_destroy(s)


@pytest.mark.parametrize("hess_start", [0, 1])
def test_server_fits_converge_and_ignore_the_mix(mocklib, hess_start):
    pb = Problem(mocklib, [2000, 1200, 1700, 800])
    rng = np.random.default_rng(2)
    starts = [pb.x_true + rng.normal(scale=0.1, size=pb.nv) for _ in pb.slots]
    s = pb.server(hess_start=hess_start)
    fid = ctypes.c_int()
    y0 = np.ascontiguousarray(starts[0])
    slot = pb.slots[0]
    assert mocklib.tpa_server_submit(s, slot[0], slot[1], float(slot[1]), _dp(y0),
                                     ctypes.byref(fid)) == 0
    _wait(mocklib, s, 1)
    alone = _result(mocklib, s, fid.value, pb.nv)
    ids = []
    for slot, y0 in zip(pb.slots, starts):
        y0 = np.ascontiguousarray(y0)
        assert mocklib.tpa_server_submit(s, slot[0], slot[1], float(slot[1]), _dp(y0),
                                         ctypes.byref(fid)) == 0
        ids.append(fid.value)
    _wait(mocklib, s, len(ids))
    mix = _result(mocklib, s, ids[0], pb.nv)
    assert np.array_equal(alone[0], mix[0]) and alone[1:] == mix[1:]
    for fid_, slot, y0 in zip(ids, pb.slots, starts):
        y, f, st, nfev, nit = _result(mocklib, s, fid_, pb.nv)
        assert st in (0, 2), st
        assert f < pb.nll(slot, y0)
        assert abs(f - pb.nll(slot, y)) <= 1e-9 * abs(f)
        _, g = _nll_grad(mocklib, s, slot, y)
        assert np.abs(g).max() < 1e-3
        assert 0.1 <= y[0] <= 3.0  # the bounded magnitude stays inside its bounds
    r, b, ev = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    ms = ctypes.c_double()
    mocklib.tpa_server_counters(s, ctypes.byref(r), ctypes.byref(b), ctypes.byref(ev),
                                ctypes.byref(ms))
    assert b.value < r.value  # batches were shared between fits
    assert ev.value > r.value
    mocklib.tpa_server_destroy(s)


def test_server_destroy_cancels_running_fits(mocklib):
    pb = Problem(mocklib, [3000])
    s = pb.server(maxiter=100000, gtol=1e-300)  # would run for a long time
    fid = ctypes.c_int()
    y0 = np.ascontiguousarray(pb.x_true + 0.3)
    for _ in range(4):
        assert mocklib.tpa_server_submit(s, 0, 3000, 3000.0, _dp(y0), ctypes.byref(fid)) == 0
    time.sleep(0.05)
    t0 = time.time()
    assert mocklib.tpa_server_destroy(s) == 0
    assert time.time() - t0 < 10


def test_server_rejects_bad_maps(mocklib):
    pb = Problem(mocklib, [100])
    pb.wave_factor[0, 0] = 99
    g = np.ascontiguousarray(pb.gram).view(np.float64).ravel()
    s = ctypes.c_void_p()
    assert mocklib.tpa_server_create(pb.pool, ctypes.byref(pb.map), _dp(g), 1e-5, 0, 8, 100.0,
                                     0, ctypes.byref(s)) != 0
    assert b"out of range" in mocklib.tpa_last_error()
