"""tpa_eval_noalloc (the allocation-free pass over a static device scratch array) against
tpa_eval on the same event sets and coefficients: equal bit for bit, in both modes, with and
without the adjoint sums, over the FP64 cache and its FP32 copy, on two calls in a row of which
the second covers fewer events (the scratch keeps the first call's partial sums, and the second
must not see them). The output buffer is device memory the test allocates itself through HIP,
as a caller that owns its memory would; the engine allocates nothing in these calls."""

import ctypes

import numpy as np
import pytest

from helpers import make_case

pytestmark = pytest.mark.gpu

_HIP_D2H = 2


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMalloc.restype = ctypes.c_int
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipFree.restype = ctypes.c_int
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemcpy.restype = ctypes.c_int
    return h


class _DeviceOut:
    """A caller-owned device buffer of n doubles."""

    def __init__(self, n):
        self.hip = _hip()
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), 8 * n) == 0

    def read(self, k):
        from tf_pwa_amd import _lib

        _lib.check(_lib.lib().tpa_device_synchronize())
        out = np.empty(k, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, 8 * k, _HIP_D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _noalloc(ev, coef, mode, flags, dev):
    from tf_pwa_amd import _lib

    c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
    _lib.check(_lib.lib().tpa_eval_noalloc(ev._h, _lib.dptr(c), mode, flags, dev.ptr, None))
    P = ev.amp.P
    return dev.read(2 + 2 * P if flags & _lib.FLAG_GRAD else 2)


def _reference(ev, coef, mode, flags):
    """tpa_eval's raw output for the same flags (value pair, then G as re/im pairs)."""
    from tf_pwa_amd import _lib

    c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
    P = ev.amp.P
    n = 2 + 2 * P if flags & _lib.FLAG_GRAD else 2
    out = np.zeros(2 + 2 * P, dtype=np.float64)
    _lib.check(_lib.lib().tpa_eval(ev._h, _lib.dptr(c), mode, flags, _lib.dptr(out), None))
    return out[:n]


@pytest.mark.parametrize("cfg", ["lc_pkpi_6res.yml", "lc_pkpi_2res.yml"])
def test_noalloc_pass_equals_tpa_eval(cfg, gpu_available):
    from tf_pwa_amd import _lib

    case = make_case(cfg, n_mc=30000, n_data=600, seed=41)
    conf = case["conf"]
    amp = conf.get_amplitude()
    big = conf.data_from_p4(case["mc"])
    small = conf.data_from_p4(case["data"], weight=np.linspace(0.5, 1.5, 600))
    coef = amp.coefficients()
    dev = _DeviceOut(2 + 2 * 32)
    try:
        for ev in (big, small):  # the second call covers fewer events than the first
            ev.ensure_built()
            ev.keep_f32(True)
            for mode in (_lib.MODE_DATA, _lib.MODE_MC):
                for flags in (_lib.FLAG_GRAD, 0, _lib.FLAG_GRAD | _lib.FLAG_F32):
                    got = _noalloc(ev, coef, mode, flags, dev)
                    ref = _reference(ev, coef, mode, flags)
                    assert np.array_equal(got, ref), (cfg, len(ev), mode, flags, got[:4], ref[:4])
                    # and through the engine's own path it is the value of the eval
                    v, _, _ = ev.eval(coef, mode, grad=bool(flags & _lib.FLAG_GRAD),
                                      f32=bool(flags & _lib.FLAG_F32))
                    assert got[0] == v
    finally:
        dev.free()


def test_noalloc_rejects_what_it_cannot_serve(gpu_available):
    from tf_pwa_amd import _lib

    case = make_case("lc_pkpi_2res.yml", n_mc=500, n_data=100, seed=43)
    conf = case["conf"]
    amp = conf.get_amplitude()
    ev = conf.data_from_p4(case["mc"])
    ev.ensure_built()
    coef = amp.coefficients()
    dev = _DeviceOut(2 + 2 * 32)
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
    L = _lib.lib()
    try:
        for flags in (_lib.FLAG_FOUT | _lib.FLAG_GRAD, _lib.FLAG_SHAPE):
            assert L.tpa_eval_noalloc(ev._h, _lib.dptr(c), _lib.MODE_MC, flags, dev.ptr, None) != 0
            assert b"TPA_FLAG_GRAD and TPA_FLAG_F32 only" in L.tpa_last_error()
        assert L.tpa_eval_noalloc(ev._h, _lib.dptr(c), 7, 0, dev.ptr, None) != 0
        assert L.tpa_eval_noalloc(ev._h, _lib.dptr(c), _lib.MODE_MC, 0, None, None) != 0
        # a set without the FP32 copy
        assert L.tpa_eval_noalloc(ev._h, _lib.dptr(c), _lib.MODE_MC, _lib.FLAG_F32, dev.ptr,
                                  None) != 0
        assert b"FP32 copy" in L.tpa_last_error()
    finally:
        dev.free()
