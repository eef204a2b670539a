"""Device-resident event sets.

An EventSet owns one data or MC sample on the GPU: the four-momenta are uploaded once, the
helicity and aligned angles are computed on the device (the role of
tf_pwa/cal_angle.py:762-865 for FCN's data and MC, tf_pwa/model/model.py:1126-1164), and the
partial-wave cache is (re)built on the device whenever the amplitude model's parameter-free
state changes. Likelihood passes are single kernel launches over the whole set.
"""

import ctypes

import numpy as np

from . import _lib


def p4_array(p4, finals, order=None):
    """Stack four-momenta into the engine layout [f*4 + c][n] with finals in ``finals``
    order. ``p4``: dict name -> (n, 4), or an (n, n_final, 4) / (n*n_final, 4) array in
    ``order`` (the dat_order of the config, like tf_pwa/data.py:300-364)."""
    names = [p.name for p in finals]
    if isinstance(p4, dict):
        arrs = [np.asarray(p4[n], dtype=np.float64) for n in names]
    else:
        a = np.asarray(p4, dtype=np.float64)
        order = order or names
        a = a.reshape(-1, len(order), 4)
        arrs = [a[:, order.index(n), :] for n in names]
    n = arrs[0].shape[0]
    out = np.empty((len(names) * 4, n), dtype=np.float64)
    for f, a in enumerate(arrs):
        out[4 * f:4 * f + 4] = a.T
    return out


def _unpack_upper(v, Q):
    M = np.zeros((Q, Q))
    iu = np.triu_indices(Q)
    M[iu] = v
    M.T[iu] = v
    return M


def decode_second_moments(out, P):
    """(value, G, Dcc, Dcb) from a tpa_hess / tpa_pool_hess result (layout in tfpwa_amd.h):
    Dcc = sum v q q^T, Dcb = sum v q q^H + sum u h h^H from the real moments O1, O2."""
    Q = 2 * P
    T = Q * (Q + 1) // 2
    G = out[2:2 + 2 * P].view(np.complex128).copy()
    O1 = _unpack_upper(out[2 + 2 * P:2 + 2 * P + T], Q)
    O2 = _unpack_upper(out[2 + 2 * P + T:2 + 2 * P + 2 * T], Q)
    a, b = slice(0, P), slice(P, Q)
    Dcc = (O1[a, a] - O1[b, b]) + 1j * (O1[a, b] + O1[b, a])
    Dcb = ((O1[a, a] + O1[b, b]) + 1j * (O1[b, a] - O1[a, b])
           + (O2[a, a] + O2[b, b]) + 1j * (O2[b, a] - O2[a, b]))
    return out[0], G, Dcc, Dcb


class EventSet:
    """One sample on the device. ``amp`` is the AmplitudeModel whose description the
    kernels use; ``weight`` defaults to 1. ``has_weight`` says whether weights were given (a
    sample dict's "weight" entry in the reference: background sets without one take -w_bkg
    in Model.get_weight_data)."""

    def __init__(self, amp, p4, weight=None, order=None):
        self.amp = amp
        self.has_weight = weight is not None
        arr = p4_array(p4, amp.finals, order)
        self.n = arr.shape[1]
        self.weight = np.ones(self.n) if weight is None else np.ascontiguousarray(
            weight, dtype=np.float64).copy()
        if self.weight.shape != (self.n,):
            raise ValueError("weight must have one entry per event")
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.tpa_events_create(amp.device(), _lib.dptr(np.ascontiguousarray(arr)),
                                       self.n, _lib.dptr(self.weight), ctypes.byref(h)))
        self._h = h
        self._dev_weight = self.weight.copy()
        self._weight_version = 0
        self._state = None

    @classmethod
    def _wrap(cls, amp, h, n, weight):
        self = cls.__new__(cls)
        self.amp = amp
        self.has_weight = weight is not None
        self.n = int(n)
        self.weight = np.ones(self.n) if weight is None else np.array(weight, dtype=np.float64)
        self._h = h
        self._dev_weight = self.weight.copy()
        self._weight_version = 0
        self._state = None
        return self

    @classmethod
    def generate(cls, amp, n, seed, weight=None, first=0, return_p4=False):
        """n flat phase-space events generated on the device (tpa_events_generate): the
        counter-based sample of (seed, first + index). With ``return_p4`` also returns the
        four-momenta [n_final * 4][n] in the engine's final order."""
        masses = np.array([p.mass for p in amp.finals], dtype=np.float64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        p4 = np.empty((4 * len(masses), n)) if return_p4 else None
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tpa_events_generate(
            amp.device(), float(amp.decay_group.top.mass), _lib.dptr(masses), int(n),
            int(seed) & (2 ** 64 - 1), int(first), _lib.dptr(w), _lib.dptr(p4), ctypes.byref(h)))
        ev = cls._wrap(amp, h, n, w)
        return (ev, p4) if return_p4 else ev

    def select(self, idx, weight=None):
        """A new event set of the events ``idx`` of this one (their angles)."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tpa_events_select(self._h, idx.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int64)), idx.size, _lib.dptr(w), ctypes.byref(h)))
        return EventSet._wrap(self.amp, h, idx.size, w)

    @classmethod
    def concat(cls, sets):
        """One event set holding the events of ``sets`` in order, with their weights
        (tpa_events_concat; tf_pwa/data.py:527-540 data_merge): one pass per evaluation
        instead of one per set."""
        sets = list(sets)
        if not sets:
            raise ValueError("no event sets to concatenate")
        amp = sets[0].amp
        if any(s.amp is not amp for s in sets):
            raise ValueError("event sets of different amplitude models")
        handles = (ctypes.c_void_p * len(sets))(*[s._h.value for s in sets])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tpa_events_concat(handles, len(sets), ctypes.byref(h)))
        w = np.concatenate([s._dev_weight for s in sets])
        ev = cls._wrap(amp, h, w.size, w)
        ev.has_weight = any(s.has_weight for s in sets)
        return ev

    def __len__(self):
        return self.n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tpa_events_destroy(h)
            self._h = None

    def set_weight(self, w):
        if w is getattr(self, "_src_weight", None):
            return
        src = w
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.shape != (self.n,):
            raise ValueError("weight must have one entry per event")
        if not np.array_equal(w, self._dev_weight):
            _lib.check(_lib.lib().tpa_events_set_weight(self._h, _lib.dptr(w)))
            self._dev_weight = w.copy()
            self._weight_version += 1
        self._src_weight = src

    def angles(self):
        if self.amp.n_final == 4:
            nf = _lib.NF4_TOPO * len(self.amp.topo) + 5
        else:
            nf = _lib.NF_TOPO * len(self.amp.topo) + 4
        out = np.empty((nf, self.n), dtype=np.float64)
        _lib.check(_lib.lib().tpa_events_get_angles(self._h, _lib.dptr(out), out.size))
        return out

    def ensure_built(self):
        self.amp.device()
        state = self.amp._dev_state
        if self._state == state:
            return
        _lib.check(_lib.lib().tpa_events_build(self._h, 1 if state[0] else 0))
        self._state = state

    def partial_waves(self):
        self.ensure_built()
        out = np.empty(2 * self.amp.P * self.n * self.amp.lam_pad, dtype=np.float64)
        _lib.check(_lib.lib().tpa_events_get_pw(self._h, _lib.dptr(out), out.size))
        return out.view(np.complex128).reshape(self.amp.P, self.n, self.amp.lam_pad)

    def keep_f32(self, on=True):
        """Keep an FP32 copy of the (folded) partial-wave cache beside the FP64 one
        (tpa_events_keep_f32), refreshed by every rebuild, for passes with ``f32=True``."""
        _lib.check(_lib.lib().tpa_events_keep_f32(self._h, 1 if on else 0))
        self._keep32 = bool(on)

    def keep_half(self, on=True, fp64=False):
        """Keep the helicity-halved copies of the (folded) cache (tpa_events_keep_half): one
        helicity slot of each mirror pair and one lineshape phase per chain and event, in FP32
        (``on``, for passes with ``half=True``) and/or FP64 (``fp64``, for ``half64=True``),
        refreshed by every rebuild. Making one checks the helicity-flip relation on every event
        (the residual is ``half_residual()``)."""
        self.ensure_built()
        _lib.check(_lib.lib().tpa_events_keep_half(self._h, (1 if on else 0) | (2 if fp64 else 0)))
        self._keep32h = bool(on)
        self._keep64h = bool(fp64)

    def half_copy(self):
        """The helicity-halved copy as the host sees it: (g [P, n, nh] complex64, rho
        [n_chain, n] complex64, slots, mirrors, signs, phi_rot [P], chain of each wave [P]),
        with h_k,o = phi_k g_k,o and h_k,-o = phi_k s_o rho conj(g_k,o), phi = -i where
        phi_rot is 1."""
        import ctypes
        P = self.amp.P
        nc = self.amp.n_chain if hasattr(self.amp, "n_chain") else None
        info = np.zeros(13 + 2 * P + 1 + 32, dtype=np.int32)
        n_chain_max = 64
        g = np.zeros(2 * P * self.n * 4, dtype=np.float32)
        rho = np.zeros(2 * n_chain_max * self.n, dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _lib.check(_lib.lib().tpa_events_get_half(
            self._h, g.ctypes.data_as(fp), g.size, rho.ctypes.data_as(fp), rho.size,
            info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), info.size))
        nh = int(info[0])
        kch = info[13 + P:13 + 2 * P].copy()
        n_chain = int(kch.max()) + 1 if nc is None else nc
        g = g[:2 * P * self.n * nh].view(np.complex64).reshape(P, self.n, nh)
        rho = rho[:2 * n_chain * self.n].view(np.complex64).reshape(n_chain, self.n)
        self._half_plan = (int(info[13 + 2 * P]), info[14 + 2 * P:14 + 2 * P + 32].copy())
        return (g, rho, info[1:1 + nh].copy(), info[5:5 + nh].copy(), info[9:9 + nh].copy(),
                info[13:13 + P].copy(), kch)

    def half_plan(self):
        """The lane plan of the chain-grouped FP32 pass over the halved copy: (waves per lane,
        the wave of each lane slot, -1 for pad), or (0, ...) when no plan fits and the pass is
        the per-wave one. Read with the copy (half_copy)."""
        self.half_copy()
        return self._half_plan

    def half_residual(self):
        """Largest relative residual of the helicity-flip relation measured by the last
        halved copy (-1 if none was made)."""
        r = np.zeros(1)
        _lib.check(_lib.lib().tpa_events_half_residual(self._h, _lib.dptr(r)))
        return float(r[0])

    def eval(self, coef, mode, grad=False, fout=False, f32=False, half=False, half64=False):
        """One pass: (value, G or None, f or None). When the model floats lineshape
        parameters, G is the tuple (G, Dm, Dg) of coupling and lineshape adjoints. ``f32``:
        read the FP32 copy of the cache (keep_f32; values widened to FP64 on load); ``half``:
        read the helicity-halved FP32 copy (keep_half); ``half64``: the halved FP64 copy."""
        self.ensure_built()
        P = self.amp.P
        shape = not self._state[0]
        if f32 and not getattr(self, "_keep32", False):
            self.keep_f32(True)
        if (half and not getattr(self, "_keep32h", False)) or (
                half64 and not getattr(self, "_keep64h", False)):
            self.keep_half(half or getattr(self, "_keep32h", False),
                           half64 or getattr(self, "_keep64h", False))
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
        out = np.zeros(2 + (6 if shape else 2) * P, dtype=np.float64)
        f = np.empty(self.n, dtype=np.float64) if fout else None
        flags = ((_lib.FLAG_GRAD if grad else 0) | (_lib.FLAG_FOUT if fout else 0)
                 | (_lib.FLAG_SHAPE if shape else 0) | (_lib.FLAG_F32 if f32 else 0)
                 | (_lib.FLAG_F32H if half else 0) | (_lib.FLAG_F64H if half64 else 0))
        _lib.check(_lib.lib().tpa_eval(self._h, _lib.dptr(c), mode, flags, _lib.dptr(out),
                                       _lib.dptr(f)))
        G = None
        if grad:
            z = out[2:].view(np.complex128)
            G = (z[:P].copy(), z[P:2 * P].copy(), z[2 * P:].copy()) if shape else z.copy()
        return out[0], G, f

    def eval_recompute(self, coef, mode, grad=False, chunk=0):
        """One pass with the partial waves rebuilt from the angle cache on this call, ``chunk``
        events at a time (tpa_eval_recompute; 0: the library's default): the event set's own
        cache is neither built nor read. (value, G or None)."""
        self.amp.device()
        P = self.amp.P
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
        out = np.zeros(2 + 2 * P, dtype=np.float64)
        _lib.check(_lib.lib().tpa_eval_recompute(self._h, _lib.dptr(c), mode,
                                                 _lib.FLAG_GRAD if grad else 0, int(chunk),
                                                 _lib.dptr(out)))
        return out[0], (out[2:].view(np.complex128).copy() if grad else None)

    def eval_fused(self, coef, mode, grad=False):
        """One pass of the fused kernel (tpa_eval_fused): the waves recomputed from the angle
        cache on every call, no partial-wave cache. (value, G or None)."""
        self.amp.device()
        P = self.amp.P
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
        out = np.zeros(2 + 2 * P, dtype=np.float64)
        _lib.check(_lib.lib().tpa_eval_fused(self._h, _lib.dptr(c), mode,
                                             _lib.FLAG_GRAD if grad else 0, _lib.dptr(out)))
        return out[0], (out[2:].view(np.complex128).copy() if grad else None)

    def hess(self, coef, mode):
        """Second-moment pass: (value, G, Dcc, Dcb) with G_k = d sum / dc_k,
        Dcc_kl = d2 sum / dc_k dc_l and Dcb_kl = d2 sum / dc_k dconj(c_l) (complex [P, P])."""
        self.ensure_built()
        if not self._state[0]:
            raise NotImplementedError("second moments need fixed lineshapes (fold = 1 cache)")
        P = self.amp.P
        Q = 2 * P
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
        out = np.zeros(2 + 2 * P + Q * (Q + 1), dtype=np.float64)
        _lib.check(_lib.lib().tpa_hess(self._h, _lib.dptr(c), mode, _lib.dptr(out)))
        return decode_second_moments(out, P)

    def hess_shape(self, coef, mode):
        """Second-moment pass with per-event lineshapes (tpa_hess_shape, fold-0 cache):
        (value, F, F2, Fcc, Fcb) over the extended waves X_(t,k) = b^t h_k (b^t the wave's
        lineshape B, dB/dm0, dB/dg0): F_j = d sum / d alpha_j (3P), F2 the moments of the
        second lineshape derivatives (3P: d2/dm0^2, d2/dm0dg0, d2/dg0^2 blocks), Fcc and Fcb
        the Wirtinger second derivatives in the extended coefficients ([3P, 3P])."""
        self.ensure_built()
        if self._state[0]:
            raise ValueError("lineshape second moments need a fold-0 cache (floating shapes)")
        P = self.amp.P
        P3, Q = 3 * P, 6 * P
        T = Q * (Q + 1) // 2
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
        out = np.zeros(2 + 12 * P + 2 * T, dtype=np.float64)
        _lib.check(_lib.lib().tpa_hess_shape(self._h, _lib.dptr(c), mode, _lib.dptr(out)))
        F = out[2:2 + 6 * P].view(np.complex128).copy()
        F2 = out[2 + 6 * P:2 + 12 * P].view(np.complex128).copy()
        O1 = _unpack_upper(out[2 + 12 * P:2 + 12 * P + T], Q)
        O2 = _unpack_upper(out[2 + 12 * P + T:2 + 12 * P + 2 * T], Q)
        a, b = slice(0, P3), slice(P3, Q)
        Fcc = (O1[a, a] - O1[b, b]) + 1j * (O1[a, b] + O1[b, a])
        Fcb = ((O1[a, a] + O1[b, b]) + 1j * (O1[b, a] - O1[a, b])
               + (O2[a, a] + O2[b, b]) + 1j * (O2[b, a] - O2[a, b]))
        return out[0], F, F2, Fcc, Fcb

    def gram(self, precision="fp64"):
        """Partial-wave Gram matrix I_kl = sum_e w_e sum_lambda conj(h_k) h_l, cached until the
        partial waves or the weights change. precision "fp32": FP32 MFMA with FP64 refolds (its
        partial sums drift by up to ~1e-8 relative for narrow peaks over a broad MC sample);
        "fp64": FP64 MFMA."""
        if precision not in ("fp32", "fp64"):
            raise ValueError("gram precision must be 'fp32' or 'fp64'")
        self.ensure_built()
        key = (self._state, self._weight_version, precision)
        if getattr(self, "_gram_key", None) != key:
            P = self.amp.P
            out = np.zeros(2 * P * P, dtype=np.float64)
            ms = ctypes.c_double(0)
            if precision == "fp64":
                _lib.check(_lib.lib().tpa_gram_f64(self._h, _lib.dptr(out), ctypes.byref(ms)))
            else:
                _lib.check(_lib.lib().tpa_gram(self._h, float(np.mean(self._dev_weight)),
                                               _lib.dptr(out), ctypes.byref(ms)))
            self._gram = out.view(np.complex128).reshape(P, P).copy()
            self._gram_key = key
            self.gram_ms = ms.value
        return self._gram

    def amp2(self, amp=None):
        _, _, f = self.eval(self.amp.coefficients(), _lib.MODE_MC, fout=True)
        return f
