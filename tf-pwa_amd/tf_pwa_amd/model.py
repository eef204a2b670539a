"""Likelihood surface of tf_pwa.model (Model, FCN) on top of the engine.

NLL = -sum_i w_i clip_log(f_i) + (sum_i w_i) log(sum_j w~_j f_j) and its gradient
(tf_pwa/model/model.py:407-449): data weights carry the alpha = sum w / sum w^2 rescale
(get_weight_data, 603-641) and MC weights are normalised to unit sum (FCN.__init__,
1126-1164). Each sum is one engine pass per event set; the gradient with respect to the
trainable variables is assembled on the host from the engine's adjoint sums G_k through
the coupling Jacobian: d sum / d theta = 2 Re sum_k (d c_k / d theta) G_k.
"""

import time

import numpy as np

from . import _lib

_MODELS = {}


def register_nll_model(name):
    """Registry of NLL models selected by ``data: model:`` (tf_pwa/model/model.py:38)."""

    def reg(cls):
        _MODELS[name] = cls
        return cls

    return reg


def get_nll_model(name):
    return _MODELS[name]


class WeightedData:
    """A list of event sets with one weight array over their concatenation (the merged
    data + background of get_weight_data)."""

    def __init__(self, sets, weight):
        self.sets = list(sets)
        self.weight = np.asarray(weight, dtype=np.float64)
        n = sum(s.n for s in self.sets)
        if self.weight.shape != (n,):
            raise ValueError("weight length {} != events {}".format(self.weight.shape, n))

    def split(self):
        out, i = [], 0
        for s in self.sets:
            out.append(self.weight[i:i + s.n])
            i += s.n
        return out


def _as_sets(x):
    if x is None:
        return []
    if isinstance(x, WeightedData):
        return x.sets
    if isinstance(x, (list, tuple)):
        return list(x)
    return [x]


@register_nll_model("default")
class Model:
    """Model.nll / nll_grad / nll_grad_batch / nll_grad_hessian (model.py:576-790)."""

    def __init__(self, amp, w_bkg=1.0, resolution_size=1, extended=False, precision="fp64",
                 gram_precision="fp64", data_storage="fp64", evaluation="stream",
                 mc_storage="fp64", **kwargs):
        """precision: "fp64" evaluates every MC event per call; "mixed" replaces the MC pass
        by the partial-wave Gram matrix built once on MFMA (valid while the partial waves,
        i.e. lineshape parameters and MC weights, are unchanged; rebuilt otherwise).
        gram_precision: the Gram's arithmetic in mixed precision, "fp64" (FP64 MFMA, the
        default: the Gram is built once per lineshape state, so its cost is off the per-call
        path) or "fp32" (FP32 MFMA with FP64 refolds; its partial sums drift by up to ~1e-8
        relative for narrow peaks over a broad MC sample).
        data_storage: the data partial-wave cache the mixed path reads, "fp64", "fp32"
        (an FP32 copy kept beside the FP64 cache, values widened to FP64 on load, every sum
        in FP64: half the bytes per data event) or "fp32h" (the helicity-halved FP32 copy:
        one helicity slot of each mirror pair plus one lineshape phase per chain and event,
        from which the mirrored slots follow; 408 instead of 712 B per event at six
        resonances); the FP32 options with fixed lineshapes, else FP64. "fp64h": the same
        halved layout in FP64 (the FP64 cache's values to rounding, 808 instead of 1416 B per
        event), for the data passes of both precisions.
        mc_storage: the MC cache the FP64 path's MC passes read, "fp64" or "fp64h" (the halved
        FP64 copy); fixed lineshapes, else FP64.
        evaluation: how a pass gets its partial waves, "stream" (read the cache built once,
        k_eval_pipe / k_eval_pipe32), "fused" (recompute every wave from the angle cache per
        event in one generated kernel, tpa_eval_fused: no partial-wave cache; three-body,
        fixed lineshapes, at most 8 waves) or "recompute" (rebuild the waves from the angle
        cache on every call in chunks of ``recompute_chunk`` events, tpa_eval_recompute: only
        the angles stay resident; three- and four-body, fixed lineshapes)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if gram_precision not in ("fp32", "fp64"):
            raise ValueError("gram_precision must be 'fp32' or 'fp64'")
        if data_storage not in ("fp32", "fp32h", "fp64", "fp64h"):
            raise ValueError("data_storage must be 'fp32', 'fp32h', 'fp64' or 'fp64h'")
        if mc_storage not in ("fp64", "fp64h"):
            raise ValueError("mc_storage must be 'fp64' or 'fp64h'")
        self.mc_storage = mc_storage
        self.recompute_chunk = int(kwargs.pop("recompute_chunk", 0))
        self.Amp = amp
        self.evaluation = evaluation
        self.precision = precision
        self.gram_precision = gram_precision
        self.data_storage = data_storage
        if resolution_size != 1:
            raise NotImplementedError("resolution folding is not supported")
        if extended:
            raise NotImplementedError("extended likelihood is not supported")
        self.vm = amp.vm
        self.w_bkg = w_bkg
        self.resolution_size = 1
        self.n_eval = 0

    def get_weight_data(self, data, weight=None, bg=None, alpha=True):
        """Concatenate data and background and apply the alpha rescale (model.py:603-641). A
        background set keeps the weights it was given (the config loader gives every
        background sample -bg_weight or minus its weight file); one without weights takes
        -w_bkg per event. Returns (WeightedData, weight)."""
        sets = _as_sets(data)
        if weight is None:
            weight = np.concatenate([s.weight for s in sets])
        elif np.isscalar(weight):
            weight = np.full(sum(s.n for s in sets), float(weight))
        weight = np.asarray(weight, dtype=np.float64)
        if bg is not None:
            bsets = _as_sets(bg)
            bw = []
            for s in bsets:
                bw.append(s.weight if getattr(s, "has_weight", True) else
                          np.full(s.n, -self.w_bkg))
            sets = sets + bsets
            weight = np.concatenate([weight] + bw)
        if alpha:
            weight = weight * (weight.sum() / (weight * weight).sum())
        return WeightedData(sets, weight), weight

    @staticmethod
    def _mc_weight(mcdata, mc_weight=None):
        sets = _as_sets(mcdata)
        if mc_weight is None or np.isscalar(mc_weight):
            w = np.concatenate([s.weight for s in sets])
        else:
            w = np.asarray(mc_weight, dtype=np.float64)
        return sets, w / w.sum()

    @property
    def evaluation(self):
        """"stream" or "fused" (see __init__)."""
        return self._evaluation

    @evaluation.setter
    def evaluation(self, value):
        if value not in ("stream", "fused", "recompute"):
            raise ValueError("evaluation must be 'stream', 'fused' or 'recompute'")
        if value == "fused" and self.Amp.P > _lib.FUSED_MAX_P:
            raise NotImplementedError(
                "the fused pass is disabled above {} partial waves (this model has {}): the "
                "kernel generated for a 22-wave model faulted, cause not identified".format(
                    _lib.FUSED_MAX_P, self.Amp.P))
        self._evaluation = value

    def _pass(self, sets, weights, coef, mode, grad):
        """Sum of one mode over event sets: value and adjoints (G, or (G, Dm, Dg) when the
        lineshapes float)."""
        shape = self.Amp.shape_mode
        P = self.Amp.P
        value = 0.0
        G = [np.zeros(P, dtype=np.complex128) for _ in range(3 if shape else 1)] if grad else None
        for s, w in zip(sets, weights):
            s.set_weight(w)
            if mode == _lib.MODE_MC and self.precision == "mixed" and not shape:
                I = s.gram(self.gram_precision)
                Ic = I @ coef
                v = float(np.real(np.vdot(coef, Ic)))
                g = (np.conj(Ic),)
            elif self.evaluation == "fused":
                if shape:
                    raise NotImplementedError("the fused pass needs fixed lineshapes")
                v, g = s.eval_fused(coef, mode, grad=grad)
                g = (g,)
            elif self.evaluation == "recompute":
                if shape:
                    raise NotImplementedError("the per-call rebuild needs fixed lineshapes")
                v, g = s.eval_recompute(coef, mode, grad=grad, chunk=self.recompute_chunk)
                g = (g,)
            else:
                v, g, _ = s.eval(coef, mode, grad=grad, f32=self._data_f32(mode),
                                 half=self._data_half(mode), half64=self._half64(mode))
                if grad and not shape:
                    g = (g,)
            value += v
            if grad:
                for acc, x in zip(G, g):
                    acc += x
        return value, G

    def _data_f32(self, mode):
        """Whether a pass of this mode reads the FP32 data caches."""
        return (mode == _lib.MODE_DATA and self.precision == "mixed"
                and self.data_storage == "fp32" and not self.Amp.shape_mode)

    def _data_half(self, mode):
        """Whether a pass of this mode reads the helicity-halved FP32 data caches."""
        return (mode == _lib.MODE_DATA and self.precision == "mixed"
                and self.data_storage == "fp32h" and not self.Amp.shape_mode)

    def _half64(self, mode):
        """Whether a pass of this mode reads the helicity-halved FP64 caches (data passes with
        data_storage "fp64h", FP64-path MC passes with mc_storage "fp64h")."""
        if self.Amp.shape_mode:
            return False
        if mode == _lib.MODE_DATA:
            return self.data_storage == "fp64h"
        return self.precision == "fp64" and self.mc_storage == "fp64h"

    @staticmethod
    def _split(sets, w):
        out, i = [], 0
        for s in sets:
            out.append(w[i:i + s.n])
            i += s.n
        return out

    def _nll_grad(self, wdata, mc_sets, mc_w, grad=True, split=None):
        """``split``: optional precomputed (data weights per set, MC weights per set, sum w)
        so repeated calls do not touch the weight arrays."""
        amp = self.Amp
        coef = amp.coefficients()
        if split is None:
            split = (wdata.split(), self._split(mc_sets, mc_w), wdata.weight.sum())
        dw, mw, sw = split
        ln, Gd = self._pass(wdata.sets, dw, coef, _lib.MODE_DATA, grad)
        int_mc, Gm = self._pass(mc_sets, mw, coef, _lib.MODE_MC, grad)
        self.n_eval += 1
        nll = -ln + sw * np.log(int_mc)
        if not grad:
            return nll, None
        comb = [-a + (sw / int_mc) * b for a, b in zip(Gd, Gm)]
        g = amp.grad_from_adjoint(comb[0])
        if len(comb) == 3:
            g = g + amp.shape_grad(*comb)
        return nll, g

    def _nll_grad_hess(self, wdata, mc_sets, mc_w, split=None):
        """NLL, gradient and exact Hessian in the trainable couplings (fixed lineshapes):
        NLL = -D + sw log S with D, S and their Wirtinger derivatives from the second-moment
        passes (the MC through its Gram matrix in mixed precision), then the coupling chain
        rule (AmplitudeModel.hess_from_adjoint). Follows BaseModel.nll_grad_hessian
        (tf_pwa/model/model.py:513-560): h = -h_data + sw (g_mc g_mc^T int_h + h_mc int_g)."""
        amp = self.Amp
        coef = amp.coefficients()
        if split is None:
            split = (wdata.split(), self._split(mc_sets, mc_w), wdata.weight.sum())
        dw, mw, sw = split
        P = amp.P
        D, Dk = 0.0, np.zeros(P, dtype=np.complex128)
        Dcc = np.zeros((P, P), dtype=np.complex128)
        Dcb = np.zeros((P, P), dtype=np.complex128)
        for s, w in zip(wdata.sets, dw):
            s.set_weight(w)
            v, g, cc, cb = s.hess(coef, _lib.MODE_DATA)
            D, Dk, Dcc, Dcb = D + v, Dk + g, Dcc + cc, Dcb + cb
        S, Sk = 0.0, np.zeros(P, dtype=np.complex128)
        Scb = np.zeros((P, P), dtype=np.complex128)
        for s, w in zip(mc_sets, mw):
            s.set_weight(w)
            if self.precision == "mixed":
                I = s.gram(self.gram_precision)
                Ic = I @ coef
                S += float(np.real(np.vdot(coef, Ic)))
                Sk += np.conj(Ic)
                Scb += I.T
            else:
                v, g, _, cb = s.hess(coef, _lib.MODE_MC)
                S, Sk, Scb = S + v, Sk + g, Scb + cb
        self.n_eval += 1
        nll = -D + sw * np.log(S)
        F1 = -Dk + (sw / S) * Sk
        Fcc = -Dcc - (sw / S ** 2) * np.outer(Sk, Sk)
        Fcb = -Dcb + sw * (Scb / S - np.outer(Sk, np.conj(Sk)) / S ** 2)
        g = amp.grad_from_adjoint(F1)
        H = amp.hess_from_adjoint(F1, Fcc, Fcb)
        return nll, g, 0.5 * (H + H.T)

    def _nll_grad_hess_shape(self, wdata, mc_sets, mc_w, split=None):
        """NLL, gradient and exact Hessian when masses or widths float: one lineshape
        second-moment pass per data and MC set (tpa_hess_shape) gives the value and the
        Wirtinger derivatives in the extended coefficients, combined as in _nll_grad_hess
        (h = -h_data + sw (h_mc / S - g_mc g_mc^T / S^2), tf_pwa/model/model.py:513-560) and
        chain-ruled to couplings, masses and widths (AmplitudeModel.hess_shape_from_adjoint)."""
        amp = self.Amp
        coef = amp.coefficients()
        if split is None:
            split = (wdata.split(), self._split(mc_sets, mc_w), wdata.weight.sum())
        dw, mw, sw = split
        P3 = 3 * amp.P

        def zero():
            return [0.0, np.zeros(2 * P3, dtype=np.complex128),
                    np.zeros((P3, P3), dtype=np.complex128),
                    np.zeros((P3, P3), dtype=np.complex128)]

        def run(sets, weights, mode):
            acc = zero()
            for st, w in zip(sets, weights):
                st.set_weight(w)
                v, F, F2, cc, cb = st.hess_shape(coef, mode)
                acc[0] += v
                acc[1] += np.concatenate([F, F2])
                acc[2] += cc
                acc[3] += cb
            return acc

        D, DF, Dcc, Dcb = run(wdata.sets, dw, _lib.MODE_DATA)
        S, SF, _, Scb = run(mc_sets, mw, _lib.MODE_MC)
        self.n_eval += 1
        nll = -D + sw * np.log(S)
        F1 = -DF + (sw / S) * SF
        s1 = SF[:P3]
        Fcc = -Dcc - (sw / S ** 2) * np.outer(s1, s1)
        Fcb = -Dcb + sw * (Scb / S - np.outer(s1, np.conj(s1)) / S ** 2)
        P = amp.P
        g = amp.grad_from_adjoint(F1[:P]) + amp.shape_grad(F1[:P], F1[P:2 * P], F1[2 * P:P3])
        H = amp.hess_shape_from_adjoint(F1, Fcc, Fcb)
        return nll, g, H

    def _hess_any(self, wdata, mc_sets, mc_w, split=None):
        if self.Amp.shape_mode:
            return self._nll_grad_hess_shape(wdata, mc_sets, mc_w, split=split)
        return self._nll_grad_hess(wdata, mc_sets, mc_w, split=split)

    def nll(self, data, mcdata, weight=1.0, batch=None, bg=None, mc_weight=1.0):
        wdata, _ = self.get_weight_data(data, None if np.isscalar(weight) and weight == 1.0
                                        else weight, bg=bg)
        mc_sets, mc_w = self._mc_weight(mcdata, None if np.isscalar(mc_weight) else mc_weight)
        return self._nll_grad(wdata, mc_sets, mc_w, grad=False)[0]

    def nll_grad(self, data, mcdata, weight=1.0, batch=65000, bg=None, mc_weight=1.0):
        wdata, _ = self.get_weight_data(data, None if np.isscalar(weight) and weight == 1.0
                                        else weight, bg=bg)
        mc_sets, mc_w = self._mc_weight(mcdata, None if np.isscalar(mc_weight) else mc_weight)
        return self._nll_grad(wdata, mc_sets, mc_w)

    def nll_grad_batch(self, data, mcdata, weight, mc_weight):
        """``data``/``mcdata``: event sets (or lists); ``weight``/``mc_weight``: the already
        rescaled / normalised weights over their concatenation (model.py:739-759)."""
        wdata = WeightedData(_as_sets(data), np.concatenate([np.atleast_1d(w) for w in weight])
                             if isinstance(weight, (list, tuple)) else weight)
        mc_sets = _as_sets(mcdata)
        mc_w = (np.concatenate([np.atleast_1d(w) for w in mc_weight])
                if isinstance(mc_weight, (list, tuple)) else np.asarray(mc_weight))
        return self._nll_grad(wdata, mc_sets, mc_w)

    def nll_grad_hessian(self, data, mcdata, weight=1.0, batch=24000, bg=None, mc_weight=1.0):
        """NLL, gradient and exact Hessian (model.py:513-560) in every trainable parameter:
        second-moment passes for couplings, lineshape second-moment passes when masses or
        widths float."""
        wdata, _ = self.get_weight_data(data, None if np.isscalar(weight) and weight == 1.0
                                        else weight, bg=bg)
        mc_sets, mc_w = self._mc_weight(mcdata, None if np.isscalar(mc_weight) else mc_weight)
        return self._hess_any(wdata, mc_sets, mc_w)

    def grad_hessp_batch(self, p, data, mcdata, weight, mc_weight):
        """Gradient and Hessian-vector product H p (BaseModel.grad_hessp_batch,
        tf_pwa/model/model.py:451-511) from the exact Hessian."""
        wdata = WeightedData(_as_sets(data), np.concatenate([np.atleast_1d(w) for w in weight])
                             if isinstance(weight, (list, tuple)) else weight)
        mc_sets = _as_sets(mcdata)
        mc_w = (np.concatenate([np.atleast_1d(w) for w in mc_weight])
                if isinstance(mc_weight, (list, tuple)) else np.asarray(mc_weight))
        p = np.asarray(p, dtype=np.float64)
        _, g, H = self._hess_any(wdata, mc_sets, mc_w)
        return g, H @ p

    def set_params(self, var):
        self.Amp.set_params(var)

    def get_params(self, trainable_only=False):
        return self.Amp.get_params(trainable_only)


class _NativeSession:
    """One tpa_fcn session: the coupling map of the current parameter store, the event sets
    with their weights in place and, in mixed precision, the summed MC Gram matrix; with
    floating masses or widths (three-body chains) also the shape map (tpa_fcn_set_shape_map).
    Valid while the store's fixed values, the trainable set, the weights and the barrier clamp
    flags are unchanged (the caller keys it on those)."""

    def __init__(self, model, data_sets, data_w, mc_sets, mc_w, sw):
        import ctypes

        from .serving import coupling_map, shape_map

        amp = model.Amp
        self.amp = amp
        self.shape = amp.shape_mode
        for s, w in zip(data_sets, data_w):
            s.set_weight(w)
            s.ensure_built()
        gram = None
        if model.precision == "mixed" and not self.shape:
            I = np.zeros((amp.P, amp.P), dtype=np.complex128)
            for s, w in zip(mc_sets, mc_w):
                s.set_weight(w)
                I = I + s.gram(model.gram_precision)
            self._gram = np.ascontiguousarray(I).view(np.float64).ravel()
            gram = _lib.dptr(self._gram)
        else:
            for s, w in zip(mc_sets, mc_w):
                s.set_weight(w)
                s.ensure_built()
        self._map, self._keep = coupling_map(amp)
        vp = ctypes.c_void_p
        self._data = (vp * len(data_sets))(*[s._h.value for s in data_sets])
        self._mc = (vp * max(1, len(mc_sets)))(*[s._h.value for s in mc_sets])
        self._sets = list(data_sets) + list(mc_sets)  # keep the handles alive
        h = vp()
        _lib.check(_lib.lib().tpa_fcn_create(ctypes.byref(self._map), self._data, len(data_sets),
                                             self._mc, len(mc_sets), gram, float(sw),
                                             ctypes.byref(h)))
        self._h = h
        if self.shape:
            self._shape = shape_map(amp)
            _lib.check(_lib.lib().tpa_fcn_set_shape_map(h, amp.device(), ctypes.byref(self._shape)))
        elif model._data_f32(_lib.MODE_DATA):
            _lib.check(_lib.lib().tpa_fcn_set_data_storage(h, 1))
            for s in data_sets:
                s._keep32 = True
        elif model._data_half(_lib.MODE_DATA):
            # (the session keeps exactly the halved copy it reads: the other one is freed)
            _lib.check(_lib.lib().tpa_fcn_set_data_half(h, 1))
            for s in data_sets:
                s._keep32h, s._keep64h = True, False
        elif model._half64(_lib.MODE_DATA):
            _lib.check(_lib.lib().tpa_fcn_set_data_half(h, 2))
            for s in data_sets:
                s._keep32h, s._keep64h = False, True
        if not self.shape and model._half64(_lib.MODE_MC) and gram is None:
            _lib.check(_lib.lib().tpa_fcn_set_mc_half(h, 1))
            for s in mc_sets:
                s._keep32h, s._keep64h = False, True
        self.n = len(self._keep["bound_kind"])
        # argument buffers with their pointers taken once (a ctypes pointer per call costs
        # ~3 us of the host's time per evaluation)
        self._ybuf = np.zeros(self.n)
        self._gbuf = np.zeros(self.n)
        self._yp, self._gp = _lib.dptr(self._ybuf), _lib.dptr(self._gbuf)
        self._nll = ctypes.c_double()
        self._nllp = ctypes.byref(self._nll)
        self._rebuilds = ctypes.c_int64(0)
        self._seen_rebuilds = 0

    def after_call(self):
        """Bring the Python host's view in line with what the session did on the device: it
        set the lineshape parameters (and, when a mass crossed a barrier clamp point, the
        clamp flags and every cache), so the description is uploaded again and the sets
        rebuild before the next Python-side pass."""
        import ctypes

        if not self.shape:
            return
        self.amp._desc_state = None
        _lib.check(_lib.lib().tpa_fcn_cache_rebuilds(self._h, ctypes.byref(self._rebuilds)))
        if self._rebuilds.value != self._seen_rebuilds:
            self._seen_rebuilds = self._rebuilds.value
            for s in self._sets:
                s._state = None

    def nll_grad(self, y, grad=True):
        self._ybuf[:] = y
        rc = _lib.lib().tpa_fcn_nll_grad(self._h, self._yp, self._nllp,
                                         self._gp if grad else None)
        self.after_call()
        _lib.check(rc)
        return self._nll.value, (self._gbuf.copy() if grad else None)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tpa_fcn_destroy(h)
        self._h = None

    def __del__(self):
        self.close()


class FCN:
    """The callable minimised by the fit drivers (tf_pwa/model/model.py:1115-1285):
    ``fcn(x)``, ``fcn.nll_grad(x)`` -> (float, ndarray in vm.trainable_vars order),
    ``fcn.nll_grad_hessian(x)``, ``fcn.grad_hessp(x, p)``."""

    def __init__(self, model, data, mcdata, bg=None, batch=65000, inmc=None, gauss_constr=None,
                 precision=None):
        if precision is not None:
            model.precision = precision
        if inmc is not None:
            raise NotImplementedError("inmc is not supported")
        self.model = model
        self.vm = model.vm
        self.n_call = 0
        self.cached_nll = None
        self.batch = batch
        self.wdata, self.weight = model.get_weight_data(data, bg=bg)
        self.data = self.wdata
        self.alpha = self.weight.sum() / (self.weight * self.weight).sum()
        self.mc_sets, self.mc_weight = model._mc_weight(mcdata)
        self.mcdata = self.mc_sets
        self.gauss_constr = dict(gauss_constr or {})
        self.last_time = 0.0
        self._split = (self.wdata.split(), Model._split(self.mc_sets, self.mc_weight),
                       float(self.weight.sum()))
        # fixed lineshapes: one native call per evaluation (tpa_fcn_*); False keeps the Python
        # host assembly (Model._nll_grad), the reference implementation of the same sums
        self.native = True
        self._native = None
        self._native_key = None
        self._set_weights = None  # (event set, its weights) of the data and MC sets

    def _set(self, x):
        if x is None or (isinstance(x, dict) and not x):
            return
        self.vm.set_all(x)

    def _constr(self):
        """Gaussian constraints sum ((x - mu)/sigma)^2 / 2 (model.py:915-987)."""
        v, g = 0.0, np.zeros(len(self.vm.trainable_vars))
        for name, (mu, sigma) in self.gauss_constr.items():
            x = self.vm.get(name)
            v += 0.5 * ((x - mu) / sigma) ** 2
            if name in self.vm.trainable_vars:
                g[self.vm.trainable_vars.index(name)] += (x - mu) / sigma ** 2
        return v, g

    def get_params(self, trainable_only=False):
        return self.vm.get_all_dic(trainable_only)

    def _native_session(self):
        """The native session (None for floating lineshapes of four-body chains), rebuilt when
        the precision, fixed values, trainable set or any set's weights change."""
        amp = self.model.Amp
        if not self.native or (amp.shape_mode and amp.n_final != 3):
            return None
        if getattr(self.model, "evaluation", "stream") in ("fused", "recompute"):
            return None  # these passes run through the host assembly (Model._pass)
        dw, mw, sw = self._split
        pairs = self._set_weights
        if pairs is None:
            pairs = self._set_weights = list(zip(list(self.wdata.sets) + list(self.mc_sets),
                                                 list(dw) + list(mw)))
        for st, w in pairs:
            st.set_weight(w)
        key = (self.model.precision, self.model.gram_precision, self.model.data_storage,
               getattr(self.model, "mc_storage", "fp64"),
               self.vm.struct_version, tuple([st._weight_version for st, _ in pairs]))
        if self._native_key != key:
            if self._native is not None:
                self._native.close()
            self._native = None
            self._native = _NativeSession(self.model, self.wdata.sets, dw, self.mc_sets, mw, sw)
            self._native_key = key
        return self._native

    def _y(self, x):
        if x is not None and not isinstance(x, dict) and len(x) == len(self.vm.trainable_vars):
            return np.asarray(x, dtype=np.float64)
        return np.array(self.vm.get_all_val())

    def _native_call(self, x, grad):
        """(nll, g) from the native session, or None when there is none."""
        nat = self._native_session()
        if nat is None:
            return None
        out = nat.nll_grad(self._y(x), grad=grad)
        self.model.n_eval += 1
        return out

    def get_nll(self, x=None):
        self._set(x)
        r = self._native_call(x, grad=False)
        if r is not None:
            nll = r[0]
        else:
            nll, _ = self.model._nll_grad(self.wdata, self.mc_sets, self.mc_weight, grad=False,
                                          split=self._split)
        self.n_call += 1
        return nll

    def __call__(self, x=None):
        self.cached_nll = self.get_nll(x) + self._constr()[0]
        return self.cached_nll

    def get_nll_grad(self, x=None):
        self._set(x)
        r = self._native_call(x, grad=True)
        if r is not None:
            nll, g = r
        else:
            nll, g = self.model._nll_grad(self.wdata, self.mc_sets, self.mc_weight,
                                          split=self._split)
        self.n_call += 1
        return nll, g

    def nll_grad(self, x=None):
        t = time.perf_counter()
        nll, g = self.get_nll_grad(x)
        if self.gauss_constr:
            cv, cg = self._constr()
            nll, g = nll + cv, g + cg
        self.cached_nll = nll
        self.last_time = time.perf_counter() - t
        return float(nll), g

    def grad(self, x=None):
        return self.nll_grad(x)[1]

    def nll_grad_hessian(self, x=None, batch=None):
        """NLL, gradient and exact Hessian in every trainable parameter (FCN.nll_grad_hessian,
        model.py:1258-1265): second-moment passes for couplings, lineshape second-moment
        passes when masses or widths float, plus the Gaussian constraints' curvature."""
        self._set(x)
        nll, g, H = self.model._hess_any(self.wdata, self.mc_sets, self.mc_weight,
                                         split=self._split)
        self.n_call += 1
        cv, cg = self._constr()
        for name, (mu, sigma) in self.gauss_constr.items():
            if name in self.vm.trainable_vars:
                i = self.vm.trainable_vars.index(name)
                H[i, i] += 1.0 / sigma ** 2
        return nll + cv, g + cg, H

    def grad_hessp(self, x, p, batch=None):
        """Gradient and Hessian-vector product H p (FCN.grad_hessp, model.py:1267-1275) from
        the exact Hessian."""
        x = np.asarray(x, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        _, g, H = self.nll_grad_hessian(x)
        return g, H @ p
