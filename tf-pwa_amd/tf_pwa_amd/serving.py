"""Many fits on one GPU: a shared device event pool, one batched launch per round of pending
NLL + gradient requests, and a native BFGS per fit.

tf-pwa runs one fit per process: fit_scipy (tf_pwa/fit.py:189-300) calls FCN.nll_grad
(tf_pwa/model/model.py:1216-1240) once per BFGS evaluation, and each call is a full pass over
that fit's data and MC. Here many fits share

- one EventPool: the data samples of all admitted fits live in one device store (angles and
  fixed-lineshape partial-wave cache per event), each fit owning a contiguous range;
- one MC normalisation: the partial-wave Gram matrix of the shared MC sample, so that
  int_mc = c^dagger I c for any coefficients;
- one batcher: the pending requests of all fits go to the GPU as a single replayed hipGraph.

The per-fit BFGS (strong-Wolfe line search, max |g| <= gtol) and the coefficient / gradient
chain rule run in the library's native host threads; this module builds the coupling map,
admits fits under an HBM budget and collects their results.
"""

import ctypes
import time

import numpy as np

from . import _lib
from .data import p4_array
from .fit import FitResult

STATUS = {0: "converged", 1: "iteration limit", 2: "line search failed (precision loss)",
          3: "cancelled", 4: "engine error"}


class PoolFull(RuntimeError):
    """No free range of the pool is long enough for the sample (admission must wait)."""


def _alpha_weights(n, weight):
    """Data weights with the alpha = sum w / sum w^2 rescale (tf_pwa/model/model.py:603-641)."""
    if n < 1:
        raise ValueError("a data sample needs at least one event")
    w = np.ones(n) if weight is None else np.ascontiguousarray(weight, dtype=np.float64).copy()
    if w.shape != (n,):
        raise ValueError("weight must have one entry per event")
    return w * (w.sum() / (w * w).sum())


class DataSlot:
    """A fit's data range in the pool: first event, size and the weight sum of the NLL."""

    def __init__(self, offset, n, sw):
        self.offset, self.n, self.sw = int(offset), int(n), float(sw)

    def __repr__(self):
        return "DataSlot(offset={}, n={})".format(self.offset, self.n)


class EventPool:
    """Device event store shared by many fits (tpa_pool_*). Lineshapes are folded into the
    cache when a sample is added, so the amplitude model must not float masses or widths."""

    def __init__(self, amp, capacity, max_items=512, data_storage="fp64"):
        """data_storage "fp32": the batched passes read an FP32 copy of the pooled partial
        waves (tpa_pool_keep_f32; FP64 arithmetic), "fp64": the FP64 cache."""
        if amp.shape_mode:
            raise NotImplementedError("the event pool caches fixed lineshapes; masses and "
                                      "widths must not be trainable")
        if data_storage not in ("fp32", "fp64"):
            raise ValueError("data_storage must be 'fp32' or 'fp64'")
        self.amp = amp
        self.capacity = int(capacity)
        self.max_items = int(max_items)
        self.data_storage = data_storage
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tpa_pool_create(amp.device(), self.capacity, self.max_items,
                                              ctypes.byref(h)))
        self._h = h
        self.capacity = self.free_events()  # the library rounds it up to a multiple of 64
        if data_storage == "fp32":
            _lib.check(_lib.lib().tpa_pool_keep_f32(h, 1))

    @staticmethod
    def bytes_per_event(amp, data_storage="fp64"):
        """Device bytes one pooled event occupies: the partial-wave cache (and its FP32 copy),
        the angles and the weight."""
        nf = _lib.NF_TOPO * len(amp.topo) + 4
        extra = 8 * amp.P * amp.lam_pad if data_storage == "fp32" else 0
        return 16 * amp.P * amp.lam_pad + extra + 8 * nf + 8

    def add(self, p4, weight=None, order=None, engine_layout=False):
        """Upload a sample into a free range. ``p4`` as for EventSet (dict or array in
        ``order``), or with ``engine_layout`` the [n_final*4][n] array of p4_array."""
        if engine_layout:
            arr = np.ascontiguousarray(p4, dtype=np.float64)
            if arr.ndim != 2 or arr.shape[0] != 4 * len(self.amp.finals):
                raise ValueError("engine layout is [n_final*4][n]")
        else:
            arr = np.ascontiguousarray(p4_array(p4, self.amp.finals, order))
        n = arr.shape[1]
        w = _alpha_weights(n, weight)
        off = ctypes.c_int64()
        rc = _lib.lib().tpa_pool_add(self._h, _lib.dptr(arr), n, _lib.dptr(w), ctypes.byref(off))
        if rc != 0:
            msg = _lib.lib().tpa_last_error().decode(errors="replace")
            if "pool full" in msg:
                raise PoolFull(msg)
            _lib.check(rc)
        return DataSlot(off.value, n, w.sum())

    def add_events(self, events, weight=None):
        """Add the events of a device EventSet of the same model (tpa_pool_add_events): its
        angles are copied on the device, so a sample generated on the GPU needs no host
        round trip. ``weight``: host [n], default the set's own weights."""
        events.amp.device()
        n = events.n
        w = _alpha_weights(n, events.weight if weight is None else weight)
        off = ctypes.c_int64()
        rc = _lib.lib().tpa_pool_add_events(self._h, events._h, _lib.dptr(w), ctypes.byref(off))
        if rc != 0:
            msg = _lib.lib().tpa_last_error().decode(errors="replace")
            if "pool full" in msg:
                raise PoolFull(msg)
            _lib.check(rc)
        return DataSlot(off.value, n, w.sum())

    def release(self, slot):
        _lib.check(_lib.lib().tpa_pool_release(self._h, slot.offset))

    def free_events(self):
        return int(_lib.lib().tpa_pool_free_events(self._h))

    def eval(self, slots, coefs, ranges=None):
        """Batched data sums: (values [m], G [m, P], graph ms) for coefficient rows ``coefs``
        ([m, P] complex) on ``slots`` (or explicit (offset, n) ``ranges``)."""
        P = self.amp.P
        if ranges is None:
            ranges = [(s.offset, s.n) for s in slots]
        r = np.ascontiguousarray(np.asarray(ranges, dtype=np.int64).reshape(-1, 2))
        m = r.shape[0]
        c = np.ascontiguousarray(np.asarray(coefs, dtype=np.complex128).reshape(m, P))
        out = np.zeros((m, 2 + 2 * P))
        ms = ctypes.c_double(0.0)
        _lib.check(_lib.lib().tpa_pool_eval(
            self._h, m, r.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            _lib.dptr(c.view(np.float64).reshape(-1)), _lib.MODE_DATA, _lib.dptr(out),
            ctypes.byref(ms)))
        G = out[:, 2:].copy().view(np.complex128)
        return out[:, 0].copy(), G, ms.value

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tpa_pool_destroy(h)
        self._h = None

    def __del__(self):
        self.close()


def shape_map(amp):
    """The TpaShapeMap of a three-body model with floating lineshapes: per chain its mass and
    width variables (trainable index or fixed value), the masses and LS lists of both vertices
    for the barrier constants, the clamp flags the caches were built with, and per partial
    wave its chain and LS indices."""
    if amp.n_final != 3:
        raise NotImplementedError("native floating lineshapes cover three-body chains")
    vm = amp.vm
    index = vm.index_map()
    names = amp._names()
    bc = amp.barrier_constants()
    fx = amp._factor_index()
    m = _lib.TpaShapeMap()
    m.n_chain = len(amp.chains)
    m.P = amp.P
    m.m_top = float(amp.decay_group.top.mass)
    shapes = amp.shape_params()
    for c, (ch, nm) in enumerate(zip(amp.chains, names)):
        R, d1, d2 = ch["R"], ch["d1"], ch["d2"]
        m.mass_var[c] = index.get(nm["mass"], -1)
        m.width_var[c] = index.get(nm["width"], -1)
        m.mass_fixed[c], m.width_fixed[c] = shapes[c]
        bach = d1.outs[1] if d1.outs[0] is R else d1.outs[0]
        m.m_bach[c] = float(bach.mass)
        m.m_y[c], m.m_z[c] = float(d2.outs[0].mass), float(d2.outs[1].mass)
        m.d[c] = float(d1.d)
        m.n_l1[c], m.n_l2[c] = len(d1.l_list), len(d2.l_list)
        for s, l in enumerate(d1.l_list):
            m.l1[c][s] = int(l)
            m.act1[c][s] = int(bc[c]["act"][0][s])
        for s, l in enumerate(d2.l_list):
            m.l2[c][s] = int(l)
            m.act2[c][s] = int(bc[c]["act"][1][s])
    for k, (c, ls) in enumerate(zip(fx["ch"], fx["ls"])):
        m.wave_chain[k] = int(c)
        m.wave_ls1[k], m.wave_ls2[k] = int(ls[0]), int(ls[1])
    return m


def coupling_map(amp):
    """The TpaCouplingMap of ``amp``'s current parameter store: per partial wave its three
    complex factors and barrier constant, per factor its trainable indices or fixed values,
    per trainable variable its bound. Returns (map, arrays kept alive for the map)."""
    vm = amp.vm
    fx = amp._factor_index()
    index = vm.index_map()
    names = fx["cnames"]
    slots = list(fx["slots"])
    unit = len(names)  # a fixed factor 1 + 0i fills the unused slots of three-body waves
    while len(slots) < _lib.CPL_SLOTS:
        slots.append(np.full(amp.P, unit))
    keep = dict(
        wave_factor=np.ascontiguousarray(np.stack(slots, 1), dtype=np.int32),
        wave_const=np.ascontiguousarray(amp._barrier_per_wave(), dtype=np.float64),
        factor_polar=np.array([int(bool(vm.complex_vars.get(n, vm.polar))) for n in names] + [0],
                              dtype=np.int32),
        factor_var=np.array([[index.get(n + "r", -1), index.get(n + "i", -1)] for n in names]
                            + [[-1, -1]], dtype=np.int32),
        factor_fixed=np.array([[vm.variables[n + "r"], vm.variables[n + "i"]] for n in names]
                              + [[1.0, 0.0]], dtype=np.float64),
    )
    kind, ab = [], []
    for n in vm.trainable_vars:
        b = vm.bnd_dic.get(n)
        if b is None or (b.lower is None and b.upper is None):
            kind.append(0)
            ab.append((0.0, 0.0))
        elif b.lower is not None and b.upper is not None:
            kind.append(1)
            ab.append((b.lower, b.upper))
        elif b.lower is not None:
            kind.append(2)
            ab.append((b.lower, 0.0))
        else:
            kind.append(3)
            ab.append((0.0, b.upper))
    keep["bound_kind"] = np.array(kind, dtype=np.int32)
    keep["bound_ab"] = np.array(ab, dtype=np.float64).reshape(-1, 2)
    i32p = ctypes.POINTER(ctypes.c_int32)
    m = _lib.TpaCouplingMap(
        P=amp.P, n_factor=len(names) + 1, n_var=len(vm.trainable_vars), reserved=0,
        wave_factor=keep["wave_factor"].ctypes.data_as(i32p),
        wave_const=_lib.dptr(keep["wave_const"]),
        factor_polar=keep["factor_polar"].ctypes.data_as(i32p),
        factor_var=keep["factor_var"].ctypes.data_as(i32p),
        factor_fixed=_lib.dptr(keep["factor_fixed"]),
        bound_kind=keep["bound_kind"].ctypes.data_as(i32p),
        bound_ab=_lib.dptr(keep["bound_ab"]))
    return m, keep


def map_coefficients(keep, y):
    """c_k from a coupling map's arrays and parameters y (the formula the native server uses;
    a host check of the map)."""
    y = np.asarray(y, dtype=np.float64)
    fv, fixed = keep["factor_var"], keep["factor_fixed"]
    r = np.where(fv[:, 0] >= 0, y[np.maximum(fv[:, 0], 0)], fixed[:, 0])
    i = np.where(fv[:, 1] >= 0, y[np.maximum(fv[:, 1], 0)], fixed[:, 1])
    v = np.where(keep["factor_polar"] == 1, r * np.exp(1j * i), r + 1j * i)
    wf = keep["wave_factor"]
    c = keep["wave_const"].astype(np.complex128)
    for a in range(wf.shape[1]):
        c = c * v[wf[:, a]]
    return c


class FitServer:
    """Serve many BFGS fits of one amplitude model at once.

    ``mc``: the shared MC EventSet (its weights are normalised to unit sum, as FCN does,
    tf_pwa/model/model.py:1126-1164); its Gram matrix is the MC normalisation of every fit.
    ``hbm_budget``: device bytes for the MC set plus the pool; the pool gets what the MC set
    leaves (or ``capacity`` events when given). Fits whose data do not fit wait for others to end.
    ``hess_start``: each fit's BFGS starts from the inverse exact Hessian at its start point
    (one second-moment pass) rather than from the identity.
    ``gram_precision``: "fp64" (FP64 MFMA) or "fp32" (FP32 MFMA with FP64 refolds) Gram.
    ``data_storage``: "fp32" keeps an FP32 copy of the pooled data waves for the batched passes
    (half the bytes per event; FP64 arithmetic), "fp64" reads the FP64 cache.
    """

    def __init__(self, amp, mc, hbm_budget=200e9, capacity=None, max_batch=512,
                 window_us=200.0, gtol=1e-5, maxiter=0, hess_start=False, gram_precision="fp64",
                 data_storage="fp64"):
        self.amp = amp
        self.vm = amp.vm
        self.mc = mc
        mc.set_weight(mc.weight / mc.weight.sum())
        self.gram = np.ascontiguousarray(mc.gram(gram_precision))
        if capacity is None:
            bpe = EventPool.bytes_per_event(amp, data_storage)
            capacity = int((hbm_budget - mc.n * EventPool.bytes_per_event(amp)) // bpe)
        if capacity < 1:
            raise ValueError("the HBM budget leaves no room for data")
        self.pool = EventPool(amp, capacity, max_batch, data_storage)
        self._map, self._keep = coupling_map(amp)
        self.n_var = len(self.vm.trainable_vars)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tpa_server_create(
            self.pool._h, ctypes.byref(self._map), _lib.dptr(self.gram.view(np.float64).ravel()),
            float(gtol), int(maxiter), int(max_batch), float(window_us), int(bool(hess_start)),
            ctypes.byref(h)))
        self._h = h
        self.slots = {}

    # -------------------------------------------------------------- data and single calls
    def add_data(self, p4, weight=None, order=None, engine_layout=False):
        return self.pool.add(p4, weight, order, engine_layout)

    def release(self, slot):
        self.pool.release(slot)

    def nll_grad(self, slot, y):
        """NLL and gradient at trainable values ``y`` (vm.trainable_vars order) on ``slot``."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        g = np.zeros(self.n_var)
        nll = ctypes.c_double()
        _lib.check(_lib.lib().tpa_server_nll_grad(self._h, slot.offset, slot.n, slot.sw,
                                                  _lib.dptr(y), ctypes.byref(nll),
                                                  _lib.dptr(g)))
        return nll.value, g

    def nll_grad_hessian(self, slot, y):
        """NLL, gradient and exact Hessian at ``y`` on ``slot`` (one second-moment pass)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        g = np.zeros(self.n_var)
        H = np.zeros((self.n_var, self.n_var))
        nll = ctypes.c_double()
        _lib.check(_lib.lib().tpa_server_nll_grad_hess(self._h, slot.offset, slot.n, slot.sw,
                                                       _lib.dptr(y), ctypes.byref(nll),
                                                       _lib.dptr(g), _lib.dptr(H)))
        return nll.value, g, H

    # -------------------------------------------------------------- fits
    def submit(self, slot, y0):
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        if y0.shape != (self.n_var,):
            raise ValueError("y0 must have one value per trainable variable")
        fid = ctypes.c_int()
        _lib.check(_lib.lib().tpa_server_submit(self._h, slot.offset, slot.n, slot.sw,
                                                _lib.dptr(y0), ctypes.byref(fid)))
        self.slots[fid.value] = slot
        return fid.value

    def poll(self, max_ids=4096):
        ids = (ctypes.c_int * max_ids)()
        n = ctypes.c_int()
        _lib.check(_lib.lib().tpa_server_poll(self._h, ids, max_ids, ctypes.byref(n)))
        return list(ids[:n.value])

    def raw_result(self, fit_id):
        y = np.zeros(self.n_var)
        nll = ctypes.c_double()
        st = ctypes.c_int()
        nfev = ctypes.c_int64()
        nit = ctypes.c_int64()
        _lib.check(_lib.lib().tpa_server_result(self._h, fit_id, _lib.dptr(y), ctypes.byref(nll),
                                                ctypes.byref(st), ctypes.byref(nfev),
                                                ctypes.byref(nit)))
        return y, nll.value, st.value, nfev.value, nit.value

    def result(self, fit_id):
        """FitResult with the full parameter dict (fixed values from the store, polar pairs
        standardised to r >= 0, phi in [-pi, pi) as in tf_pwa/variable.py:716-741)."""
        y, nll, st, nfev, nit = self.raw_result(fit_id)
        params = self.vm.get_all_dic()
        params.update(zip(self.vm.trainable_vars, y.tolist()))
        for name, polar in self.vm.complex_vars.items():
            if not polar:
                continue
            r, p = params[name + "r"], params[name + "i"]
            if r < 0:
                r, p = -r, p + np.pi
            params[name + "r"], params[name + "i"] = r, (p + np.pi) % (2 * np.pi) - np.pi
        r = FitResult(params, None, nll, ndf=self.n_var, success=st == 0,
                      message=STATUS.get(st, str(st)), nfev=nfev, ngrad=nfev)
        r.niter = nit
        return r

    # -------------------------------------------------------------- statistics
    def latencies_us(self):
        n = ctypes.c_int64()
        _lib.check(_lib.lib().tpa_server_latencies(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1))
        _lib.check(_lib.lib().tpa_server_latencies(self._h, _lib.dptr(out), n.value,
                                                   ctypes.byref(n)))
        return out[:n.value]

    def counters(self):
        r, b, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        ms = ctypes.c_double()
        _lib.check(_lib.lib().tpa_server_counters(self._h, ctypes.byref(r), ctypes.byref(b),
                                                  ctypes.byref(e), ctypes.byref(ms)))
        return dict(requests=r.value, batches=b.value, events=e.value, gpu_ms=ms.value)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tpa_server_destroy(h)
        self._h = None
        if getattr(self, "pool", None) is not None:
            self.pool.close()

    def __del__(self):
        self.close()

    # -------------------------------------------------------------- the serving loop
    def serve(self, jobs, arrivals, poll_s=2e-4, timeout_s=None):
        """Run ``jobs`` (each a dict with "p4", "y0" and optional "weight", "order",
        "engine_layout"; "p4" may be a callable that draws the sample when the fit is admitted;
        "events", a device EventSet of the model or a callable making one, in place of "p4")
        arriving at ``arrivals`` seconds from now. Admission is first come, first served: the head of the
        queue waits until the pool has a free range for its sample. Returns (results in job
        order, per-fit dict of times: arrival, admitted, done)."""
        t0 = time.perf_counter()
        order = np.argsort(np.asarray(arrivals), kind="stable")
        queue = list(order)
        waiting = []  # arrived, not admitted
        staged = {}  # p4 of waiting jobs whose sample was already drawn
        running = {}  # fit id -> job index
        results = [None] * len(jobs)
        times = [dict(arrival=float(arrivals[j])) for j in range(len(jobs))]
        left = len(jobs)
        peak_events = 0
        while left:
            now = time.perf_counter() - t0
            while queue and arrivals[queue[0]] <= now:
                waiting.append(queue.pop(0))
            while waiting:
                j = waiting[0]
                if j not in staged:
                    src = jobs[j]["events"] if "events" in jobs[j] else jobs[j]["p4"]
                    staged[j] = src() if callable(src) else src
                try:
                    if "events" in jobs[j]:
                        slot = self.pool.add_events(staged[j], jobs[j].get("weight"))
                    else:
                        slot = self.add_data(staged[j], jobs[j].get("weight"),
                                             jobs[j].get("order"),
                                             jobs[j].get("engine_layout", False))
                except PoolFull:
                    if not running:
                        raise
                    break
                del staged[j]
                waiting.pop(0)
                fid = self.submit(slot, jobs[j]["y0"])
                running[fid] = j
                times[j]["admitted"] = time.perf_counter() - t0
                peak_events = max(peak_events, self.pool.capacity - self.pool.free_events())
            for fid in self.poll():
                j = running.pop(fid)
                results[j] = self.result(fid)
                times[j]["done"] = time.perf_counter() - t0
                self.release(self.slots.pop(fid))
                left -= 1
            if timeout_s is not None and time.perf_counter() - t0 > timeout_s:
                raise TimeoutError("serving loop exceeded {} s".format(timeout_s))
            if left:
                time.sleep(poll_s)
        self.peak_events = peak_events
        return results, times
