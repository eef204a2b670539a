"""Fit driver: scipy BFGS over the bound-transformed parameters (tf_pwa/fit.py:189-420) with
the NaN-gradient patch of Cached_FG (tf_pwa/fit_improve.py:23-66)."""

import json
import time

import numpy as np


class LargeNumberError(ValueError):
    """Raised when a parameter runs away (tf_pwa/fit.py:13, 257-260)."""


class FitResult:
    """Result of a fit (tf_pwa/fit.py:508-539)."""

    def __init__(self, params, fcn, min_nll, ndf=0, success=True, message="", hess_inv=None,
                 nfev=0, ngrad=0, time_s=0.0):
        self.params = params
        self.error = {}
        self.fcn = fcn
        self.min_nll = float(min_nll)
        self.ndf = ndf
        self.success = success
        self.message = message
        self.hess_inv = hess_inv
        self.nfev = nfev
        self.ngrad = ngrad
        self.time_s = time_s

    def save_as(self, file_name, save_hess=False):
        s = {"value": self.params, "error": self.error, "status": {
            "success": self.success, "NLL": self.min_nll, "Ndf": self.ndf,
            "message": str(self.message)}}
        with open(file_name, "w") as f:
            json.dump(s, f, indent=2)

    def set_error(self, error):
        self.error = dict(error)


class CachedFG:
    """Caches f and g per x; a NaN gradient component is replaced by a central difference."""

    def __init__(self, f_g, eps=1e-6):
        self.f_g = f_g
        self.eps = eps
        self.x = None
        self.fg = None
        self.ncall = 0

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        if self.x is not None and np.array_equal(x, self.x):
            return self.fg
        f, g = self.f_g(x)
        self.ncall += 1
        g = np.array(g, dtype=np.float64)
        bad = ~np.isfinite(g)
        for i in np.nonzero(bad)[0]:
            h = self.eps * max(1.0, abs(x[i]))
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            g[i] = (self.f_g(xp)[0] - self.f_g(xm)[0]) / (2 * h)
        self.x, self.fg = x.copy(), (f, g)
        return self.fg

    def fun(self, x):
        return self(x)[0]

    def jac(self, x):
        return self(x)[1]


def fit_scipy(fcn, method="BFGS", maxiter=None, callback=None, check_large=1e7):
    """Minimise ``fcn`` from the current parameter values; returns a FitResult."""
    from scipy.optimize import minimize

    vm = fcn.vm
    f_g = vm.trans_fcn_grad(fcn.nll_grad)
    cached = CachedFG(f_g)
    x0 = vm.get_x_from_y()
    t0 = time.perf_counter()

    def cb(x):
        if np.any(np.abs(x) > check_large):
            raise LargeNumberError("parameter above {}".format(check_large))
        if callback is not None:
            callback(x, cached.fg[0])

    opts = {} if maxiter is None else {"maxiter": maxiter}
    res = minimize(cached.fun, x0, jac=cached.jac, method=method, callback=cb, options=opts)
    x = res.x
    yvals = x.copy()
    for i, name in enumerate(vm.trainable_vars):
        if name in vm.bnd_dic:
            yvals[i] = vm.bnd_dic[name].get_x2y(x[i])
    vm.set_all(yvals)
    vm.std_polar_all()
    return FitResult(vm.get_all_dic(), fcn, res.fun, ndf=len(x), success=res.success,
                     message=res.message, hess_inv=getattr(res, "hess_inv", None),
                     nfev=res.nfev, ngrad=getattr(res, "njev", 0) or 0,
                     time_s=time.perf_counter() - t0)


_NATIVE_MESSAGES = {0: "Optimization terminated successfully.",
                    1: "Maximum number of iterations has been exceeded.",
                    2: "Desired error not necessarily achieved due to precision loss."}


def fit_native(fcn, maxiter=None, gtol=1e-5):
    """Minimise ``fcn`` from the current parameter values with the engine's BFGS
    (tpa_fcn_minimize): the scheme of fit_scipy's scipy BFGS (same bound transforms, strong-
    Wolfe line search, |grad|_inf <= gtol in the unbounded variables) run natively, one engine
    evaluation per step (a mass that crosses a barrier clamp point rebuilds the caches inside
    the session). Falls back to fit_scipy when the FCN has no native session (Gaussian
    constraints, four-body floating lineshapes, ``fcn.native = False``). Returns a FitResult."""
    import ctypes

    from . import _lib

    vm = fcn.vm
    if fcn.gauss_constr:
        return fit_scipy(fcn, maxiter=maxiter)
    fcn._set(None)
    nat = fcn._native_session()
    if nat is None:
        return fit_scipy(fcn, maxiter=maxiter)
    y0 = np.ascontiguousarray(vm.get_all_val(), dtype=np.float64)
    y = np.zeros_like(y0)
    nll = ctypes.c_double()
    status = ctypes.c_int()
    nfev = ctypes.c_int64()
    niter = ctypes.c_int64()
    t0 = time.perf_counter()
    rc = _lib.lib().tpa_fcn_minimize(nat._h, _lib.dptr(y0), float(gtol),
                                     int(maxiter or 0), _lib.dptr(y), ctypes.byref(nll),
                                     ctypes.byref(status), ctypes.byref(nfev), ctypes.byref(niter))
    nat.after_call()
    fcn.n_call += int(nfev.value)
    fcn.model.n_eval += int(nfev.value)
    _lib.check(rc)
    vm.set_all(y)
    vm.std_polar_all()
    fcn.cached_nll = nll.value
    return FitResult(vm.get_all_dic(), fcn, nll.value, ndf=len(y), success=status.value == 0,
                     message=_NATIVE_MESSAGES.get(status.value, str(status.value)),
                     nfev=int(nfev.value), ngrad=int(nfev.value),
                     time_s=time.perf_counter() - t0)
