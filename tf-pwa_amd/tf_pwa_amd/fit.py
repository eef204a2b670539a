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
