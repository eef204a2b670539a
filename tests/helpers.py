"""Shared fixtures of the parity tests: the same seeded events and parameters for the
oracle and the engine."""

import numpy as np
import torch

from conftest import config_path


def make_case(cfg, n_mc, n_data, seed=7, weight_seed=None):
    from oracle import tfpwa_oracle as O
    from tf_pwa_amd import ConfigLoader

    path = config_path(cfg)
    om = O.Model(path)
    om.set_params(O.random_couplings(om, seed))
    conf = ConfigLoader(path, seed=seed)
    conf.set_params(om.values)
    mc = O.gen_mc(om, n_mc, seed + 1)
    data = O.gen_data(om, n_data, seed + 2, batch=max(20000, 4 * n_data))
    wd = None
    if weight_seed is not None:
        wd = np.random.default_rng(weight_seed).uniform(0.5, 1.5, n_data)
    return dict(om=om, conf=conf, mc=mc, data=data, wd=wd, O=O)


def oracle_nll_grad(case, hessian=False):
    O, om = case["O"], case["om"]
    ad = O.cal_angle(om, case["data"])
    am = O.cal_angle(om, case["mc"])
    n_d = len(next(iter(case["data"].values())))
    n_m = len(next(iter(case["mc"].values())))
    w, wm = O.fcn_weights(n_d, case["wd"], n_m)
    if hessian:
        return O.nll_grad_hessian(om, ad, am, w, wm)
    return O.nll_grad(om, ad, am, w, wm)


def oracle_f(case, key):
    O, om = case["O"], case["om"]
    ang = O.cal_angle(om, case[key])
    return O.sum_amp(om, ang, om.var_tensors()).numpy(), ang


def oracle_angle_layout(case, ang, amp):
    """The oracle's angles in the engine's angle-cache layout [field][event]."""
    om = case["om"]
    names = [p.name for p in amp.finals]
    n = ang["mass"][ang["top_key"]].shape[0]
    out = np.zeros((18 * len(amp.topo) + 4, n))
    keys = list(ang["topo"].keys())
    for t, T in enumerate(amp.topo):
        rep = amp.reps[t]
        tk = rep.topology_key()
        assert tk in keys
        fy, fz, fx = names[T["y"]], names[T["z"]], names[T["x"]]
        kR = frozenset([fy, fz])
        top = ang["top_key"]
        a = ang["topo"][tk]["ang"]
        out[18 * t + 0] = ang["mass"][kR].numpy()
        out[18 * t + 1] = a[top][kR]["alpha"].numpy()
        out[18 * t + 2] = a[top][kR]["beta"].numpy()
        out[18 * t + 3] = a[top][frozenset([fx])]["alpha"].numpy()
        out[18 * t + 4] = a[top][frozenset([fx])]["beta"].numpy()
        out[18 * t + 5] = a[kR][frozenset([fy])]["alpha"].numpy()
        out[18 * t + 6] = a[kR][frozenset([fy])]["beta"].numpy()
        out[18 * t + 7] = a[kR][frozenset([fz])]["alpha"].numpy()
        out[18 * t + 8] = a[kR][frozenset([fz])]["beta"].numpy()
        for f, nm in enumerate(names):
            al = ang["aligned"][tk].get(nm)
            if al is not None:
                out[18 * t + 9 + 3 * f] = al["alpha"].numpy()
                out[18 * t + 10 + 3 * f] = al["beta"].numpy()
                out[18 * t + 11 + 3 * f] = al["gamma"].numpy()
    g = 18 * len(amp.topo)
    out[g] = ang["mass"][ang["top_key"]].numpy()
    for f, nm in enumerate(names):
        out[g + 1 + f] = ang["mass"][frozenset([nm])].numpy()
    return out

