"""The per-call rebuild (tpa_eval_recompute, Model(evaluation="recompute")): the partial waves
built from the angle cache on every call, chunk by chunk, and streamed.

Bars: against the streaming pass over the resident FP64 cache, value 1e-12 relative and G 1e-11
of its largest entry (the same arithmetic, summed per chunk); through FCN, the oracle's NLL to
1e-6 and gradient to 1e-7 relative (the FP64 bar) and the streaming FCN to 1e-8, three-body and
four-body.
"""

import numpy as np
import pytest

from helpers import make_case, oracle_nll_grad
from test_gpu_half import case4  # noqa: F401  (the four-body fixture)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case1():
    return make_case("lc_pkpi_2res.yml", n_mc=6000, n_data=800, seed=11)


@pytest.fixture(scope="module")
def case2():
    return make_case("lc_pkpi_6res.yml", n_mc=6000, n_data=800, seed=21, weight_seed=3)


@pytest.mark.parametrize("which,n,chunk", [("case2", 1, 0), ("case2", 1000, 300),
                                           ("case2", 20000, 0), ("case1", 3001, 1000),
                                           ("case4", 3001, 700)])
def test_recompute_pass_matches_streaming_pass(which, n, chunk, request, gpu_available):
    from tf_pwa_amd import _lib

    case = request.getfixturevalue(which)
    conf = case["conf"]
    full = conf.data_from_p4(case["mc"])
    idx = np.arange(n) % full.n
    ev = full.select(idx)            # streamed: its cache is built
    fresh = full.select(idx)         # rebuilt per call: its cache is never built
    w = np.random.default_rng(n + 9).uniform(-0.5, 1.5, n)
    ev.set_weight(w)
    fresh.set_weight(w)
    amp = conf.get_amplitude()
    name = [v for v in conf.vm.variables if v.endswith("_mass")][0]
    m0 = conf.vm.get(name)
    try:
        for shift in (0.0, 0.01):
            conf.vm.set(name, m0 + shift)
            c = amp.coefficients()
            for mode in (_lib.MODE_DATA, _lib.MODE_MC):
                v, G = fresh.eval_recompute(c, mode, grad=True, chunk=chunk)
                v0, G0, _ = ev.eval(c, mode, grad=True)
                scale = max(1.0, np.abs(w).sum())
                assert abs(v - v0) <= 1e-12 * scale * max(1.0, abs(v0) / scale)
                np.testing.assert_allclose(G, G0, rtol=0, atol=1e-11 * np.abs(G0).max())
                vv, GG = fresh.eval_recompute(c, mode, grad=False, chunk=chunk)
                assert GG is None and abs(vv - v) <= 1e-14 * max(1.0, abs(v))
    finally:
        conf.vm.set(name, m0)


@pytest.mark.parametrize("which", ["case1", "case2", "case4"])
def test_recompute_fcn_matches_oracle(which, request, gpu_available):
    from tf_pwa_amd import FCN

    case = request.getfixturevalue(which)
    conf = case["conf"]
    data = conf.data_from_p4(case["data"], weight=case["wd"])
    mc = conf.data_from_p4(case["mc"])
    m = conf.get_model()
    m.evaluation = "recompute"
    m.recompute_chunk = 1500
    fr = FCN(m, data, mc)
    nll, g = fr.nll_grad()
    assert fr._native_session() is None
    n0, g0 = FCN(conf.get_model(), data, mc).nll_grad()
    assert abs(nll - n0) < 1e-8
    assert np.linalg.norm(g - g0) <= 1e-10 * np.linalg.norm(g0)
    o_nll, o_g = oracle_nll_grad(case)
    assert abs(nll - o_nll) < 1e-6, (nll, o_nll)
    assert np.linalg.norm(g - o_g) <= 1e-7 * np.linalg.norm(o_g)
