"""The helicity-halved FP32 copy (tpa_events_keep_half) and its pass (TPA_FLAG_F32H).

The copy keeps one helicity slot of each mirror pair of every partial wave in FP32 and one
phase per chain and event; the mirrored slots follow from h_k,-o = phi_k s_o rho conj(g_k,o).
Three bars:
- the relation itself on the FP64 cache: the copy's measured residual <= 1e-12, and the full
  cache rebuilt from the copy within FP32 rounding (3e-7 of the event's largest wave);
- the pass on the copy against the oracle's C restatement run on that rebuilt cache (the same
  FP32 inputs, FP64 arithmetic): value 1e-12 relative, G 1e-11 of its largest entry, f 1e-12;
- the mixed path with data_storage="fp32h" against the FP64 path and the oracle: |dNLL| < 1e-5
  and the gradient within 1e-6 relative at these sizes (the metric's bar is 1e-3).
"""

import ctypes

import numpy as np
import pytest

from helpers import make_case, oracle_nll_grad

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case1():
    return make_case("lc_pkpi_2res.yml", n_mc=6000, n_data=800, seed=11)


@pytest.fixture(scope="module")
def case2():
    return make_case("lc_pkpi_6res.yml", n_mc=6000, n_data=800, seed=21, weight_seed=3)


@pytest.fixture(scope="module")
def case4():
    from conftest import config_path
    from oracle import tfpwa_oracle as O
    from tf_pwa_amd import ConfigLoader
    from tf_pwa_amd.phasespace import generate_phsp

    path = config_path("jpsi_kkpipi_12ch.yml")
    om = O.Model(path)
    om.set_params(O.random_couplings(om, 17))
    conf = ConfigLoader(path, seed=17)
    conf.set_params(om.values)
    order = om.config["data"]["dat_order"]
    masses = [om.particles[n].mass for n in order]
    mc = dict(zip(order, generate_phsp(om.top.mass, masses, 4000, np.random.default_rng(1))))
    data = dict(zip(order, generate_phsp(om.top.mass, masses, 700, np.random.default_rng(2))))
    wd = np.random.default_rng(3).uniform(0.5, 1.5, 700)
    return dict(om=om, conf=conf, mc=mc, data=data, wd=wd, O=O)


def rebuild_from_half(ev):
    """The full cache [P, n, lam_pad] from the halved copy, in FP64 from its FP32 values."""
    g, rho, slots, mirrors, sgn, rot, kch = ev.half_copy()
    P, n, nh = g.shape
    phi = np.where(rot == 1, -1j, 1.0).astype(np.complex128)
    g = g.astype(np.complex128)
    rho = rho.astype(np.complex128)
    out = np.zeros((P, n, ev.amp.lam_pad), dtype=np.complex128)
    for q in range(nh):
        out[:, :, slots[q]] = phi[:, None] * g[:, :, q]
        out[:, :, mirrors[q]] = phi[:, None] * sgn[q] * rho[kch] * np.conj(g[:, :, q])
    return out


@pytest.mark.parametrize("which", ["case1", "case2", "case4"])
def test_half_copy_reproduces_the_mirrored_slots(which, request, gpu_available):
    case = request.getfixturevalue(which)
    conf = case["conf"]
    ev = conf.data_from_p4(case["mc"])
    ev.keep_half(True)
    assert 0.0 <= ev.half_residual() <= 1e-12
    h = ev.partial_waves()
    rec = rebuild_from_half(ev)
    scale = np.abs(h).max(axis=(0, 2))
    err = np.abs(rec - h).max(axis=(0, 2))
    assert np.all(err <= 3e-7 * np.maximum(scale, 1e-300))
    # every slot is covered: the stored slots and their mirrors
    g, rho, slots, mirrors = ev.half_copy()[:4]
    assert sorted(list(slots) + list(mirrors)) == list(range(conf.get_amplitude().lam))


def _cpu_pass(L, h, w, c, mode):
    dp = ctypes.POINTER(ctypes.c_double)
    h = np.ascontiguousarray(h)
    P, N, LG = h.shape
    cc = np.ascontiguousarray(c).view(np.float64)
    out = np.zeros(2 + 2 * P)
    assert L.cpu_eval(h.view(np.float64).ctypes.data_as(dp), w.ctypes.data_as(dp), N, P, LG,
                      cc.ctypes.data_as(dp), mode, out.ctypes.data_as(dp)) == 0
    return out


@pytest.mark.parametrize("which,n", [("case2", 1), ("case2", 257), ("case2", 20000),
                                     ("case2", 200003), ("case1", 1000), ("case4", 3001)])
def test_half_pass_matches_oracle_on_rebuilt_waves(which, n, request, gpu_available):
    """tpa_eval with TPA_FLAG_F32H against the oracle's C restatement on the cache rebuilt
    from the same copy; the copy follows a rebuild of the cache (a fixed mass moved)."""
    from oracle import build_cpu
    from tf_pwa_amd import _lib

    case = request.getfixturevalue(which)
    L = build_cpu.load()
    conf = case["conf"]
    full = conf.data_from_p4(case["mc"])
    ev = full.select(np.arange(n) % full.n)
    w = np.random.default_rng(n + 1).uniform(-0.5, 1.5, n)
    ev.set_weight(w)
    ev.keep_half(True)
    amp = conf.get_amplitude()
    name = [v for v in conf.vm.variables if v.endswith("_mass")][0]
    m0 = conf.vm.get(name)
    try:
        for shift in (0.0, 0.01):
            conf.vm.set(name, m0 + shift)
            c = amp.coefficients()
            for mode in (_lib.MODE_DATA, _lib.MODE_MC):
                v, G, f = ev.eval(c, mode, grad=True, fout=True, half=True)
                assert ev.half_residual() <= 1e-12
                h = rebuild_from_half(ev)
                out = _cpu_pass(L, h, w, c, 0 if mode == _lib.MODE_DATA else 1)
                scale = max(1.0, np.abs(w).sum())
                assert abs(out[0] - v) <= 1e-12 * scale * max(1.0, abs(v) / scale)
                np.testing.assert_allclose(out[2:].view(np.complex128), G, rtol=0,
                                           atol=1e-11 * max(1e-300, np.abs(G).max()))
                A = np.einsum("k,kel->el", c, h)
                np.testing.assert_allclose(f, (np.abs(A) ** 2).sum(1), rtol=1e-12)
    finally:
        conf.vm.set(name, m0)


@pytest.mark.parametrize("which", ["case1", "case2", "case4"])
def test_half_storage_mixed_path_close_to_fp64_and_oracle(which, request, gpu_available):
    from tf_pwa_amd import FCN

    case = request.getfixturevalue(which)
    conf = case["conf"]
    data = conf.data_from_p4(case["data"], weight=case["wd"])
    mc = conf.data_from_p4(case["mc"])
    n1, g1 = FCN(conf.get_model(), data, mc).nll_grad()
    mh = conf.get_model()
    mh.data_storage = "fp32h"
    fs = FCN(mh, data, mc, precision="mixed")
    n2, g2 = fs.nll_grad()
    assert fs._native is not None
    assert abs(n1 - n2) < 1e-5, abs(n1 - n2)
    assert np.linalg.norm(g1 - g2) <= 1e-6 * np.linalg.norm(g1)
    fs.native = False  # the Python host assembly reads the same halved copies
    n3, g3 = fs.nll_grad()
    assert abs(n3 - n2) <= 1e-12 * abs(n2)
    assert np.abs(g3 - g2).max() <= 1e-10 * np.abs(g2).max()
    o_nll, o_g = oracle_nll_grad(case)
    assert abs(n2 - o_nll) < 1e-5
    assert np.linalg.norm(g2 - o_g) <= 1e-6 * np.linalg.norm(o_g)


@pytest.mark.parametrize("which,n", [("case2", 1), ("case2", 257), ("case2", 20000),
                                     ("case1", 1000), ("case4", 3001)])
def test_half64_pass_matches_the_fp64_cache_pass(which, n, request, gpu_available):
    """TPA_FLAG_F64H: the halved FP64 copy reproduces the FP64 cache's pass to rounding
    (value 1e-12 relative, G 1e-11 of its largest entry, f 1e-12), both modes, after a rebuild."""
    from tf_pwa_amd import _lib

    case = request.getfixturevalue(which)
    conf = case["conf"]
    full = conf.data_from_p4(case["mc"])
    ev = full.select(np.arange(n) % full.n)
    w = np.random.default_rng(n + 5).uniform(-0.5, 1.5, n)
    ev.set_weight(w)
    ev.keep_half(False, fp64=True)
    amp = conf.get_amplitude()
    name = [v for v in conf.vm.variables if v.endswith("_mass")][0]
    m0 = conf.vm.get(name)
    try:
        for shift in (0.0, 0.01):
            conf.vm.set(name, m0 + shift)
            c = amp.coefficients()
            for mode in (_lib.MODE_DATA, _lib.MODE_MC):
                v, G, f = ev.eval(c, mode, grad=True, fout=True, half64=True)
                assert 0.0 <= ev.half_residual() <= 1e-12
                v0, G0, f0 = ev.eval(c, mode, grad=True, fout=True)
                scale = max(1.0, np.abs(w).sum())
                assert abs(v - v0) <= 1e-12 * scale * max(1.0, abs(v0) / scale)
                np.testing.assert_allclose(G, G0, rtol=0, atol=1e-11 * np.abs(G0).max())
                np.testing.assert_allclose(f, f0, rtol=1e-12)
    finally:
        conf.vm.set(name, m0)


@pytest.mark.parametrize("which", ["case1", "case2", "case4"])
def test_half64_storage_fp64_and_mixed_paths(which, request, gpu_available):
    """FP64 path with halved FP64 data and MC copies, and the mixed path with halved FP64
    data: the FP64 cache's results to rounding (|dNLL| < 1e-8, gradient 1e-10 relative), the
    oracle at the FP64 bar (1e-6, 1e-7)."""
    from tf_pwa_amd import FCN

    case = request.getfixturevalue(which)
    conf = case["conf"]
    data = conf.data_from_p4(case["data"], weight=case["wd"])
    mc = conf.data_from_p4(case["mc"])
    n1, g1 = FCN(conf.get_model(), data, mc).nll_grad()
    m = conf.get_model()
    m.data_storage = "fp64h"
    m.mc_storage = "fp64h"
    fs = FCN(m, data, mc)
    n2, g2 = fs.nll_grad()
    assert fs._native is not None
    assert abs(n1 - n2) < 1e-8, abs(n1 - n2)
    assert np.linalg.norm(g1 - g2) <= 1e-10 * np.linalg.norm(g1)
    fs.native = False  # the Python host assembly reads the same halved copies
    n3, g3 = fs.nll_grad()
    assert abs(n3 - n2) <= 1e-12 * abs(n2)
    assert np.abs(g3 - g2).max() <= 1e-10 * np.abs(g2).max()
    mm = conf.get_model()
    mm.data_storage = "fp64h"
    n4, g4 = FCN(mm, data, mc, precision="mixed").nll_grad()
    assert abs(n1 - n4) < 1e-8
    assert np.linalg.norm(g1 - g4) <= 1e-10 * np.linalg.norm(g1)
    o_nll, o_g = oracle_nll_grad(case)
    assert abs(n2 - o_nll) < 1e-6
    assert np.linalg.norm(g2 - o_g) <= 1e-7 * np.linalg.norm(o_g)



@pytest.mark.parametrize("n", [257, 70001])
@pytest.mark.parametrize("cfg,planned", [("lc_pkpi_7res.yml", 8), ("lc_pkpi_9res.yml", 0)])
def test_half_pass_larger_models(cfg, planned, n, gpu_available):
    """The halved FP32 pass at 26 partial waves in seven chains (the chain-grouped pass with
    eight waves per lane) and at 32 waves in nine chains (more chains than the eight lane runs:
    no plan fits, the per-wave pass), against the oracle's C restatement on the cache rebuilt
    from the same copy (value 1e-12 relative, G 1e-11 of its largest entry, f 1e-12), both
    modes. The lane plan covers every wave once."""
    import os

    from oracle import build_cpu
    from tf_pwa_amd import ConfigLoader, _lib
    from tf_pwa_amd.phasespace import generate_phsp

    L = build_cpu.load()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs", cfg)
    conf = ConfigLoader(path, seed=3)
    order = conf.dat_order
    masses = [conf.particles[p].mass for p in order]
    top = conf.particles[conf.top_name].mass
    p4 = dict(zip(order, generate_phsp(top, masses, n, np.random.default_rng(n))))
    ev = conf.data_from_p4(p4)
    w = np.random.default_rng(n + 5).uniform(-0.5, 1.5, n)
    ev.set_weight(w)
    ev.keep_half(True)
    assert 0.0 <= ev.half_residual() <= 1e-12
    amp = conf.get_amplitude()
    pml, plan = ev.half_plan()
    assert pml == planned
    if planned:
        waves = [k for k in plan[:4 * pml] if k >= 0]
        assert sorted(waves) == list(range(amp.P))
    c = amp.coefficients()
    h = rebuild_from_half(ev)
    for mode in (_lib.MODE_DATA, _lib.MODE_MC):
        v, G, f = ev.eval(c, mode, grad=True, fout=True, half=True)
        out = _cpu_pass(L, h, w, c, 0 if mode == _lib.MODE_DATA else 1)
        scale = max(1.0, np.abs(w).sum())
        assert abs(out[0] - v) <= 1e-12 * scale * max(1.0, abs(v) / scale)
        np.testing.assert_allclose(out[2:].view(np.complex128), G, rtol=0,
                                   atol=1e-11 * max(1e-300, np.abs(G).max()))
        A = np.einsum("k,kel->el", c, h)
        np.testing.assert_allclose(f, (np.abs(A) ** 2).sum(1), rtol=1e-12)


def test_sampled_pass_timing(case2, gpu_available):
    """tpa_timing_enable(k) times one tpa_eval pass in k, with events from the chain-grouped
    pass's own dispatch; timed and untimed passes return the same sums."""
    from tf_pwa_amd import _lib

    L = _lib.lib()
    conf = case2["conf"]
    full = conf.data_from_p4(case2["mc"])
    n = 20000
    ev = full.select(np.arange(n) % full.n)
    ev.set_weight(np.random.default_rng(5).uniform(0.5, 1.5, n))
    ev.keep_half(True)
    assert ev.half_plan()[0] > 0  # the chain-grouped pass serves this model
    c = conf.get_amplitude().coefficients()
    v0, G0, _ = ev.eval(c, _lib.MODE_DATA, grad=True, half=True)
    ms, nl = ctypes.c_double(0), ctypes.c_int64(0)
    try:
        for every, calls, timed in ((1, 4, 4), (3, 7, 3), (8, 5, 1)):
            _lib.check(L.tpa_timing_reset())
            _lib.check(L.tpa_timing_enable(every))
            for _ in range(calls):
                v, G, _ = ev.eval(c, _lib.MODE_DATA, grad=True, half=True)
                assert abs(v - v0) <= 1e-14 * abs(v0)
                np.testing.assert_allclose(G, G0, rtol=0, atol=1e-14 * np.abs(G0).max())
            _lib.check(L.tpa_timing_enable(0))
            _lib.check(L.tpa_timing_read(ctypes.byref(ms), ctypes.byref(nl)))
            assert nl.value == timed
            assert 0.0 < ms.value < 1e3 * timed
    finally:
        L.tpa_timing_enable(0)
        L.tpa_timing_reset()
