"""Engine (HIP, FP64) against the oracle on identical seeded events and parameters.

Tolerances: angles and per-event |A|^2 are compared to 1e-9 (relative for |A|^2); the NLL
to 1e-6 absolute and the gradient to 1e-7 relative to its norm, the FP64 bar of the
reference's own float64 path.
"""

import numpy as np
import pytest

from helpers import make_case, oracle_angle_layout, oracle_f, oracle_nll_grad

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case1():
    return make_case("lc_pkpi_2res.yml", n_mc=6000, n_data=800, seed=11)


@pytest.fixture(scope="module")
def case2():
    return make_case("lc_pkpi_6res.yml", n_mc=6000, n_data=800, seed=21, weight_seed=3)


def _half_angle_close(a, b, atol):
    """alpha folded by 2 pi differs only at branch cuts; compare e^{i a/2}."""
    return np.abs(np.exp(0.5j * a) - np.exp(0.5j * b)).max() < atol


@pytest.mark.parametrize("which", ["case1", "case2"])
def test_angles_match_oracle(which, request, gpu_available):
    case = request.getfixturevalue(which)
    conf = case["conf"]
    amp = conf.get_amplitude()
    ev = conf.data_from_p4(case["mc"])
    got = ev.angles()
    _, ang = oracle_f(case, "mc")
    ref = oracle_angle_layout(case, ang, amp)
    nt = len(amp.topo)
    for t in range(nt):
        b = 18 * t
        np.testing.assert_allclose(got[b], ref[b], rtol=1e-12, atol=1e-12)
        for fld in (1, 3, 5, 7):
            assert _half_angle_close(got[b + fld], ref[b + fld], 1e-9), (t, fld)
            np.testing.assert_allclose(got[b + fld + 1], ref[b + fld + 1], atol=1e-9)
        for f in range(3):
            al, be, ga = got[b + 9 + 3 * f: b + 12 + 3 * f]
            ral, rbe, rga = ref[b + 9 + 3 * f: b + 12 + 3 * f]
            np.testing.assert_allclose(be, rbe, atol=1e-7)
            assert _half_angle_close(al + ga, ral + rga, 1e-7)
            assert _half_angle_close(al - ga, ral - rga, 1e-7)
    np.testing.assert_allclose(got[18 * nt:], ref[18 * nt:], rtol=1e-13)


@pytest.mark.parametrize("which", ["case1", "case2"])
def test_amp2_matches_oracle(which, request, gpu_available):
    case = request.getfixturevalue(which)
    conf = case["conf"]
    ev = conf.data_from_p4(case["mc"])
    f = conf.get_amplitude()(ev)
    ref, _ = oracle_f(case, "mc")
    assert np.all(f >= 0)
    np.testing.assert_allclose(f, ref, rtol=1e-9, atol=1e-12 * ref.max())


@pytest.mark.parametrize("which", ["case1", "case2"])
def test_nll_grad_matches_oracle(which, request, gpu_available):
    from tf_pwa_amd import FCN

    case = request.getfixturevalue(which)
    conf = case["conf"]
    data = conf.data_from_p4(case["data"], weight=case["wd"])
    mc = conf.data_from_p4(case["mc"])
    fcn = FCN(conf.get_model(), data, mc)
    nll, g = fcn.nll_grad()
    ref_nll, ref_g = oracle_nll_grad(case)
    assert abs(nll - ref_nll) < 1e-6, (nll, ref_nll)
    assert np.linalg.norm(g - ref_g) <= 1e-7 * np.linalg.norm(ref_g), (g, ref_g)
    # value-only path agrees with the gradient path
    assert abs(fcn(None) - nll) < 1e-9 * max(1.0, abs(nll))


@pytest.mark.parametrize("which", ["case1", "case2"])
def test_exact_hessian_matches_oracle(which, request, gpu_available):
    """Second-moment passes + coupling chain rule against the oracle's autograd Hessian
    (FP64 path), and the mixed path's Hessian against the FP64 one."""
    from tf_pwa_amd import FCN

    case = request.getfixturevalue(which)
    conf = case["conf"]
    data = conf.data_from_p4(case["data"], weight=case["wd"])
    mc = conf.data_from_p4(case["mc"])
    fcn = FCN(conf.get_model(), data, mc)
    nll, g, H = fcn.nll_grad_hessian()
    ref_nll, ref_g, ref_H = oracle_nll_grad(case, hessian=True)
    assert abs(nll - ref_nll) < 1e-6
    assert np.linalg.norm(g - ref_g) <= 1e-7 * np.linalg.norm(ref_g)
    scale = np.abs(ref_H).max()
    assert np.abs(H - ref_H).max() <= 1e-7 * scale, np.abs(H - ref_H).max() / scale
    # the engine's own gradient path agrees with the second-moment pass
    n2, g2 = fcn.nll_grad()
    assert abs(n2 - nll) < 1e-9 * abs(nll) and np.abs(g2 - g).max() < 1e-9 * np.abs(g).max()
    mix = FCN(conf.get_model(), data, mc, precision="mixed")
    _, _, Hm = mix.nll_grad_hessian()
    assert np.abs(Hm - H).max() <= 1e-4 * scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_native_session_matches_python_assembly(case2, precision, gpu_available):
    """FCN's one-call native session (tpa_fcn_*) against the Python host assembly of the same
    engine passes (Model._nll_grad), at several parameter points and after a weight change."""
    from tf_pwa_amd import FCN

    conf = case2["conf"]
    data = conf.data_from_p4(case2["data"], weight=case2["wd"])
    mc = conf.data_from_p4(case2["mc"])
    nat = FCN(conf.get_model(), data, mc, precision=precision)
    ref = FCN(conf.get_model(), data, mc, precision=precision)
    ref.native = False
    x0 = np.array(conf.vm.get_all_val())
    rng = np.random.default_rng(12)
    for _ in range(3):
        x = x0 + rng.normal(scale=0.2, size=x0.size)
        n1, g1 = nat.nll_grad(x)
        n2, g2 = ref.nll_grad(x)
        assert abs(n1 - n2) <= 1e-12 * abs(n2)
        assert np.abs(g1 - g2).max() <= 1e-10 * np.abs(g2).max()
        assert abs(nat(x) - n1) <= 1e-12 * abs(n1)
    assert nat._native is not None and ref._native is None
    conf.vm.set_all(x0)


@pytest.mark.parametrize("name", ["lc_pkpi_2res", "lc_pkpi_6res", "jpsi_kkpipi_12ch"])
def test_engine_matches_committed_golden_vectors(name, gpu_available):
    """The FP64 engine on the committed golden inputs against the oracle's committed outputs:
    per-event |A|^2 (1e-9 relative), NLL (1e-6 absolute), gradient (1e-7 relative) and the
    exact Hessian (1e-7 of its largest entry)."""
    import json
    import os

    from conftest import config_path
    from tf_pwa_amd import FCN, ConfigLoader

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "golden_{}.json".format(name))) as fh:
        fx = json.load(fh)
    inp, ref = fx["inputs"], fx["outputs"]
    conf = ConfigLoader(config_path(inp["config"]), seed=0)
    conf.set_params(inp["params"])
    assert conf.vm.trainable_vars == ref["trainable"]
    data = conf.data_from_p4({k: np.array(v) for k, v in inp["data_p4"].items()},
                             weight=np.array(inp["data_weight"]))
    mc = conf.data_from_p4({k: np.array(v) for k, v in inp["mc_p4"].items()})
    amp = conf.get_amplitude()
    np.testing.assert_allclose(amp(data), ref["f_data"], rtol=1e-9)
    np.testing.assert_allclose(amp(mc), ref["f_mc"], rtol=1e-9)
    fcn = FCN(conf.get_model(), data, mc)
    nll, g = fcn.nll_grad()
    assert abs(nll - ref["nll"]) < 1e-6
    rg = np.array(ref["grad"])
    assert np.linalg.norm(g - rg) <= 1e-7 * np.linalg.norm(rg)
    _, _, H = fcn.nll_grad_hessian()
    rH = np.array(ref["hessian"])
    assert np.abs(H - rH).max() <= 1e-7 * np.abs(rH).max()
    p = np.random.default_rng(0).normal(size=len(g))
    g2, Hp = fcn.grad_hessp(np.array(conf.vm.get_all_val()), p)
    assert np.linalg.norm(g2 - rg) <= 1e-7 * np.linalg.norm(rg)
    assert np.abs(Hp - rH @ p).max() <= 1e-7 * np.abs(rH @ p).max()


def test_gram_matches_partial_waves(case2, gpu_available):
    """FP32 MFMA Gram against the FP64 Gram of the same partial waves (read back)."""
    conf = case2["conf"]
    mc = conf.data_from_p4(case2["mc"])
    w = np.random.default_rng(1).uniform(0.5, 1.5, mc.n)
    mc.set_weight(w)
    h = mc.partial_waves()  # [P][n][lam_pad]
    ref = np.einsum("e,kel,jel->kj", w, np.conj(h), h)
    got = mc.gram("fp32")
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() < 1e-6 * scale
    np.testing.assert_allclose(got, np.conj(got.T), atol=1e-12 * scale)
    # c^H I c equals the engine's direct weighted MC sum
    c = conf.get_amplitude().coefficients()
    from tf_pwa_amd import _lib

    v, G, _ = mc.eval(c, _lib.MODE_MC, grad=True)
    assert abs(np.real(np.vdot(c, got @ c)) - v) < 1e-6 * abs(v)
    np.testing.assert_allclose(np.conj(got @ c), G, rtol=0, atol=1e-6 * np.abs(G).max())


@pytest.mark.parametrize("which", ["case1", "case2", "case4"])
def test_gram_f64_matches_partial_waves(which, request, gpu_available):
    """FP64 MFMA Gram (one to four 16-column tiles: 2, 6 resonances and the 12-chain four-body
    model, ragged last row block) against the FP64 Gram of the same partial waves read back."""
    from tf_pwa_amd import _lib

    case = request.getfixturevalue(which)
    conf = case["conf"]
    mc = conf.data_from_p4(case["mc"])
    w = np.random.default_rng(1).uniform(0.5, 1.5, mc.n)
    mc.set_weight(w)
    h = mc.partial_waves()
    ref = np.einsum("e,kel,jel->kj", w, np.conj(h), h)
    got = mc.gram("fp64")
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() < 1e-12 * scale
    np.testing.assert_array_equal(got, np.conj(got.T))
    c = conf.get_amplitude().coefficients()
    v, G, _ = mc.eval(c, _lib.MODE_MC, grad=True)
    assert abs(np.real(np.vdot(c, got @ c)) - v) < 1e-12 * abs(v)
    np.testing.assert_allclose(np.conj(got @ c), G, rtol=0, atol=1e-12 * np.abs(G).max())


def test_mixed_path_close_to_fp64(case2, gpu_available):
    from tf_pwa_amd import FCN

    conf = case2["conf"]
    data = conf.data_from_p4(case2["data"], weight=case2["wd"])
    mc = conf.data_from_p4(case2["mc"])
    f64 = FCN(conf.get_model(), data, mc)
    n1, g1 = f64.nll_grad()
    # FP64-MFMA Gram (default): the same sums in another order
    n2, g2 = FCN(conf.get_model(), data, mc, precision="mixed").nll_grad()
    assert abs(n1 - n2) < 1e-9
    assert np.linalg.norm(g1 - g2) < 1e-10 * np.linalg.norm(g1)
    # FP32-MFMA Gram
    model32 = conf.get_model()
    model32.gram_precision = "fp32"
    n3, g3 = FCN(model32, data, mc, precision="mixed").nll_grad()
    assert abs(n1 - n3) < 1e-3
    assert np.linalg.norm(g1 - g3) < 1e-4 * np.linalg.norm(g1)
    # FP32-stored data waves (FP64 sums): the bar of the mixed path, |dNLL| < 1e-3, with the
    # rounding's expected size (~1e-7 per event, random signs) well inside it
    m32 = conf.get_model()
    m32.data_storage = "fp32"
    fs = FCN(m32, data, mc, precision="mixed")
    n4, g4 = fs.nll_grad()
    assert fs._native is not None
    assert abs(n1 - n4) < 1e-5, abs(n1 - n4)
    assert np.linalg.norm(g1 - g4) < 1e-6 * np.linalg.norm(g1)
    fs.native = False  # the Python host assembly reads the same FP32 copies
    n5, g5 = fs.nll_grad()
    assert abs(n5 - n4) <= 1e-12 * abs(n4)
    assert np.abs(g5 - g4).max() <= 1e-10 * np.abs(g4).max()


@pytest.mark.parametrize("n", [1, 257, 20000])
def test_fp32_cache_pass_matches_oracle_on_rounded_waves(case2, n, gpu_available):
    """tpa_eval with TPA_FLAG_F32 computes in FP64 from the FP32-rounded partial waves: the
    oracle's C restatement on the same rounded waves gives the same sums to FP64 rounding.
    The copy follows a rebuild of the cache (a fixed mass moved)."""
    import ctypes

    from oracle import build_cpu
    from tf_pwa_amd import _lib

    L = build_cpu.load()
    conf = case2["conf"]
    full = conf.data_from_p4(case2["mc"])
    ev = full.select(np.arange(n) % full.n)
    w = np.random.default_rng(n + 1).uniform(-0.5, 1.5, n)
    ev.set_weight(w)
    ev.keep_f32(True)
    amp = conf.get_amplitude()
    dp = ctypes.POINTER(ctypes.c_double)
    name = [v for v in conf.vm.variables if v.endswith("_mass")][0]
    m0 = conf.vm.get(name)
    try:
        for shift in (0.0, 0.01):
            conf.vm.set(name, m0 + shift)
            c = amp.coefficients()
            h = ev.partial_waves().astype(np.complex64).astype(np.complex128)
            h = np.ascontiguousarray(h)
            P, N, LG = h.shape
            cc = np.ascontiguousarray(c).view(np.float64)
            for mode in (_lib.MODE_DATA, _lib.MODE_MC):
                v, G, f = ev.eval(c, mode, grad=True, fout=True, f32=True)
                out = np.zeros(2 + 2 * P)
                assert L.cpu_eval(h.view(np.float64).ctypes.data_as(dp), w.ctypes.data_as(dp), N,
                                  P, LG, cc.ctypes.data_as(dp), 0 if mode == _lib.MODE_DATA else 1,
                                  out.ctypes.data_as(dp)) == 0
                scale = max(1.0, np.abs(w).sum())
                assert abs(out[0] - v) <= 1e-12 * scale * max(1.0, abs(v) / scale)
                np.testing.assert_allclose(out[2:].view(np.complex128), G, rtol=0,
                                           atol=1e-11 * max(1e-300, np.abs(G).max()))
                A = np.einsum("k,kel->el", c, h)
                np.testing.assert_allclose(f, (np.abs(A) ** 2).sum(1), rtol=1e-12)
    finally:
        conf.vm.set(name, m0)


@pytest.mark.parametrize("cfg,res", [
    ("lc_pkpi_2res.yml", ("Kst", "L1520")),
    ("lc_pkpi_6res.yml", ("Kst", "K0st", "L1520", "D1232", "L1405", "L1820")),
])
def test_floating_lineshapes_match_oracle(cfg, res, gpu_available):
    """Masses and widths trainable: per-event lineshapes with their derivatives, against the
    oracle's autograd gradient (includes the L(1405) below-threshold q0 clamp)."""
    import yaml

    from conftest import config_path
    from oracle import tfpwa_oracle as O
    from tf_pwa_amd import FCN, ConfigLoader

    conf_d = yaml.safe_load(open(config_path(cfg)))
    for r in res:
        conf_d["particle"][r]["float"] = "mg"
    om = O.Model(conf_d)
    om.set_params(O.random_couplings(om, 31))
    conf = ConfigLoader(conf_d, seed=31)
    conf.set_params(om.values)
    assert conf.vm.trainable_vars == om.trainable
    assert conf.get_amplitude().shape_mode
    mc = O.gen_mc(om, 5000, 32)
    data = O.gen_data(om, 600, 33, batch=20000)
    fcn = FCN(conf.get_model(), conf.data_from_p4(data), conf.data_from_p4(mc))
    nll, g = fcn.nll_grad()
    w, wm = O.fcn_weights(600, None, 5000)
    ref_nll, ref_g = O.nll_grad(om, O.cal_angle(om, data), O.cal_angle(om, mc), w, wm)
    assert abs(nll - ref_nll) < 1e-6, (nll, ref_nll)
    assert np.linalg.norm(g - ref_g) <= 1e-7 * np.linalg.norm(ref_g), (g - ref_g)


def test_fit_fractions_from_the_gram_matrix(case2, gpu_available):
    """Gram-matrix fit fractions against direct MC sums of the masked amplitude (the
    definition of tf_pwa/fitfractions.py:178-258), the identity sum FF_R + sum FF_RS = 1,
    gradients against central differences, and errors from the exact-Hessian error matrix."""
    from tf_pwa_amd import FCN, _lib
    from tf_pwa_amd.fitfractions import cal_fitfractions, gram_fp64

    conf = case2["conf"]
    amp = conf.get_amplitude()
    mc = conf.data_from_p4(case2["mc"])
    I = gram_fp64(mc)
    c = amp.coefficients()
    v, _, _ = mc.eval(c, _lib.MODE_MC)
    assert abs(np.real(np.vdot(c, I @ c)) - v) < 1e-11 * v
    np.testing.assert_allclose(I, I.conj().T, rtol=0, atol=1e-13 * np.abs(I).max())
    frac, grad = cal_fitfractions(amp, mc)
    names = [ch["R"].name for ch in amp.chains]
    for r in set(names):
        cm = np.array([c[k] if names[w.chain] == r else 0 for k, w in enumerate(amp.waves)])
        vr, _, _ = mc.eval(cm, _lib.MODE_MC)
        assert abs(frac[r] - vr / v) < 1e-10, (r, frac[r], vr / v)
    total = sum(val for key, val in frac.items() if key != "sum_diag")
    assert abs(total - 1.0) < 1e-10
    x0 = np.array(conf.vm.get_all_val())
    for key in ("Kst", ("L1520", "Kst")):
        for i in (0, 5, 17):
            h = 1e-6
            xp, xm = x0.copy(), x0.copy()
            xp[i] += h
            xm[i] -= h
            conf.vm.set_all(xp)
            fp = cal_fitfractions(amp, gram=I)[0][key]
            conf.vm.set_all(xm)
            fm = cal_fitfractions(amp, gram=I)[0][key]
            conf.vm.set_all(x0)
            assert abs((fp - fm) / (2 * h) - grad[key][i]) < 1e-6
    data = conf.data_from_p4(case2["data"], weight=case2["wd"])
    err = conf.get_params_error(fcn=FCN(conf.get_model(), data, mc))
    assert set(err) == set(conf.vm.trainable_vars) and all(np.isfinite(list(err.values())))
    frac2, ferr = conf.cal_fitfractions(mc)  # MC weights now unit-sum (set by FCN)
    assert all(abs(frac2[k] - frac[k]) < 1e-12 for k in frac)
    assert all(e >= 0 and np.isfinite(e) for e in ferr.values())
    # the FP32-MFMA Gram gives the same fractions to the mixed path's accuracy
    mc.set_weight(np.ones(mc.n))
    frac32, _ = cal_fitfractions(amp, gram=mc.gram("fp32"))
    for key, val in frac.items():
        assert abs(frac32[key] - val) < 1e-5


@pytest.fixture(scope="module")
def case4():
    """J/psi -> K+ K- pi+ pi- (12 chains, 6 topologies): phase-space MC and weighted
    phase-space "data" (the likelihood does not assume the data follow the model)."""
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


def test_four_body_angles_match_oracle(case4, gpu_available):
    conf, O, om = case4["conf"], case4["O"], case4["om"]
    amp = conf.get_amplitude()
    ev = conf.data_from_p4(case4["mc"])
    got = ev.angles()
    ang = O.cal_angle(om, case4["mc"])
    names = [p.name for p in amp.finals]

    def key(mask):
        return frozenset(n for i, n in enumerate(names) if mask >> i & 1)

    for t, (T, rep) in enumerate(zip(amp.topo, amp.reps)):
        a = ang["topo"][rep.topology_key()]["ang"]
        for v in range(3):
            b = 15 * t + 5 * v
            np.testing.assert_allclose(got[b], ang["mass"][key(T["core"][v])].numpy(), rtol=1e-12)
            for o in range(2):
                ref = a[key(T["core"][v])][key(T["out"][v][o])]
                al, be = got[b + 1 + 2 * o], got[b + 2 + 2 * o]
                assert np.abs(np.exp(0.5j * al) - np.exp(0.5j * ref["alpha"].numpy())).max() < 1e-9
                np.testing.assert_allclose(be, ref["beta"].numpy(), atol=1e-9)


def test_four_body_amplitude_nll_gradient_hessian_match_oracle(case4, gpu_available):
    from tf_pwa_amd import FCN

    conf = case4["conf"]
    amp = conf.get_amplitude()
    mc = conf.data_from_p4(case4["mc"])
    ref_f, _ = oracle_f(case4, "mc")
    np.testing.assert_allclose(amp(mc), ref_f, rtol=1e-9, atol=1e-12 * ref_f.max())
    data = conf.data_from_p4(case4["data"], weight=case4["wd"])
    fcn = FCN(conf.get_model(), data, mc)
    nll, g = fcn.nll_grad()
    ref_nll, ref_g, ref_H = oracle_nll_grad(case4, hessian=True)
    assert abs(nll - ref_nll) < 1e-6, (nll, ref_nll)
    assert np.linalg.norm(g - ref_g) <= 1e-7 * np.linalg.norm(ref_g)
    _, _, H = fcn.nll_grad_hessian()
    assert np.abs(H - ref_H).max() <= 1e-7 * np.abs(ref_H).max()
    model32 = conf.get_model()
    model32.gram_precision = "fp32"
    n2, g2 = FCN(model32, data, mc, precision="mixed").nll_grad()
    assert abs(n2 - nll) < 1e-3 and np.linalg.norm(g2 - g) < 1e-4 * np.linalg.norm(g)
    n3, g3 = FCN(conf.get_model(), data, mc, precision="mixed").nll_grad()
    assert abs(n3 - nll) < 1e-9 and np.linalg.norm(g3 - g) < 1e-10 * np.linalg.norm(g)


@pytest.mark.parametrize("cfg", ["lc_pkpi_6res.yml", "jpsi_kkpipi_12ch.yml"])
def test_device_phase_space_and_selection(cfg, gpu_available):
    """Device GENBOD sample: four-momentum conservation and masses, the same invariant-mass
    distributions as the host generators (two-sample KS), reproducible per (seed, index) and
    in pieces, angles equal to those of an event set made from the same four-momenta; a
    selected subset carries exactly the angles of its source events."""
    from scipy.stats import ks_2samp

    from conftest import config_path
    from tf_pwa_amd import ConfigLoader, EventSet
    from tf_pwa_amd.data import p4_array
    from tf_pwa_amd.phasespace import generate_phsp

    conf = ConfigLoader(config_path(cfg), seed=2)
    amp = conf.get_amplitude()
    nfin = len(amp.finals)
    masses = [p.mass for p in amp.finals]
    top = amp.decay_group.top.mass
    n = 30000
    ev, p4 = EventSet.generate(amp, n, seed=1234, return_p4=True)
    p = [p4[4 * f:4 * f + 4].T for f in range(nfin)]
    tot = sum(p)
    np.testing.assert_allclose(tot[:, 0], top, atol=1e-12)
    np.testing.assert_allclose(tot[:, 1:], 0.0, atol=1e-12)
    for x, m in zip(p, masses):
        np.testing.assert_allclose(np.sqrt(np.abs(x[:, 0] ** 2 - (x[:, 1:] ** 2).sum(1))), m,
                                   atol=1e-10)
    ref = generate_phsp(top, masses, n, np.random.default_rng(5))

    def m2(a, b):
        s = a + b
        return s[:, 0] ** 2 - (s[:, 1:] ** 2).sum(1)

    for i in range(nfin):
        for j in range(i + 1, nfin):
            assert ks_2samp(m2(p[i], p[j]), m2(ref[i], ref[j])).pvalue > 1e-4, (i, j)
    # reproducible, and the same events when generated in two pieces
    _, p4b = EventSet.generate(amp, 1000, seed=1234, first=500, return_p4=True)
    np.testing.assert_array_equal(p4b, p4[:, 500:1500])
    # the device sample's angles equal those of an event set uploaded from its four-momenta
    up = EventSet(amp, {f.name: p4[4 * i:4 * i + 4].T[:2000] for i, f in enumerate(amp.finals)})
    np.testing.assert_array_equal(up.angles(), ev.select(np.arange(2000)).angles())
    idx = np.random.default_rng(0).choice(n, 777, replace=False)
    sub = ev.select(idx, weight=np.full(777, 2.0))
    np.testing.assert_array_equal(sub.angles(), ev.angles()[:, idx])
    assert sub.n == 777 and np.all(sub.weight == 2.0)


def test_concat_event_sets(case2, gpu_available):
    """tpa_events_concat: the angles and weights of the sources side by side (bit-identical),
    and one pass over the merged set equals the sum of the passes over its parts."""
    from tf_pwa_amd import EventSet, _lib

    conf = case2["conf"]
    amp = conf.get_amplitude()
    mc = conf.data_from_p4(case2["mc"])
    rng = np.random.default_rng(5)
    parts = [mc.select(np.arange(0, 1000), rng.uniform(0.5, 1.5, 1000)),
             mc.select(np.arange(3000, 3001)),
             mc.select(np.arange(1000, 2777), rng.uniform(-1.0, 1.0, 1777))]
    merged = EventSet.concat(parts)
    assert merged.n == 2778
    np.testing.assert_array_equal(merged.angles(), np.concatenate([p.angles() for p in parts], 1))
    np.testing.assert_array_equal(merged._dev_weight, np.concatenate([p._dev_weight for p in parts]))
    c = amp.coefficients()
    for mode in (_lib.MODE_DATA, _lib.MODE_MC):
        v, G, _ = merged.eval(c, mode, grad=True)
        ref = [p.eval(c, mode, grad=True) for p in parts]
        vr = sum(r[0] for r in ref)
        Gr = sum(r[1] for r in ref)
        assert abs(v - vr) <= 1e-12 * sum(abs(r[0]) for r in ref)
        assert np.abs(G - Gr).max() <= 1e-12 * sum(np.abs(r[1]).max() for r in ref)
    with pytest.raises(Exception):
        EventSet.concat([])


@pytest.mark.parametrize("mode_name", ["MODE_DATA", "MODE_MC"])
def test_cpu_port_matches_engine_pass(case2, mode_name, gpu_available):
    """The oracle's C restatement (the CPU baseline) and the engine's pass agree on the same
    partial waves: the baseline measures the same computation."""
    import ctypes

    from oracle import build_cpu
    from tf_pwa_amd import _lib

    mode = getattr(_lib, mode_name)
    L = build_cpu.load()
    conf = case2["conf"]
    mc = conf.data_from_p4(case2["mc"])
    w = np.random.default_rng(9).uniform(-0.5, 1.5, mc.n)
    mc.set_weight(w)
    c = conf.get_amplitude().coefficients()
    v, G, _ = mc.eval(c, mode, grad=True)
    h = np.ascontiguousarray(mc.partial_waves())
    P, N, LG = h.shape
    out = np.zeros(2 + 2 * P)
    dp = ctypes.POINTER(ctypes.c_double)
    cc = np.ascontiguousarray(c).view(np.float64)
    assert L.cpu_eval(h.view(np.float64).ctypes.data_as(dp), w.ctypes.data_as(dp), N, P, LG,
                      cc.ctypes.data_as(dp), 0 if mode == _lib.MODE_DATA else 1,
                      out.ctypes.data_as(dp)) == 0
    assert abs(out[0] - v) <= 1e-12 * np.abs(w).sum() * max(1.0, abs(v) / np.abs(w).sum())
    np.testing.assert_allclose(out[2:].view(np.complex128), G, rtol=0,
                               atol=1e-11 * np.abs(G).max())


@pytest.mark.parametrize("n", [1, 3, 17, 255, 257, 4097])
def test_eval_pass_ragged_event_counts(case2, n, gpu_available):
    """The pipelined pass at event counts that end inside a wave's group, inside a block's step
    and past one grid stride (clamped loads, the early exit of the second stage), in both modes,
    against the oracle's C restatement on the same partial waves."""
    import ctypes

    from oracle import build_cpu
    from tf_pwa_amd import _lib

    L = build_cpu.load()
    conf = case2["conf"]
    full = conf.data_from_p4(case2["mc"])
    ev = full.select(np.arange(n) % full.n)
    w = np.random.default_rng(n).uniform(-0.5, 1.5, n)
    ev.set_weight(w)
    c = conf.get_amplitude().coefficients()
    h = np.ascontiguousarray(ev.partial_waves())
    P, N, LG = h.shape
    dp = ctypes.POINTER(ctypes.c_double)
    cc = np.ascontiguousarray(c).view(np.float64)
    for mode in (_lib.MODE_DATA, _lib.MODE_MC):
        v, G, f = ev.eval(c, mode, grad=True, fout=True)
        out = np.zeros(2 + 2 * P)
        assert L.cpu_eval(h.view(np.float64).ctypes.data_as(dp), w.ctypes.data_as(dp), N, P, LG,
                          cc.ctypes.data_as(dp), 0 if mode == _lib.MODE_DATA else 1,
                          out.ctypes.data_as(dp)) == 0
        scale = max(1.0, np.abs(w).sum())
        assert abs(out[0] - v) <= 1e-12 * scale * max(1.0, abs(v) / scale)
        np.testing.assert_allclose(out[2:].view(np.complex128), G, rtol=0,
                                   atol=1e-11 * max(1e-300, np.abs(G).max()))
        A = np.einsum("k,kel->el", c, h)
        np.testing.assert_allclose(f, (np.abs(A) ** 2).sum(1), rtol=1e-12)


@pytest.mark.parametrize("cfg", ["lc_pkpi_2res.yml", "lc_pkpi_6res.yml"])
def test_native_shape_session_matches_python_assembly(cfg, gpu_available):
    """Floating masses and widths: the native session (barrier constants, lineshape parameters
    and the shape gradient in C++) against the Python host assembly at several points,
    including a mass step large enough to move q0 and every barrier constant."""
    import yaml

    from conftest import config_path
    from tf_pwa_amd import FCN, ConfigLoader
    from tf_pwa_amd.phasespace import generate_phsp

    c = yaml.safe_load(open(config_path(cfg)))
    for name, p in c["particle"].items():
        if not name.startswith("$") and isinstance(p, dict) and "width" in p:
            p["float"] = "mg"
    conf = ConfigLoader(c, seed=4)
    order = conf.dat_order
    masses = [conf.particles[n].mass for n in order]
    top = conf.particles[conf.top_name].mass
    rng = np.random.default_rng(4)
    data = conf.data_from_p4(dict(zip(order, generate_phsp(top, masses, 3000, rng))),
                             weight=rng.uniform(0.5, 1.5, 3000))
    mc = conf.data_from_p4(dict(zip(order, generate_phsp(top, masses, 20000, rng))))
    fn = FCN(conf.get_model(), data, mc)
    fp = FCN(conf.get_model(), data, mc)
    fp.native = False
    x0 = np.array(conf.vm.get_all_val())
    names = conf.vm.trainable_vars
    for step in range(4):
        x = x0.copy()
        for i, n in enumerate(names):
            if n.endswith("_mass"):
                x[i] += 0.004 * (step + 1) * (1 if i % 2 else -1)
            elif n.endswith("_width"):
                x[i] *= 1.0 + 0.05 * step
        n1, g1 = fn.nll_grad(x)
        assert fn._native is not None  # the native path ran
        n2, g2 = fp.nll_grad(x)
        assert abs(n1 - n2) <= 1e-11 * max(1.0, abs(n2)), (n1, n2)
        np.testing.assert_allclose(g1, g2, rtol=0, atol=1e-9 * np.abs(g2).max())


def test_native_session_rebuilds_on_clamp_flip(gpu_available):
    """K*(1430)'s production-vertex barrier flag flips between m = 1.4481 and 1.4482 GeV
    (P_l(q0^2 d^2) changes sign): the native session rebuilds the caches itself when a mass
    crosses it, and the Python host, sharing the same event sets, sees its caches as stale.
    Order P(lo), N(lo), N(hi), P(lo): the last Python pass reads caches the native session
    rebuilt with the other flags unless it rebuilds them."""
    import yaml

    from conftest import config_path
    from tf_pwa_amd import FCN, ConfigLoader
    from tf_pwa_amd.phasespace import generate_phsp

    c = yaml.safe_load(open(config_path("lc_pkpi_6res.yml")))
    for name, p in c["particle"].items():
        if not name.startswith("$") and isinstance(p, dict) and "width" in p:
            p["float"] = "mg"
    conf = ConfigLoader(c, seed=4)
    order = conf.dat_order
    masses = [conf.particles[n].mass for n in order]
    top = conf.particles[conf.top_name].mass
    rng = np.random.default_rng(5)
    data = conf.data_from_p4(dict(zip(order, generate_phsp(top, masses, 3000, rng))))
    mc = conf.data_from_p4(dict(zip(order, generate_phsp(top, masses, 20000, rng))))
    fn = FCN(conf.get_model(), data, mc)
    fp = FCN(conf.get_model(), data, mc)
    fp.native = False
    names = conf.vm.trainable_vars
    i = names.index("K0st_mass")
    lo = np.array(conf.vm.get_all_val())
    lo[i] = 1.440
    hi = lo.copy()
    hi[i] = 1.456
    amp = conf.get_amplitude()
    flags = []
    for x in (lo, hi):
        conf.vm.set_all(x)
        flags.append([tuple(b["act1"]) for b in amp.barrier_constants()])
    assert flags[0] != flags[1]  # the two points sit on either side of the flip

    def close(a, b):
        assert abs(a[0] - b[0]) <= 1e-11 * max(1.0, abs(b[0])), (a[0], b[0])
        np.testing.assert_allclose(a[1], b[1], rtol=0, atol=1e-9 * np.abs(b[1]).max())

    p_lo = fp.nll_grad(lo)
    n_lo = fn.nll_grad(lo)
    close(n_lo, p_lo)
    n_hi = fn.nll_grad(hi)
    assert fn._native._seen_rebuilds == 1
    close(fp.nll_grad(lo), p_lo)
    close(n_hi, fp.nll_grad(hi))
    close(fn.nll_grad(lo), p_lo)
    assert fn._native._seen_rebuilds == 2


def test_torch_autograd_gradient_matches_central_differences(gpu_available):
    """The NLL as a torch operation (tf_pwa_amd.autograd.nll): its backward pass, the engine's
    adjoint for couplings, masses and widths (46 parameters of the six-resonance model), equals
    central differences of the NLL with h = 1e-6 relative, and composes with a torch term."""
    import torch
    import yaml

    from conftest import config_path
    from tf_pwa_amd import FCN, ConfigLoader
    from tf_pwa_amd.autograd import nll
    from tf_pwa_amd.phasespace import generate_phsp

    c = yaml.safe_load(open(config_path("lc_pkpi_6res.yml")))
    for name, p in c["particle"].items():
        if not name.startswith("$") and isinstance(p, dict) and "width" in p:
            p["float"] = "mg"
    conf = ConfigLoader(c, seed=6)
    order = conf.dat_order
    masses = [conf.particles[n].mass for n in order]
    top = conf.particles[conf.top_name].mass
    rng = np.random.default_rng(6)
    data = conf.data_from_p4(dict(zip(order, generate_phsp(top, masses, 3000, rng))),
                             weight=rng.uniform(0.5, 1.5, 3000))
    mc = conf.data_from_p4(dict(zip(order, generate_phsp(top, masses, 20000, rng))))
    fcn = FCN(conf.get_model(), data, mc)
    x0 = np.array(conf.vm.get_all_val())
    assert len(x0) == 46
    x = torch.tensor(x0, requires_grad=True)
    y = nll(fcn, x)
    (y + 0.5 * (x * x).sum()).backward()
    g = x.grad.numpy() - x0
    fd = np.zeros_like(x0)
    for i in range(len(x0)):
        h = 1e-6 * max(abs(x0[i]), 1e-3)
        xp, xm = x0.copy(), x0.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (fcn(xp) - fcn(xm)) / (2 * h)
    conf.vm.set_all(x0)
    err = np.abs(g - fd)
    assert err.max() <= 1e-5 * np.abs(g).max() + 1e-5, (err.max(), np.abs(g).max())


def test_native_fit_matches_scipy_fit(case1, gpu_available):
    """fit_native (the engine's BFGS in tpa_fcn_minimize) and fit_scipy reach the same minimum
    from the same start (couplings of the 2-resonance model, 800 weighted data events)."""
    from tf_pwa_amd import FCN
    from tf_pwa_amd.fit import fit_native, fit_scipy

    conf = case1["conf"]
    data = conf.data_from_p4(case1["data"])
    mc = conf.data_from_p4(case1["mc"])
    fcn = FCN(conf.get_model(), data, mc)
    start = conf.vm.get_all_dic()
    rn = fit_native(fcn)
    conf.set_params(start)
    rs = fit_scipy(fcn)
    assert rn.success
    assert abs(rn.min_nll - rs.min_nll) < 1e-6 * max(1.0, abs(rs.min_nll)), (rn.min_nll, rs.min_nll)
    assert rn.nfev > 0


@pytest.mark.parametrize("n", [1, 257, 70001])
def test_fp32_cache_pass_four_body_matches_oracle_on_rounded_waves(n, gpu_available):
    """The four-body model's pass over the full FP32 copy (two helicity slots, 25 waves: the
    deferred-value pass k_eval_pipe32d) in FP64 from the FP32-rounded waves: the oracle's C
    restatement on the same rounded waves gives the same sums to FP64 rounding (value 1e-12
    relative, G 1e-11 of its largest entry, f 1e-12), both modes, ragged counts."""
    import ctypes

    from conftest import config_path
    from oracle import build_cpu
    from tf_pwa_amd import ConfigLoader, EventSet, _lib

    L = build_cpu.load()
    conf = ConfigLoader(config_path("jpsi_kkpipi_12ch.yml"), seed=9)
    amp = conf.get_amplitude()
    ev = EventSet.generate(amp, n, seed=77)
    w = np.random.default_rng(n + 3).uniform(-0.5, 1.5, n)
    ev.set_weight(w)
    ev.keep_f32(True)
    dp = ctypes.POINTER(ctypes.c_double)
    c = amp.coefficients()
    h = np.ascontiguousarray(ev.partial_waves().astype(np.complex64).astype(np.complex128))
    P, N, LG = h.shape
    assert LG == 2 and 16 < P <= 32
    cc = np.ascontiguousarray(c).view(np.float64)
    for mode in (_lib.MODE_DATA, _lib.MODE_MC):
        v, G, f = ev.eval(c, mode, grad=True, fout=True, f32=True)
        v2, _, _ = ev.eval(c, mode, grad=False, f32=True)
        assert v2 == v
        out = np.zeros(2 + 2 * P)
        assert L.cpu_eval(h.view(np.float64).ctypes.data_as(dp), w.ctypes.data_as(dp), N, P, LG,
                          cc.ctypes.data_as(dp), 0 if mode == _lib.MODE_DATA else 1,
                          out.ctypes.data_as(dp)) == 0
        scale = max(1.0, np.abs(w).sum())
        assert abs(out[0] - v) <= 1e-12 * scale * max(1.0, abs(v) / scale)
        np.testing.assert_allclose(out[2:].view(np.complex128), G, rtol=0,
                                   atol=1e-11 * max(1e-300, np.abs(G).max()))
        A = np.einsum("k,kel->el", c, h)
        np.testing.assert_allclose(f, (np.abs(A) ** 2).sum(1), rtol=1e-12)
