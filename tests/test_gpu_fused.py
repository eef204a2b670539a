"""The fused pass (tpa_eval_fused: every partial wave recomputed from the angle cache per event,
in one kernel generated for the model and compiled with hipRTC) against the streaming pass over
the partial-wave cache (tpa_eval) on the same event sets and coefficients, and through
FCN(Model(evaluation="fused")) against the oracle.

Bars: fused vs streaming 1e-11 relative (value) and 1e-10 of the largest |G_k| (both FP64, the
same factors in another order); NLL 1e-6 absolute and gradient 1e-7 relative against the oracle
(the FP64 bar).

The 22-wave Lambda_c model is refused by the library (TPA_FUSED_MAX_P = 8): its generated kernel
took a memory-aperture fault on the GPU with no cause found in the generated source, the compiled
code or the launch, so its cases are skipped and the refusal itself is tested on the CPU."""

import copy

import numpy as np
import pytest
import yaml

from conftest import config_path
from helpers import make_case, oracle_nll_grad

pytestmark = pytest.mark.gpu

_FAULTED = pytest.mark.skip(
    reason="22-wave fused kernel took a memory-aperture fault on gfx950, cause not identified; "
    "the library refuses models above TPA_FUSED_MAX_P waves")


def _compare(ev, coef, mode):
    v_s, G_s, _ = ev.eval(coef, mode, grad=True)
    v_f, G_f = ev.eval_fused(coef, mode, grad=True)
    assert abs(v_f - v_s) <= 1e-11 * max(1.0, abs(v_s)), (mode, v_f, v_s)
    assert np.abs(G_f - G_s).max() <= 1e-10 * np.abs(G_s).max(), (mode, np.abs(G_f - G_s).max())
    v0, G0 = ev.eval_fused(coef, mode, grad=False)
    assert v0 == v_f and G0 is None


@pytest.mark.parametrize("cfg,seed", [("lc_pkpi_2res.yml", 121),
                                      pytest.param("lc_pkpi_6res.yml", 122, marks=_FAULTED)])
def test_fused_pass_equals_streaming_pass(cfg, seed, gpu_available):
    from tf_pwa_amd import _lib

    case = make_case(cfg, n_mc=20000, n_data=700, seed=seed, weight_seed=seed)
    conf = case["conf"]
    amp = conf.get_amplitude()
    coef = amp.coefficients()
    for p4, w in ((case["mc"], None), (case["data"], case["wd"])):
        ev = conf.data_from_p4(p4, weight=w)
        for mode in (_lib.MODE_DATA, _lib.MODE_MC):
            _compare(ev, coef, mode)
    # one event and a ragged count (a partly filled block)
    one = conf.data_from_p4({k: v[:1] for k, v in case["mc"].items()})
    _compare(one, coef, _lib.MODE_DATA)
    rag = conf.data_from_p4({k: v[:257] for k, v in case["mc"].items()})
    _compare(rag, coef, _lib.MODE_MC)
    secs = np.zeros(1)
    _lib.check(_lib.lib().tpa_fused_compile_seconds(amp.device(), _lib.dptr(secs)))
    assert 0 < secs[0] < 60


@pytest.mark.parametrize("cfg", ["lc_pkpi_2res.yml",
                                 pytest.param("lc_pkpi_6res.yml", marks=_FAULTED)])
def test_fused_fcn_matches_oracle(cfg, gpu_available):
    from tf_pwa_amd import FCN

    case = make_case(cfg, n_mc=6000, n_data=800, seed=131, weight_seed=5)
    conf = case["conf"]
    model = conf.get_model()
    model.evaluation = "fused"
    fcn = FCN(model, conf.data_from_p4(case["data"], weight=case["wd"]),
              conf.data_from_p4(case["mc"]))
    nll, g = fcn.nll_grad()
    ref_nll, ref_g = oracle_nll_grad(case)
    assert abs(nll - ref_nll) < 1e-6, (nll, ref_nll)
    assert np.linalg.norm(g - ref_g) <= 1e-7 * np.linalg.norm(ref_g)


def _options_case(cfg_name, clamp_name, width_name, seed):
    from oracle import tfpwa_oracle as O
    from tf_pwa_amd import ConfigLoader

    cfg = yaml.safe_load(open(config_path(cfg_name)))
    P = cfg["particle"]
    if cfg_name == "lc_pkpi_6res.yml":
        P["Kst"]["running_width"] = False
        P["L1520"]["width_norm"] = True
        P["D1232"]["bw_l"] = 3
        P["L1820"]["model"] = "BW"
    else:
        P["Kst"]["model"] = "BW"
        P["L1520"]["width_norm"] = True
        P["L1520"]["bw_l"] = 3
    cfg["constrains"]["decay"]["decay_d"] = 2.5
    om = O.Model(copy.deepcopy(cfg))
    om.set_params(O.random_couplings(om, seed))
    conf = ConfigLoader(copy.deepcopy(cfg), seed=seed)
    conf.set_params(om.values)
    return conf, O.gen_mc(om, 8000, seed + 1), clamp_name, width_name


@pytest.mark.parametrize("cfg_name,clamp_name,width_name", [
    ("lc_pkpi_2res.yml", "Kst_mass", "L1520_width"),
    pytest.param("lc_pkpi_6res.yml", "K0st_mass", "Kst_width", marks=_FAULTED)])
def test_fused_pass_follows_lineshape_options_and_parameter_changes(cfg_name, clamp_name,
                                                                    width_name, gpu_available):
    """A config with the lineshape options and decay_d: the generated kernel bakes the codes,
    the kernel arguments carry m0, g0 and the barrier clamp flags, so a changed mass (taken
    across the production-vertex clamp point m0 + m_p = 2.286 GeV) and a changed width give the
    streaming pass's result after its rebuild."""
    from tf_pwa_amd import _lib

    conf, p4, clamp_name, width_name = _options_case(cfg_name, clamp_name, width_name, 141)
    amp = conf.get_amplitude()
    ev = conf.data_from_p4(p4)
    _compare(ev, amp.coefficients(), _lib.MODE_DATA)
    states = set()
    for name, value in ((clamp_name, 1.30), (clamp_name, 1.60), (width_name, 0.07)):
        conf.vm.set(name, value)
        _compare(ev, amp.coefficients(), _lib.MODE_MC)
        states.add(tuple(tuple(tuple(a) for a in b["act"]) for b in amp.barrier_constants()))
    assert len(states) == 2  # the production-vertex clamp flag differs at 1.30 and 1.60
