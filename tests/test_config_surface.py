"""The config.yml boundary on the CPU: every key the engine implements changes the model or the
parameters as the reference's ConfigLoader does (tf_pwa/config_loader/config_loader.py:258-537,
tf_pwa/amp/core.py:366-418), and every key it does not implement raises instead of being
dropped. The GPU tests hold the same configs to the oracle's likelihood."""

import copy
import os

import numpy as np
import pytest
import torch
import yaml

from conftest import config_path
from oracle import tfpwa_oracle as O


def _base(name="lc_pkpi_6res.yml"):
    with open(config_path(name)) as f:
        return yaml.safe_load(f)


def _loader(cfg, seed=0):
    from tf_pwa_amd import ConfigLoader

    return ConfigLoader(copy.deepcopy(cfg), seed=seed)


def test_unsupported_lineshape_model_raises():
    cfg = _base()
    cfg["particle"]["K0st"] = dict(cfg["particle"]["K0st"], model="GS")
    c = _loader(cfg)
    with pytest.raises(NotImplementedError, match="GS"):
        c.get_amplitude()
    om = O.Model(copy.deepcopy(cfg))
    p = O.gen_mc(om, 20, 1)
    with pytest.raises(NotImplementedError):
        O.sum_amp(om, O.cal_angle(om, p), om.var_tensors())


@pytest.mark.parametrize("where,key,value", [
    ("particle", "coef_head", "Kst"),
    ("particle", "polarization", "vector"),
    ("decay", "has_barrier_factor", False),
    ("decay", "barrier_factor_norm", True),
    ("decay", "model", "helicity_full"),
    ("constrains", "pre_trans", {}),
    ("data", "center_mass", True),
    ("data", "identical_particles", [["K", "pi"]]),
])
def test_unimplemented_options_raise(where, key, value):
    cfg = _base()
    if where == "particle":
        cfg["particle"]["Kst"][key] = value
    elif where == "decay":
        cfg["decay"]["Kst"] = ["K", "pi", {"p_break": True, key: value}]
    elif where == "constrains":
        cfg["constrains"][key] = value
    else:
        cfg["data"][key] = value
    with pytest.raises(NotImplementedError, match=key):
        c = _loader(cfg)
        c.get_model()


def test_display_keys_are_accepted():
    cfg = _base()
    cfg["particle"]["Kst"]["display"] = "$K^*(892)$"
    cfg["decay"]["Kst"] = ["K", "pi", {"p_break": True, "curve_style": "g-"}]
    ref = _loader(_base()).vm.trainable_vars
    assert _loader(cfg).vm.trainable_vars == ref


def test_lineshape_options_lower_to_codes():
    """running_width: False -> BW; width_norm -> times the width (bit 2); bw_l -> the running
    width's L; model BW ignores both options (ParticleBW.get_amp)."""
    from tf_pwa_amd.amp import SHAPE_BW, SHAPE_BWR, SHAPE_WIDTH_NORM

    cfg = _base()
    P = cfg["particle"]
    P["Kst"]["running_width"] = False
    P["L1520"]["width_norm"] = True
    P["D1232"]["bw_l"] = 3
    P["L1820"] = dict(P["L1820"], model="BW", width_norm=True, running_width=True)
    amp = _loader(cfg).get_amplitude()
    M = amp.describe()
    got = {ch["R"].name: (M.chain[i].shape, M.chain[i].bw_l) for i, ch in enumerate(amp.chains)}
    assert got["Kst"][0] == SHAPE_BW
    assert got["L1520"][0] == SHAPE_BWR | SHAPE_WIDTH_NORM
    assert got["D1232"] == (SHAPE_BWR, 3)
    assert got["L1820"][0] == SHAPE_BW
    assert got["K0st"] == (SHAPE_BWR, 0)


def test_oracle_lineshape_options_follow_the_reference_switches():
    """particle_amp restates Particle.get_amp's switches: BW without the running width, the
    width factor with width_norm, and bw_l as the running width's L (the BW and BWR formulas
    themselves are pinned by the reference's BW vector and Gamma(m0) = Gamma0)."""
    m = torch.linspace(0.7, 1.2, 7, dtype=O.DT)
    m0, g0 = torch.tensor(0.892, dtype=O.DT), torch.tensor(0.05, dtype=O.DT)
    q = O.get_relative_p(m, torch.tensor(0.494, dtype=O.DT), torch.tensor(0.14, dtype=O.DT))
    q0 = O.get_relative_p(m0, torch.tensor(0.494, dtype=O.DT), torch.tensor(0.14, dtype=O.DT))
    r = O.Particle("R", J=1, mass=0.892, width=0.05)
    r.decays = [O.Decay(r, [O.Particle("a"), O.Particle("b")], p_break=True)]
    bwr = O.particle_amp(r, m, m0, g0, q, q0)
    assert torch.allclose(bwr, O.BWR(m, m0, g0, q, q0, 1, 3.0))
    r.running_width = False
    assert torch.allclose(O.particle_amp(r, m, m0, g0, q, q0), O.BW(m, m0, g0))
    r.width_norm = True
    assert torch.allclose(O.particle_amp(r, m, m0, g0, q, q0), 0.05 * O.BW(m, m0, g0))
    r.running_width, r.width_norm, r.bw_l = True, False, 2
    b2 = O.particle_amp(r, m, m0, g0, q, q0)
    assert torch.allclose(b2, O.BWR(m, m0, g0, q, q0, 2, 3.0))
    assert not torch.allclose(b2, bwr)
    r.model = "BW"
    r.width_norm, r.running_width = True, True
    assert torch.allclose(O.particle_amp(r, m, m0, g0, q, q0), O.BW(m, m0, g0))


def test_fix_free_and_range_constraints():
    cfg = _base()
    cfg["constrains"].update({
        "fix_var": {"Lc->Kst.p_g_ls_1r": 0.7},
        "free_var": ["Kst_mass"],
        "var_range": {"Kst_mass": [0.85, 0.95], "Lc->L1520.piL1520->p.K_total_0r": [0.0, None]},
    })
    c = _loader(cfg)
    vm = c.vm
    assert "Lc->Kst.p_g_ls_1r" not in vm.trainable_vars
    assert vm.get("Lc->Kst.p_g_ls_1r") == 0.7
    assert vm.trainable_vars[-1] == "Kst_mass"
    assert c.get_amplitude().shape_mode
    assert (vm.bnd_dic["Kst_mass"].lower, vm.bnd_dic["Kst_mass"].upper) == (0.85, 0.95)
    assert vm.bnd_dic["Lc->L1520.piL1520->p.K_total_0r"].upper is None
    cfg["constrains"]["fix_var"] = {"no_such_variable": 1.0}
    with pytest.raises(KeyError):
        _loader(cfg)


def test_var_equal_shares_one_trainable_variable():
    """var_equal (VarsManager.set_same): the later names leave trainable_vars and follow the
    first one's value; d c / d(shared) is the sum over every wave that uses any of the names
    (checked against finite differences of the coefficients)."""
    a, b = "Lc->L1520.piL1520->p.K_total_0r", "Lc->L1405.piL1405->p.K_total_0r"
    cfg = _base()
    cfg["constrains"]["var_equal"] = [[a, b]]
    c = _loader(cfg, seed=3)
    vm = c.vm
    assert a in vm.trainable_vars and b not in vm.trainable_vars
    assert vm.get(a) == vm.get(b)
    amp = c.get_amplitude()
    x0 = np.array(vm.get_all_val())
    i = vm.trainable_vars.index(a)
    vm.set_all(x0 + np.eye(len(x0))[i] * 0.3)
    assert vm.get(b) == vm.get(a) == x0[i] + 0.3
    vm.set_all(x0)
    J = amp.jacobian()
    h = 1e-6
    cp = (vm.set_all(x0 + np.eye(len(x0))[i] * h), amp.coefficients())[1]
    cm = (vm.set_all(x0 - np.eye(len(x0))[i] * h), amp.coefficients())[1]
    vm.set_all(x0)
    np.testing.assert_allclose(J[i], (cp - cm) / (2 * h), rtol=1e-7, atol=1e-9)
    # the waves of both chains move with it
    waves = {amp.chains[w.chain]["R"].name for w, jv in zip(amp.waves, J[i]) if abs(jv) > 0}
    assert waves == {"L1520", "L1405"}


def test_particle_equal_and_float_share_a_mass():
    cfg = _base()
    for r in ("L1520", "L1405"):
        cfg["particle"][r]["float"] = "m"
    cfg["constrains"]["particle"] = {"equal": {"mass": [["L1520", "L1405"]]}}
    c = _loader(cfg)
    tv = c.vm.trainable_vars
    assert "L1520_mass" in tv and "L1405_mass" not in tv
    assert c.vm.get("L1405_mass") == c.vm.get("L1520_mass")
    assert c.vm.index_map()["L1405_mass"] == tv.index("L1520_mass")
    from tf_pwa_amd.serving import shape_map

    sm = shape_map(c.get_amplitude())
    names = [ch["R"].name for ch in c.get_amplitude().chains]
    assert sm.mass_var[names.index("L1405")] == sm.mass_var[names.index("L1520")] >= 0


def test_gaussian_constraints_from_every_source():
    cfg = _base()
    cfg["particle"]["Kst"].update({"float": "mg", "gauss_constr": {"m": 0.002}})
    cfg["particle"]["L1520"].update({"float": "m", "mass_constr": True, "mass_sigma": 0.003})
    cfg["constrains"]["gauss_constr"] = {"Lc->L1520.piL1520->p.K_total_0r": [1.0, 0.5]}
    c = _loader(cfg)
    g = c.gauss_constr
    assert g["Kst_mass"] == (0.892, 0.002)
    assert g["L1520_mass"] == (1.519, 0.003)
    assert g["Lc->L1520.piL1520->p.K_total_0r"] == (1.0, 0.5)
    cfg["particle"]["K0st"]["gauss_constr"] = {"g": None}
    with pytest.raises(ValueError):
        _loader(cfg)


def test_decay_d_sets_every_barrier_radius():
    cfg = _base()
    cfg["constrains"]["decay"]["decay_d"] = 2.5
    amp = _loader(cfg).get_amplitude()
    M = amp.describe()
    assert all(M.chain[i].d == 2.5 for i in range(len(amp.chains)))
    om = O.Model(copy.deepcopy(cfg))
    assert all(d.d == 2.5 for d in om.decays)
    cfg["constrains"]["decay"]["decay_d"] = [3.0, 1.5]
    with pytest.raises(NotImplementedError, match="decay_d"):
        _loader(cfg)


class _Set:
    def __init__(self, n, weight=None):
        self.n = n
        self.has_weight = weight is not None
        self.weight = np.ones(n) if weight is None else np.asarray(weight, dtype=np.float64)


def test_background_weights_follow_the_loader_sign(tmp_path):
    """Background samples loaded from the config carry -bg_weight (a number) or minus their
    weight file; get_weight_data keeps them, and gives -w_bkg only to a background set that
    carries no weights (model.py:603-641, config_loader/data.py:149-251)."""
    from tf_pwa_amd.model import Model

    wfile = os.path.join(tmp_path, "bg_w.dat")
    bw = np.linspace(0.2, 0.9, 7)
    np.savetxt(wfile, bw)
    cfg = _base()
    cfg["data"]["bg_weight"] = 0.731
    c = _loader(cfg)
    np.testing.assert_array_equal(c.sample_weight("bg", 5), O.loaded_weight("bg", 5, 0.731))
    assert c.sample_weight("data", 5) is None
    assert c.get_model().w_bkg == 0.731
    cfg["data"]["bg_weight"] = wfile
    c = _loader(cfg)
    np.testing.assert_array_equal(c.sample_weight("bg", 7), O.loaded_weight("bg", 7, wfile))
    np.testing.assert_array_equal(c.sample_weight("bg", 7), -bw)
    model = Model.__new__(Model)
    model.w_bkg = 0.731
    wd = np.linspace(0.5, 1.5, 11)
    for bg in (_Set(7, -bw), _Set(7, None)):
        _, w = Model.get_weight_data(model, _Set(11, wd), bg=bg)
        ref = O.get_weight_data(11, wd, 7, None if not bg.has_weight else -bw, w_bkg=0.731)
        np.testing.assert_allclose(w, ref.numpy(), rtol=1e-15)
    # the old all-ones heuristic would have treated a background of weight +1 as unweighted
    _, w = Model.get_weight_data(model, _Set(3), bg=_Set(2, np.ones(2)))
    assert np.all(w[3:] > 0)


def test_one_sided_bound_clamps_and_rejects_inverted_ranges():
    from tf_pwa_amd.variable import Bound

    b = Bound(0.5, None)
    assert b.get_y2x(0.2) == 0.0 and np.isfinite(b.get_y2x(0.2))
    assert b.get_x2y(b.get_y2x(1.7)) == pytest.approx(1.7)
    u = Bound(None, 2.0)
    assert u.get_y2x(3.0) == 0.0
    with pytest.raises(ValueError):
        Bound(1.0, 0.5)
