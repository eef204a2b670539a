"""The helicity-flip relation the halved data storage rests on, pinned on the oracle's waves.

For a single partial wave (one LS coupling per vertex, one chain), flipping every helicity
conjugates the angular part: with A_o = c h_o,

    A_-o = eta s_o (B / conj B) (c / conj c) conj(A_o),

where B is the chain's lineshape product, s_o = (-1)^(sum of the slot's helicities) and
eta = +-1 depends on the wave only. This follows from d^j_{-m,-n} = (-1)^(m-n) d^j_{m,n}, the
CG symmetry and real barrier factors; the test checks it on the oracle's own chain amplitudes
(tf_pwa/amp/core.py:1392-1445 restated), including the aligned rotation of the proton, and
checks that a planted mistake (a sign slip in one d-function element) breaks it.
"""

import itertools

import numpy as np
import pytest
import torch

from conftest import config_path


def _wave_ratios(om, O, ch, ang, lam_top):
    """X[wave][event, o] = A_-o / conj(A_o) / (c / conj c) / (B / conj B) for every single
    wave of chain ch, and the slot helicity sums."""
    base = dict(om.values)
    vlists = [d.gls_names for d in ch.decays]
    var0 = om.var_tensors()
    # the chain's lineshape product B at each event, as chain_amplitude builds it
    fs = ch.final_sets()
    mass = ang["mass"]
    B = None
    for r in ch.inner:
        dec = [d for d in ch.decays if d.core is r][0]
        y, z = dec.outs
        m = mass[fs[r]]
        q = O.get_relative_p(m, mass[fs[y]], mass[fs[z]])
        m0 = var0[r.name + "_mass"]
        q0 = O.get_relative_p(m0, torch.tensor(y.mass, dtype=O.DT),
                              torch.tensor(z.mass, dtype=O.DT))
        b = O.particle_amp(r, m, m0, var0.get(r.name + "_width"), q, q0)
        if not torch.is_complex(b):
            b = torch.complex(b, torch.zeros_like(b))
        B = b if B is None else B * b
    B = B.numpy()
    out = {}
    for combo in itertools.product(*[range(len(v)) for v in vlists]):
        vals = dict(base)
        t = ch.total_name
        c = base[t + "r"] * np.exp(1j * base[t + "i"])
        for v, s in zip(vlists, combo):
            for i, n in enumerate(v):
                if i != s:
                    vals[n + "r"] = 0.0
            c *= vals[v[s] + "r"] * np.exp(1j * vals[v[s] + "i"])
        om.values = vals
        try:
            A = O.chain_amplitude(om, ch, ang, om.var_tensors()).detach().numpy()
        finally:
            om.values = base
        A = A.reshape(A.shape[0], -1)
        # helicity lists are sorted and symmetric: slot o's mirror is the reversed index
        out[combo] = A[:, ::-1] / np.conj(A) / (c / np.conj(c)) / (B / np.conj(B))[:, None]
    return out


def _slot_signs(om, lam_top):
    """(-1)^(sum of helicities) of every outer slot, finals in the oracle's sorted order."""
    lists = [lam_top] + [list(f.spins) for f in om.finals]
    return np.array([(-1.0) ** round(sum(x)) for x in itertools.product(*lists)])


def _check_model(cfg, n, gen):
    from oracle import tfpwa_oracle as O

    om = O.Model(config_path(cfg))
    om.set_params(O.random_couplings(om, 3))
    p4 = gen(O, om, n)
    ang = O.cal_angle(om, p4)
    s = _slot_signs(om, list(om.top.spins))
    for ch in om.chains:
        X = _wave_ratios(om, O, ch, ang, list(om.top.spins))
        for combo, x in X.items():
            # x = eta s_o: the same +-1 for every event, its slot pattern the closed form
            eta = np.round(np.median(x[:, 0].real)) * s[0]
            assert eta in (-1.0, 1.0)
            np.testing.assert_allclose(x, np.broadcast_to(eta * s + 0j, x.shape), atol=1e-11,
                                       rtol=0)
    return om


def _gen3(O, om, n):
    return O.gen_mc(om, n, 11)


def _gen4(O, om, n):
    from tf_pwa_amd.phasespace import generate_phsp

    order = om.config["data"]["dat_order"]
    masses = [om.particles[k].mass for k in order]
    return dict(zip(order, generate_phsp(om.top.mass, masses, n, np.random.default_rng(1))))


def test_helicity_flip_relation_three_body():
    """Six resonances in three topologies, the proton's helicity aligned in two of them."""
    _check_model("lc_pkpi_6res.yml", 200, _gen3)


def test_helicity_flip_relation_four_body():
    """Twelve quasi-two-body and sequential chains of J/psi -> K+ K- pi+ pi-."""
    _check_model("jpsi_kkpipi_12ch.yml", 60, _gen4)


def test_helicity_flip_relation_fails_a_planted_mistake(monkeypatch):
    """A sign slip in one element of the d-functions (d^1_{1,0} -> -d^1_{1,0}) breaks the
    relation: the check is not satisfied by construction."""
    from oracle import tfpwa_oracle as O

    orig = O.small_d_matrix

    def slipped(theta, j):
        d = orig(theta, j)
        if j == 2:  # the oracle indexes spins doubled
            d = d.clone()
            d[..., 2, 1] = -d[..., 2, 1]
        return d

    monkeypatch.setattr(O, "small_d_matrix", slipped)
    with pytest.raises(AssertionError):
        _check_model("lc_pkpi_6res.yml", 60, _gen3)
