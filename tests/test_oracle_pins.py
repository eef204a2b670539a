"""Pins of the oracle against what fixes it independently of its own formulas: the
reference's golden values and data files, textbook closed forms, exact rational
arithmetic, invariants and finite differences."""

import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import torch

from conftest import config_path
from oracle import tfpwa_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_cg_against_reference_table():
    """tf_pwa/cg_table.json values (reference data file) for j1, j2 <= 2."""
    rows = json.load(open(os.path.join(GOLDEN, "cg_table_subset.json")))["rows"]
    assert len(rows) > 100
    for j1, j2, m1, m2, J, M, v in rows:
        assert abs(O.cg_coef(j1, j2, m1, m2, J, M) - v) < 1e-12


def test_cg_half_integer_closed_forms():
    # <1/2 1/2; 1/2 -1/2 | 1 0> = 1/sqrt(2), <1/2 1/2; 1/2 -1/2 | 0 0> = 1/sqrt(2)
    assert abs(O.cg_coef(0.5, 0.5, 0.5, -0.5, 1, 0) - 2 ** -0.5) < 1e-14
    assert abs(O.cg_coef(0.5, 0.5, 0.5, -0.5, 0, 0) - 2 ** -0.5) < 1e-14
    assert abs(O.cg_coef(0.5, 0.5, -0.5, 0.5, 0, 0) + 2 ** -0.5) < 1e-14
    # <1 1; 1/2 -1/2 | 1/2 1/2> = sqrt(2/3)
    assert abs(O.cg_coef(1, 0.5, 1, -0.5, 0.5, 0.5) - math.sqrt(2 / 3)) < 1e-14


def test_ls_count_reference():
    """tf_pwa/tests/test_particle.py:30-36: 1- -> 1- 1- (C parity not imposed, c_break
    defaults to True) has 4 LS couplings; imposing C = +1 keeps (l, s) = (1, 1) only."""
    assert O.ls_list(1, 1, 1, -1, -1, -1, p_break=False) == [(1, 0), (1, 1), (1, 2), (3, 2)]
    assert O.ls_list(1, 1, 1, -1, -1, -1, p_break=False, ca=1) == [(1, 1)]
    # spin-parity rule examples
    assert O.ls_list(0.5, 1, 0.5, 1, -1, 1, p_break=False) == [(1, 0.5), (1, 1.5)]
    assert O.ls_list(0.5, 1, 0.5, p_break=True) == [(0, 0.5), (1, 0.5), (1, 1.5), (2, 1.5)]


@pytest.mark.parametrize("beta", [0.0, 0.3, 1.7, 3.0])
def test_small_d_closed_forms(beta):
    b = torch.tensor([beta], dtype=O.DT)
    c, s = math.cos(beta / 2), math.sin(beta / 2)
    d = O.small_d_matrix(b, 1)[0].numpy()  # j = 1/2, index (m + j)
    np.testing.assert_allclose(d, [[c, s], [-s, c]], atol=1e-15)
    cb, sb = math.cos(beta), math.sin(beta)
    d = O.small_d_matrix(b, 2)[0].numpy()  # j = 1
    ref = np.array([[(1 + cb) / 2, sb / math.sqrt(2), (1 - cb) / 2],
                    [-sb / math.sqrt(2), cb, sb / math.sqrt(2)],
                    [(1 - cb) / 2, -sb / math.sqrt(2), (1 + cb) / 2]])
    np.testing.assert_allclose(d, ref, atol=1e-14)
    d = O.small_d_matrix(b, 3)[0].numpy()  # j = 3/2: d_{3/2,3/2}, d_{3/2,1/2}, d_{1/2,1/2}
    np.testing.assert_allclose(d[3, 3], c ** 3, atol=1e-14)
    np.testing.assert_allclose(d[3, 2], -math.sqrt(3) * c * c * s, atol=1e-14)
    np.testing.assert_allclose(d[2, 2], c * (3 * c * c - 2), atol=1e-14)
    for j in range(0, 8):
        d = O.small_d_matrix(b, j)[0].numpy()
        np.testing.assert_allclose(d @ d.T, np.eye(j + 1), atol=1e-12)


def test_small_d_sympy():
    """Independent library: sympy's Wigner small-d (Rotation.d) at a generic angle."""
    from sympy import Rational, N
    from sympy.physics.quantum.spin import Rotation

    beta = 0.731
    for j2 in (1, 3, 5):
        d = O.small_d_matrix(torch.tensor([beta], dtype=O.DT), j2)[0].numpy()
        for a in range(j2 + 1):
            for b in range(j2 + 1):
                m, n = Rational(2 * a - j2, 2), Rational(2 * b - j2, 2)
                refc = complex(N(Rotation.d(Rational(j2, 2), m, n, beta).doit()))
                assert abs(refc.imag) < 1e-12
                ref = refc.real
                assert abs(abs(d[a, b]) - abs(ref)) < 1e-12
                assert abs(d[a, b] - ref) < 1e-12 or abs(d[a, b] + ref) < 1e-12


def test_bw_reference_vector():
    """tf_pwa/tests/test_amp.py:302: -BW(m=[1,2], m0=2, G=2) = [-0.12-0.16i, -0.25i]."""
    m = torch.tensor([1.0, 2.0], dtype=O.DT)
    v = -O.BW(m, torch.tensor(2.0, dtype=O.DT), torch.tensor(2.0, dtype=O.DT)).numpy()
    np.testing.assert_allclose(v, [-0.12 - 0.16j, -0.25j], atol=1e-15)


def test_bwr_width_at_pole():
    m0, g0, L, d = 1.519, 0.016, 1, 3.0
    m = torch.tensor([m0], dtype=O.DT)
    q = O.get_relative_p(m, 0.938, 0.494)
    q0 = O.get_relative_p(torch.tensor(m0, dtype=O.DT), 0.938, 0.494)
    g = O.Gamma(m, g0, q, q0, L, m0, d)
    assert abs(float(g) - g0) < 1e-15
    bw = O.BWR(m, torch.tensor(m0, dtype=O.DT), g0, q, q0, L, d)
    assert abs(complex(bw) - 1j / (m0 * g0)) < 1e-9


def _reverse_bessel_sq(l):
    """|theta_l(i w)|^2 coefficients in powers of w^2 from exact rational arithmetic."""
    theta = [Fraction(math.factorial(l + k), math.factorial(l - k) * math.factorial(k) * 2 ** k)
             for k in range(l + 1)]  # coefficient of x^(l-k)
    # theta(i w) = sum_k a_k (i w)^(l-k); |.|^2 = sum_{k,k'} a_k a_k' i^(l-k) (-i)^(l-k') w^..
    res = {}
    for k, a in enumerate(theta):
        for k2, b in enumerate(theta):
            p1, p2 = l - k, l - k2
            ph = (1j) ** p1 * (-1j) ** p2
            res[p1 + p2] = res.get(p1 + p2, 0) + a * b * complex(ph).real
    return [res.get(2 * (l - i), 0) for i in range(l + 1)]


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4, 5])
def test_barrier_polynomial_is_reverse_bessel(l):
    ref = _reverse_bessel_sq(l)
    for got, r in zip(O.BPRIME_COEFF[l], ref):
        assert abs(got - float(r)) < 1e-9 * max(1.0, abs(float(r)))
    # B'_1 = sqrt((z0+1)/(z+1))
    z0, z = 0.3, 0.7
    bp = O.Bprime_q2(1, torch.tensor([z / 9], dtype=O.DT), torch.tensor(z0 / 9, dtype=O.DT), 3.0)
    assert abs(float(bp) - math.sqrt((z0 + 1) / (z + 1))) < 1e-14


def test_clip_log_closed_form():
    x = torch.tensor([2.0, 1e-6, 5e-7, -1e-7], dtype=O.DT)
    y = O.clip_log(x).numpy()
    eps = 1e-6
    assert y[0] == math.log(2.0)
    for xi, yi in zip(x.numpy()[1:], y[1:]):
        t = (xi - eps) / eps
        assert abs(yi - (math.log(eps) + t - t * t / 2)) < 1e-12


def test_phase_space_invariants():
    m = O.Model(config_path("lc_pkpi_2res.yml"))
    p = O.gen_mc(m, 2000, 1)
    tot = sum(p.values())
    np.testing.assert_allclose(tot[:, 0], 2.286, atol=1e-12)
    np.testing.assert_allclose(tot[:, 1:], 0.0, atol=1e-12)
    for k, v in p.items():
        mass = np.sqrt(np.abs(v[:, 0] ** 2 - (v[:, 1:] ** 2).sum(1)))
        np.testing.assert_allclose(mass, m.particles[k].mass, atol=1e-12)


@pytest.fixture(scope="module")
def small_model():
    m = O.Model(config_path("lc_pkpi_6res.yml"))
    m.set_params(O.random_couplings(m, 3))
    p = O.gen_mc(m, 400, 2)
    return m, p


def test_helicity_gamma_zero_and_momentum_reconstruction(small_model):
    """test_cal_angle.py:34 (gamma == 0) and: the top-frame direction of R is
    (sin b cos a, sin b sin a, cos b) in the (x, y, z) base frame at rest."""
    m, p = small_model
    ang = O.cal_angle(m, p)
    top = ang["top_key"]
    for tk, tv in ang["topo"].items():
        for core, outs in tv["ang"].items():
            for k, a in outs.items():
                assert torch.all(a["gamma"] == 0)
        for k, a in tv["ang"][top].items():
            pr = sum(torch.as_tensor(p[n]) for n in k)
            ptop = sum(torch.as_tensor(v) for v in p.values())
            v = O.lv_rest_vector(ptop, pr)[:, 1:]
            u = v / torch.linalg.vector_norm(v, dim=-1, keepdim=True)
            al, be = a["alpha"], a["beta"]
            ref = torch.stack([torch.sin(be) * torch.cos(al), torch.sin(be) * torch.sin(al),
                               torch.cos(be)], -1)
            assert torch.allclose(u, ref, atol=1e-10)


def _rotate(p, R):
    return {k: np.concatenate([v[:, :1], v[:, 1:] @ R.T], axis=1) for k, v in p.items()}


def test_amplitude_rotation_invariance(small_model):
    """sum over helicities of |A|^2 of an unpolarised parent is invariant under a rotation
    of the whole event; this ties the helicity frames of all three topologies together
    through the aligned angles."""
    m, p = small_model
    var = m.var_tensors()
    f0 = O.sum_amp(m, O.cal_angle(m, p), var).numpy()
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    f1 = O.sum_amp(m, O.cal_angle(m, _rotate(p, q)), var).numpy()
    assert np.all(f0 >= 0)
    np.testing.assert_allclose(f1, f0, rtol=1e-9)


def test_oracle_gradient_finite_difference(small_model):
    m, p = small_model
    d = {k: v[:150] for k, v in p.items()}
    ad, am = O.cal_angle(m, d), O.cal_angle(m, p)
    w, wm = O.fcn_weights(150, None, 400)
    f = O.nll_fn(m, ad, am, w, wm)
    x0 = torch.tensor(m.trainable_values(), dtype=O.DT)
    nll, g = O.nll_grad(m, ad, am, w, wm)
    for i in range(0, len(x0), 5):
        h = 1e-6 * max(1.0, abs(float(x0[i])))
        xp, xm = x0.clone(), x0.clone()
        xp[i] += h
        xm[i] -= h
        fd = (float(f(xp)) - float(f(xm))) / (2 * h)
        assert abs(fd - g[i]) < 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


@pytest.mark.parametrize("name", ["lc_pkpi_2res", "lc_pkpi_6res", "jpsi_kkpipi_12ch"])
def test_oracle_reproduces_committed_golden_vectors(name):
    """The committed golden vectors are the oracle's outputs on their inputs: re-running the
    oracle must give the same f, NLL, gradient and Hessian (a regression pin of the oracle)."""
    import json
    import os

    from golden.make_golden import compute

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "golden_{}.json".format(name))) as fh:
        fx = json.load(fh)
    out = compute(name, fx["inputs"])
    ref = fx["outputs"]
    assert out["trainable"] == ref["trainable"]
    np.testing.assert_allclose(out["f_data"], ref["f_data"], rtol=1e-12)
    np.testing.assert_allclose(out["f_mc"], ref["f_mc"], rtol=1e-12)
    assert abs(out["nll"] - ref["nll"]) < 1e-12 * abs(ref["nll"])
    np.testing.assert_allclose(out["grad"], ref["grad"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(out["hessian"], ref["hessian"], rtol=1e-9, atol=1e-10)


def test_oracle_four_body_rotation_invariance():
    """J/psi -> K+ K- pi+ pi- through 12 chains in 6 topologies (quasi-two-body and sequential):
    with every top helicity summed, |A|^2 must not change under a rotation of whole events,
    which ties the frame conventions of all topologies together; with helicities +-1 only it is
    invariant under rotations about the z axis alone."""
    import yaml
    from scipy.spatial.transform import Rotation

    from tf_pwa_amd.phasespace import generate_phsp

    cfg = yaml.safe_load(open(config_path("jpsi_kkpipi_12ch.yml")))
    for full in (True, False):
        c = dict(cfg)
        c["particle"] = {k: (dict(v) if isinstance(v, dict) else v) for k, v in cfg["particle"].items()}
        if full:
            c["particle"]["$top"] = {"Jpsi": {k: v for k, v in cfg["particle"]["$top"]["Jpsi"].items()
                                              if k != "spins"}}
        om = O.Model(c)
        om.set_params(O.random_couplings(om, 5))
        order = om.config["data"]["dat_order"]
        ps = generate_phsp(om.top.mass, [om.particles[n].mass for n in order], 400,
                           np.random.default_rng(3))
        p4 = dict(zip(order, ps))
        f = O.sum_amp(om, O.cal_angle(om, p4), om.var_tensors()).numpy()
        phi = 0.9
        Rz = np.array([[np.cos(phi), -np.sin(phi), 0], [np.sin(phi), np.cos(phi), 0], [0, 0, 1]])
        R = Rotation.random(random_state=2).as_matrix() if full else Rz
        p4r = {k: np.concatenate([v[:, :1], v[:, 1:] @ R.T], 1) for k, v in p4.items()}
        fr = O.sum_amp(om, O.cal_angle(om, p4r), om.var_tensors()).numpy()
        assert np.abs(fr - f).max() < 1e-10 * f.max()
        if not full:
            Rx = Rotation.from_euler("x", 0.7).as_matrix()
            p4x = {k: np.concatenate([v[:, :1], v[:, 1:] @ Rx.T], 1) for k, v in p4.items()}
            fx = O.sum_amp(om, O.cal_angle(om, p4x), om.var_tensors()).numpy()
            assert np.abs(fx - f).max() > 1e-3 * f.max()  # helicity +-1 only: not invariant


def _boost_to_rest(p, frame):
    """Pure boost of four-vectors p (n, 4) into the rest frame of `frame` (n, 4); the axes
    keep their directions (no rotation)."""
    b = frame[:, 1:] / frame[:, :1]
    b2 = (b * b).sum(1)
    g = 1.0 / np.sqrt(1.0 - b2)
    bp = (b * p[:, 1:]).sum(1)
    k = np.where(b2 > 0, (g - 1.0) / np.where(b2 > 0, b2, 1.0), 0.0)
    out = np.empty_like(p)
    out[:, 0] = g * (p[:, 0] - bp)
    out[:, 1:] = p[:, 1:] + (k * bp - g * p[:, 0])[:, None] * b
    return out


def _rotate_pair_in_rest_frame(pa, pb, R):
    """Rotate the back-to-back momenta of the pair (a, b) by R in the pair's rest frame (pure
    boost there and back): the pair's four-momentum, every invariant mass built from the pair
    or from it and other particles whole, stays; the pair's decay angles change."""
    P = pa + pb
    a = _boost_to_rest(pa, P)
    b = _boost_to_rest(pb, P)
    a[:, 1:] = a[:, 1:] @ R.T
    b[:, 1:] = b[:, 1:] @ R.T
    back = P.copy()
    back[:, 1:] = -back[:, 1:]
    return _boost_to_rest(a, back), _boost_to_rest(b, back)


@pytest.mark.parametrize("kind", ["quasi-two-body", "sequential"])
def test_oracle_four_body_single_chain_closed_form(kind):
    """One four-body chain whose every vertex has a single LS coupling with a helicity-
    independent helicity amplitude: J/psi(+-1) -> phi f0, phi -> K+ K- (l = 1), f0 -> pi pi;
    and J/psi(+-1) -> K1+ K- (S wave), K1+ -> K*0 pi+ (S wave), K*0 -> K+ pi- (l = 1). The sum
    over the intermediate helicities of the product of the vertices' D functions is then the D
    function of the composed rotation, so sum_{lambda = +-1} |A|^2 = C(masses) sin^2(beta),
    with beta the polar angle (to the J/psi z axis) of the K+ after the chain of pure boosts
    into the last resonance's rest frame. Rotating the last pair inside its own rest frame keeps
    every mass of the chain, so f'/f must equal sin^2(beta') / sin^2(beta)."""
    from scipy.spatial.transform import Rotation

    from tf_pwa_amd.phasespace import generate_phsp

    cfg = O.yaml.safe_load(open(config_path("jpsi_kkpipi_12ch.yml")))
    parts = cfg["particle"]
    if kind == "quasi-two-body":
        decay = {"Jpsi": [["phi", "f0pp", {"l_list": [0]}]], "phi": ["Kp", "Km"],
                 "f0pp": ["pip", "pim"]}
        keep = ["phi", "f0pp"]
    else:
        decay = {"Jpsi": [["K1p", "Km", {"l_list": [0]}]],
                 "K1p": [["Kst0", "pip", {"l_list": [0]}]], "Kst0": ["Kp", "pim"]}
        keep = ["K1p", "Kst0"]
    c = {"data": cfg["data"], "decay": decay,
         "particle": {"$top": parts["$top"], "$finals": parts["$finals"],
                      **{k: parts[k] for k in keep}}}
    om = O.Model(c)
    assert len(om.chains) == 1
    om.set_params(O.random_couplings(om, 11))
    order = om.config["data"]["dat_order"]
    ps = generate_phsp(om.top.mass, [om.particles[n].mass for n in order], 300,
                       np.random.default_rng(7))
    p4 = dict(zip(order, ps))

    def f_of(p):
        return O.sum_amp(om, O.cal_angle(om, p), om.var_tensors()).numpy()

    def sin2beta(p):
        if kind == "quasi-two-body":
            k = _boost_to_rest(p["Kp"], p["Kp"] + p["Km"])
        else:
            k1 = p["Kp"] + p["pim"] + p["pip"]
            kst = _boost_to_rest(p["Kp"] + p["pim"], k1)
            k = _boost_to_rest(_boost_to_rest(p["Kp"], k1), kst)
        n = k[:, 1:] / np.linalg.norm(k[:, 1:], axis=1, keepdims=True)
        return 1.0 - n[:, 2] ** 2

    R = Rotation.random(random_state=4).as_matrix()
    a, b = ("Kp", "Km") if kind == "quasi-two-body" else ("Kp", "pim")
    pr = dict(p4)
    pr[a], pr[b] = _rotate_pair_in_rest_frame(p4[a], p4[b], R)
    f, fr = f_of(p4), f_of(pr)
    s, sr = sin2beta(p4), sin2beta(pr)
    ok = (s > 1e-3) & (sr > 1e-3)
    np.testing.assert_allclose(fr[ok] / f[ok], sr[ok] / s[ok], rtol=1e-9)
    # the closed form is not trivially satisfied: the angles did change
    assert np.abs(sr / s - 1.0)[ok].max() > 0.1
