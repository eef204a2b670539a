"""Pins of the oracle's three-body helicity frames, aligned (Wigner-rotation) angles and helicity
assembly against constructions that share none of its code:

* the helicity frame of every final in every topology as a product of 4x4 Lorentz matrices
  (passive rotation R_y(-theta) R_z(-phi) to the daughter's direction, then the boost along z to
  its rest frame), built from the four-momenta alone: the Wigner rotation between two topologies
  W = L_ref L_t^-1 must be a pure rotation equal to R_z(-gamma) R_y(-beta) R_z(-alpha) of the
  oracle's aligned angles (tf_pwa/cal_angle.py:444-474 with r_boost and final_rest);
* the same frames as products of SL(2, C) matrices in tf-pwa's step convention (rotation_z(a) =
  diag(e^{-ia/2}, e^{ia/2}), rotation_y(b), boost_z(w) = diag(e^{-w/2}, e^{w/2}), alpha folded
  per out as cal_angle.py:291-303 does): the oracle's D^{1/2}(gamma, beta, alpha) must equal that
  product, sign included (SU2M.get_euler_angle, tf_pwa/angle.py:374-392);
* single-chain closed forms: Lc -> K*(892) p, K* -> K pi gives f = a cos^2(theta_K) +
  b sin^2(theta_K) and Lc -> L(1520) pi, L(1520) -> p K gives
  f = sum_{l, l'} |H1_l|^2 |H2_l'|^2 d^{3/2}_{l l'}(theta_p)^2, with the helicity couplings built
  here from sympy's Clebsch-Gordan coefficients and textbook Blatt-Weisskopf factors
  (tf_pwa/amp/core.py:851-900, 1025-1097, 1392-1445);
* invariances of |A|^2 that hold only when the aligned D factors are consistent with the
  vertex D functions: a 2 pi shift of one vertex's alpha (the other SU(2) branch of that
  vertex's rotation), and a change of the topology that defines the proton's reference frame
  (the aligned D must be a representation: D(W_32) = D(W_12) D(W_31)).

Planted mistakes (a negated aligned beta, an aligned alpha shifted by 2 pi, a transposed aligned
D factor) are shown to fail these pins."""

import math

import numpy as np
import pytest
import sympy
import torch
from sympy.physics.quantum.cg import CG

from conftest import config_path
from oracle import tfpwa_oracle as O

ETA = np.diag([1.0, -1.0, -1.0, -1.0])


# ----------------------------------------------------------------- 4x4 Lorentz algebra
def _rz(a):
    c, s = np.cos(a), np.sin(a)
    m = np.zeros(a.shape + (4, 4))
    m[..., 0, 0] = m[..., 3, 3] = 1.0
    m[..., 1, 1], m[..., 1, 2], m[..., 2, 1], m[..., 2, 2] = c, -s, s, c
    return m


def _ry(b):
    c, s = np.cos(b), np.sin(b)
    m = np.zeros(b.shape + (4, 4))
    m[..., 0, 0] = m[..., 2, 2] = 1.0
    m[..., 1, 1], m[..., 1, 3], m[..., 3, 1], m[..., 3, 3] = c, s, -s, c
    return m


def _boost_z_to_rest(p):
    """Boost along z into the rest frame of p (p along z in the current coordinates)."""
    bz = p[..., 3] / p[..., 0]
    g = 1.0 / np.sqrt(1.0 - bz * bz)
    m = np.zeros(p.shape[:-1] + (4, 4))
    m[..., 1, 1] = m[..., 2, 2] = 1.0
    m[..., 0, 0] = m[..., 3, 3] = g
    m[..., 0, 3] = m[..., 3, 0] = -g * bz
    return m


def _mv(m, v):
    return np.einsum("...ij,...j->...i", m, v)


def _mm(a, b):
    return np.einsum("...ij,...jk->...ik", a, b)


def _angles(p):
    """(phi, theta) of the three-momentum of p in the current coordinates."""
    return np.arctan2(p[..., 2], p[..., 1]), np.arctan2(np.hypot(p[..., 1], p[..., 2]), p[..., 3])


def _path(rep, final):
    """The (decay, out) steps from the top down to a final in a representative chain."""
    fs = rep.final_sets()
    for d in rep.decays:
        if d.core is rep.top:
            top = d
    steps = []
    node = top
    while True:
        nxt = [o for o in node.outs if final in fs[o]][0]
        steps.append((node, nxt))
        if nxt.name == final:
            return steps
        node = [d for d in rep.decays if d.core is nxt][0]


def _lab(p4, rep, particle):
    fs = rep.final_sets()
    return sum(np.asarray(p4[n]) for n in sorted(fs[particle]))


def frame_4x4(p4, rep, final):
    """Passive Lorentz matrix from the top rest frame to the final's helicity rest frame."""
    L = None
    for d, out in _path(rep, final):
        p = _lab(p4, rep, out)
        if L is not None:
            p = _mv(L, p)
        phi, th = _angles(p)
        R = _mm(_ry(-th), _rz(-phi))
        step = _mm(_boost_z_to_rest(_mv(R, p)), R)
        L = step if L is None else _mm(step, L)
    return L


def _so3_from_euler(al, be, ga):
    """R_z(-gamma) R_y(-beta) R_z(-alpha), the 3x3 block."""
    return _mm(_mm(_rz(-ga), _ry(-be)), _rz(-al))[..., 1:, 1:]


# ------------------------------------------------------------------ SL(2, C) algebra
def _su2_rz(a):
    z = np.zeros(a.shape + (2, 2), dtype=complex)
    z[..., 0, 0], z[..., 1, 1] = np.exp(-0.5j * a), np.exp(0.5j * a)
    return z


def _su2_ry(b):
    z = np.zeros(b.shape + (2, 2), dtype=complex)
    c, s = np.cos(0.5 * b), np.sin(0.5 * b)
    z[..., 0, 0], z[..., 0, 1], z[..., 1, 0], z[..., 1, 1] = c, -s, s, c
    return z


def _su2_boost(w):
    z = np.zeros(w.shape + (2, 2), dtype=complex)
    z[..., 0, 0], z[..., 1, 1] = np.exp(-0.5 * w), np.exp(0.5 * w)
    return z


def _inv2(m):
    out = np.empty_like(m)
    out[..., 0, 0], out[..., 1, 1] = m[..., 1, 1], m[..., 0, 0]
    out[..., 0, 1], out[..., 1, 0] = -m[..., 0, 1], -m[..., 1, 0]
    return out / (m[..., 0, 0] * m[..., 1, 1] - m[..., 0, 1] * m[..., 1, 0])[..., None, None]


def frame_sl2(p4, rep, final):
    """The same frame as an SL(2, C) product in tf-pwa's convention: per step the rotation
    rotation_y(theta) rotation_z(alpha) with alpha folded into [-pi, pi) for outs[0] and
    [-2 pi, 0) for outs[1], then boost_z(omega) of the daughter."""
    L4 = None
    M = None
    for d, out in _path(rep, final):
        p = _lab(p4, rep, out)
        if L4 is not None:
            p = _mv(L4, p)
        phi, th = _angles(p)
        bias = -math.pi * (1 + list(d.outs).index(out))
        al = np.remainder(phi - bias, 2 * math.pi) + bias
        R4 = _mm(_ry(-th), _rz(-phi))
        pr = _mv(R4, p)
        step4 = _mm(_boost_z_to_rest(pr), R4)
        L4 = step4 if L4 is None else _mm(step4, L4)
        w = np.arcsinh(np.abs(pr[..., 3]) / np.sqrt(pr[..., 0] ** 2 - pr[..., 3] ** 2))
        rot = _mm(_su2_ry(th), _su2_rz(al))
        # tf-pwa keeps the boost of each level apart from the rotations and multiplies the
        # previous level's boost in when it descends: r[out] = rot . b[core] . r[core]
        if M is None:
            r, b = rot, _su2_boost(w)
        else:
            r, b = _mm(_mm(rot, M[1]), M[0]), _su2_boost(w)
        M = (r, b)
    return _mm(M[1], M[0])


def _d_half_from_euler(al, be, ga):
    """The SU(2) matrix the oracle's Euler triple stands for, through the oracle's own
    D^{1/2*} (indices m = -1/2, +1/2): conj(D*(gamma, beta, alpha)) with both axes in the
    (+1/2, -1/2) order of the SU(2) matrices."""
    t = lambda x: torch.as_tensor(np.asarray(x), dtype=O.DT)
    dc = O.D_matrix_conj(t(ga), t(be), t(al), 1).numpy()
    return np.conj(dc)[..., ::-1, ::-1]


# ----------------------------------------------------------------------- fixtures
@pytest.fixture(scope="module")
def six():
    m = O.Model(config_path("lc_pkpi_6res.yml"))
    m.set_params(O.random_couplings(m, 17))
    p = O.gen_mc(m, 400, 33)
    ptop = sum(np.asarray(v) for v in p.values())
    assert np.abs(ptop[:, 1:]).max() < 1e-9  # events in the parent's rest frame
    return m, p


def alignment_errors(model, p4):
    """(max deviation of W from a pure rotation, max |W3 - SO3(euler)|, max |SU(2) product -
    D^{1/2}(euler)|, max |det - 1|) over every final, topology and event."""
    ang = O.cal_angle(model, p4)
    reps = {r.topology_key(): r for r in model.topologies()}
    errs = [0.0, 0.0, 0.0, 0.0]
    n_checked = 0
    for tk, al in ang["aligned"].items():
        for fname, e in al.items():
            ref_key = [k for k in reps if fname not in ang["aligned"][k]][0]
            L_t = frame_4x4(p4, reps[tk], fname)
            L_r = frame_4x4(p4, reps[ref_key], fname)
            W = _mm(L_r, ETA @ np.swapaxes(L_t, -1, -2) @ ETA)
            pure = max(np.abs(W[:, 0, 0] - 1).max(), np.abs(W[:, 0, 1:]).max(),
                       np.abs(W[:, 1:, 0]).max())
            a, b, g = (e[k].numpy() for k in ("alpha", "beta", "gamma"))
            so3 = np.abs(W[:, 1:, 1:] - _so3_from_euler(a, b, g)).max()
            M_t = frame_sl2(p4, reps[tk], fname)
            M_r = frame_sl2(p4, reps[ref_key], fname)
            Rs = _mm(M_r, _inv2(M_t))
            det = np.abs(Rs[:, 0, 0] * Rs[:, 1, 1] - Rs[:, 0, 1] * Rs[:, 1, 0] - 1).max()
            su2 = np.abs(Rs - _d_half_from_euler(a, b, g)).max()
            errs = [max(errs[0], pure), max(errs[1], so3), max(errs[2], su2), max(errs[3], det)]
            n_checked += 1
    assert n_checked >= 4
    return errs


def test_wigner_rotation_from_lorentz_matrices(six):
    """W = L_ref L_t^-1 is a pure rotation and equals the oracle's aligned Euler angles
    (three-body, three topologies, every final), and the SL(2, C) frame product equals the
    oracle's D^{1/2}(gamma, beta, alpha) sign included."""
    m, p = six
    pure, so3, su2, det = alignment_errors(m, p)
    assert pure < 1e-10
    assert det < 1e-10
    assert so3 < 1e-9
    assert su2 < 1e-9


def _planted(monkeypatch, change):
    orig = O.SU2.get_euler_angle

    def wrong(self):
        e = orig(self)
        return change(dict(e))

    monkeypatch.setattr(O.SU2, "get_euler_angle", wrong)


def test_planted_negated_beta_fails_the_frame_pins(six, monkeypatch):
    m, p = six
    _planted(monkeypatch, lambda e: {**e, "beta": -e["beta"]})
    _, so3, su2, _ = alignment_errors(m, p)
    assert so3 > 1e-3 and su2 > 1e-3


def test_planted_2pi_alpha_fails_the_su2_pin(six, monkeypatch):
    m, p = six
    _planted(monkeypatch, lambda e: {**e, "alpha": e["alpha"] + 2 * math.pi})
    _, so3, su2, _ = alignment_errors(m, p)
    assert so3 < 1e-9  # invisible in SO(3) ...
    assert su2 > 1.0  # ... but the SU(2) element changes sign


# ------------------------------------------------------------------- invariances
def _f(model, p4, **kw):
    return O.sum_amp(model, O.cal_angle(model, p4, **kw), model.var_tensors()).numpy()


def _vertex_keys(model):
    """(topology key, core key, [out keys]) of every vertex of every topology."""
    out = []
    for r in model.topologies():
        fs = r.final_sets()
        top = frozenset(f.name for f in model.finals)
        for d in r.decays:
            kc = top if d.core is r.top else fs[d.core]
            out.append((r.topology_key(), kc, [fs[o] for o in d.outs]))
    return out


def test_vertex_branch_shift_leaves_amp2_unchanged(six):
    """Rotating one vertex's two-body system by 2 pi more (both outs' alpha + 2 pi) flips the
    sign of its half-integer D function and of every frame below it; the aligned D factors must
    absorb this, so |A|^2 does not change, vertex by vertex."""
    m, p = six
    f0 = _f(m, p)
    for tk, kc, outs in _vertex_keys(m):
        shift = {(tk, kc, ko): 2 * math.pi for ko in outs}
        f1 = _f(m, p, alpha_shift=shift)
        np.testing.assert_allclose(f1, f0, rtol=1e-9, err_msg=str((tk, kc)))
    # a shift of one out alone breaks the back-to-back relation of the two outs' alpha, and
    # then the amplitude does change (the pin above is not vacuous)
    tk, kc, outs = _vertex_keys(m)[0]
    f2 = _f(m, p, alpha_shift={(tk, kc, outs[1]): 2 * math.pi})
    assert np.abs(f2 / f0 - 1).max() > 1e-3


def _proton_ref_keys(model):
    return [r.topology_key() for r in model.topologies()]


def ref_change_error(model, p4):
    f0 = _f(model, p4)
    err = 0.0
    for tk in _proton_ref_keys(model):
        f1 = _f(model, p4, ref_topology={"p": tk})
        err = max(err, np.abs(f1 / f0 - 1).max())
    return err


def test_reference_frame_choice_leaves_amp2_unchanged(six):
    """Any topology may define the proton's reference helicity frame: the choice multiplies the
    summed amplitude by one unitary matrix in the proton index, so sum |A|^2 is unchanged,
    which holds only if W -> aligned D is a representation (composition, sign included)."""
    m, p = six
    assert ref_change_error(m, p) < 1e-9


def test_planted_2pi_alpha_fails_the_reference_change(six, monkeypatch):
    m, p = six
    _planted(monkeypatch, lambda e: {**e, "alpha": e["alpha"] + 2 * math.pi})
    assert ref_change_error(m, p) > 1e-2


def test_planted_transposed_aligned_D_fails_the_reference_change(six, monkeypatch):
    m, p = six
    orig = O.aligned_D
    monkeypatch.setattr(O, "aligned_D", lambda al, f: orig(al, f).transpose(-1, -2))
    assert ref_change_error(m, p) > 1e-2


# ------------------------------------------------------------------- closed forms
def _cg(j1, m1, j2, m2, J, M):
    S = lambda x: sympy.Rational(int(round(2 * x)), 2)
    return float(CG(S(j1), S(m1), S(j2), S(m2), S(J), S(M)).doit())


def _bprime(l, q2, q02, d=3.0):
    """Blatt-Weisskopf ratio sqrt(P_l(z0) / P_l(z)), z = q^2 d^2 (textbook, l <= 2)."""
    P = {0: lambda z: 1.0 + 0 * z, 1: lambda z: 1.0 + z, 2: lambda z: 9.0 + 3.0 * z + z * z}[l]
    return np.sqrt(P(q02 * d * d) / P(q2 * d * d))


def _q2(M, m1, m2):
    return (M * M - (m1 + m2) ** 2) * (M * M - (m1 - m2) ** 2) / (4 * M * M)


def _h_couplings(model, decay, ja, jb, jc, lam_b, lam_c, q2, q02):
    """H[lambda_b, lambda_c] = sum_ls g_ls q^l B'_l sqrt((2l+1)/(2ja+1))
    <jb lb; jc -lc | s d> <l 0; s d | ja d>, d = lb - lc, per event."""
    H = 0.0
    for (l, s), n in zip(decay.ls, decay.gls_names):
        g = model.values[n + "r"] * np.exp(1j * model.values[n + "i"])
        dl = lam_b - lam_c
        if abs(dl) > s or abs(dl) > ja:
            continue
        cg = math.sqrt((2 * l + 1) / (2 * ja + 1)) * _cg(jb, lam_b, jc, -lam_c, s, dl) * \
            _cg(l, 0, s, dl, ja, dl)
        H = H + g * q2 ** (l / 2) * _bprime(l, q2, q02) * cg
    return H


def _boost_rest(p_frame, p):
    """p in the rest frame of p_frame (pure boost)."""
    b = p_frame[:, 1:] / p_frame[:, :1]
    b2 = (b * b).sum(-1)
    g = 1.0 / np.sqrt(1 - b2)
    bp = -(b * p[:, 1:]).sum(-1)
    g2 = (g - 1) / b2
    vec = p[:, 1:] + (g2 * bp)[:, None] * (-b) + (g * p[:, 0])[:, None] * (-b)
    return vec


def _cos_helicity(p4, pair, daughter):
    """cos of the daughter's helicity angle in the pair's rest frame, against the pair's
    direction in the parent's rest frame."""
    pr = sum(np.asarray(p4[n]) for n in pair)
    v = _boost_rest(pr, np.asarray(p4[daughter]))
    u = pr[:, 1:]
    return (v * u).sum(-1) / np.linalg.norm(v, axis=-1) / np.linalg.norm(u, axis=-1)


def _single_chain_model(chain):
    parts = {
        "$top": {"Lc": {"J": "1/2", "P": 1, "mass": 2.286}},
        "$finals": {"p": {"J": "1/2", "P": 1, "mass": 0.938},
                    "K": {"J": 0, "P": -1, "mass": 0.494},
                    "pi": {"J": 0, "P": -1, "mass": 0.140}},
    }
    if chain == "Kst":
        decay = {"Lc": [["Kst", "p", {"p_break": True}]], "Kst": [["K", "pi", {"p_break": True}]]}
        parts["Kst"] = {"J": 1, "P": -1, "mass": 0.892, "width": 0.050}
    else:
        decay = {"Lc": [["L1520", "pi", {"p_break": True}]],
                 "L1520": [["p", "K", {"p_break": True}]]}
        parts["L1520"] = {"J": "3/2", "P": -1, "mass": 1.519, "width": 0.016}
    cfg = {"data": {"dat_order": ["p", "K", "pi"]}, "decay": decay, "particle": parts,
           "constrains": {"decay": {"fix_chain_idx": 0, "fix_chain_val": 1.0}}}
    m = O.Model(cfg)
    m.set_params(O.random_couplings(m, 29))
    return m


def _bwr_abs2(model, p4, r, y, z, m_y, m_z):
    """|BWR|^2 of resonance r at the events' pair mass (the oracle's BWR is pinned separately:
    Gamma(m0) = Gamma0 and the reference's BW vector)."""
    pr = sum(torch.as_tensor(np.asarray(p4[n])) for n in (y, z))
    m = O.lv_M(pr)
    m0 = torch.tensor(model.values[r + "_mass"], dtype=O.DT)
    g0 = torch.tensor(model.values[r + "_width"], dtype=O.DT)
    q = O.get_relative_p(m, torch.tensor(m_y, dtype=O.DT), torch.tensor(m_z, dtype=O.DT))
    q0 = O.get_relative_p(m0, torch.tensor(m_y, dtype=O.DT), torch.tensor(m_z, dtype=O.DT))
    L = min(l for l, s in [d for d in model.decays if d.core.name == r][0].ls)
    bw = O.BWR(m, m0, g0, q, q0, L, 3.0)
    return (bw.real ** 2 + bw.imag ** 2).numpy(), m.numpy()


def test_single_chain_kstar_closed_form():
    """Lc -> K*(892) p, K* -> K pi: summed over the Lc and proton helicities the top D functions
    are unitary, so sum |A|^2 = |BW|^2 |H2|^2 [a cos^2 th_K + b sin^2 th_K] with
    a = sum_lp |H1(0, lp)|^2, b = (|H1(1, 1/2)|^2 + |H1(-1, -1/2)|^2) / 2."""
    m = _single_chain_model("Kst")
    p = O.gen_mc(m, 300, 41)
    f = _f(m, p)
    bw2, mkpi = _bwr_abs2(m, p, "Kst", "K", "pi", 0.494, 0.140)
    top = [d for d in m.decays if d.core.name == "Lc"][0]
    dec = [d for d in m.decays if d.core.name == "Kst"][0]
    q2_1 = _q2(2.286, mkpi, 0.938)
    q02_1 = _q2(2.286, m.values["Kst_mass"], 0.938)
    q2_2 = _q2(mkpi, 0.494, 0.140)
    q02_2 = _q2(m.values["Kst_mass"], 0.494, 0.140)
    H1 = {(lr, lp): _h_couplings(m, top, 0.5, 1, 0.5, lr, lp, q2_1, q02_1)
          for lr in (-1, 0, 1) for lp in (-0.5, 0.5)}
    H2 = _h_couplings(m, dec, 1, 0, 0, 0, 0, q2_2, q02_2)
    a = np.abs(H1[(0, 0.5)]) ** 2 + np.abs(H1[(0, -0.5)]) ** 2
    b = 0.5 * (np.abs(H1[(1, 0.5)]) ** 2 + np.abs(H1[(-1, -0.5)]) ** 2)
    c = _cos_helicity(p, ("K", "pi"), "K")
    tot = m.values[m.chains[0].total_name + "r"] ** 2
    ref = tot * bw2 * np.abs(H2) ** 2 * (a * c * c + b * (1 - c * c))
    np.testing.assert_allclose(f, ref, rtol=1e-9)
    # the distribution really is of the stated form and not flat in cos(theta_K)
    assert np.std(a * c * c + b * (1 - c * c)) > 1e-3 * np.mean(a + b)


def test_single_chain_lambda1520_closed_form():
    """Lc -> L(1520) pi, L(1520) -> p K: with the top D functions summed away,
    sum |A|^2 = |BW|^2 sum_{l, lp} |H1(l)|^2 |H2(lp)|^2 d^{3/2}_{l, lp}(theta_p)^2 over
    l, lp = +-1/2, with the textbook d^{3/2} functions of the proton's helicity angle."""
    m = _single_chain_model("L1520")
    p = O.gen_mc(m, 300, 43)
    f = _f(m, p)
    bw2, mpk = _bwr_abs2(m, p, "L1520", "p", "K", 0.938, 0.494)
    top = [d for d in m.decays if d.core.name == "Lc"][0]
    dec = [d for d in m.decays if d.core.name == "L1520"][0]
    q2_1 = _q2(2.286, mpk, 0.140)
    q02_1 = _q2(2.286, m.values["L1520_mass"], 0.140)
    q2_2 = _q2(mpk, 0.938, 0.494)
    q02_2 = _q2(m.values["L1520_mass"], 0.938, 0.494)
    H1 = {lr: _h_couplings(m, top, 0.5, 1.5, 0, lr, 0, q2_1, q02_1) for lr in (-0.5, 0.5)}
    H2 = {lp: _h_couplings(m, dec, 1.5, 0.5, 0, lp, 0, q2_2, q02_2) for lp in (-0.5, 0.5)}
    c = _cos_helicity(p, ("p", "K"), "p")
    th = np.arccos(np.clip(c, -1, 1))
    ch, sh = np.cos(th / 2), np.sin(th / 2)
    d = {(0.5, 0.5): 0.5 * (3 * c - 1) * ch, (0.5, -0.5): -0.5 * (3 * c + 1) * sh,
         (-0.5, 0.5): 0.5 * (3 * c + 1) * sh, (-0.5, -0.5): 0.5 * (3 * c - 1) * ch}
    s = sum(np.abs(H1[l]) ** 2 * np.abs(H2[lp]) ** 2 * d[(l, lp)] ** 2
            for l in (-0.5, 0.5) for lp in (-0.5, 0.5))
    tot = m.values[m.chains[0].total_name + "r"] ** 2
    np.testing.assert_allclose(f, tot * bw2 * s, rtol=1e-9)


def test_planted_cg_helicity_sign_fails_the_closed_form(monkeypatch):
    """<jb lb; jc +lc | s, lb + lc> in place of -lc in the CG factor of the helicity couplings
    (a plausible index slip) must break the K* closed form."""

    def wrong_cg(self):
        ja, jb, jc = self.core.J, self.outs[0].J, self.outs[1].J
        ret = np.zeros((len(self.ls), len(self.outs[0].spins), len(self.outs[1].spins)))
        for i, (l, s) in enumerate(self.ls):
            for i1, lb in enumerate(self.outs[0].spins):
                for i2, lc in enumerate(self.outs[1].spins):
                    dl = lb + lc
                    if abs(dl) > s or abs(dl) > ja:
                        continue
                    ret[i, i1, i2] = (math.sqrt((2 * l + 1) / (2 * ja + 1))
                                      * O.cg_coef(jb, jc, lb, lc, s, dl)
                                      * O.cg_coef(l, s, 0, dl, ja, dl))
        return ret

    monkeypatch.setattr(O.Decay, "cg_matrix", property(wrong_cg))
    with pytest.raises(AssertionError):
        test_single_chain_kstar_closed_form()
