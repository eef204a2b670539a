"""CPU float64 restatement of tf-pwa's NLL / NLL-gradient / Hessian hot path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; the product never imports it.

Every function follows one reference function (paths relative to the tf-pwa tree) and uses
eager torch ops on CPU in float64/complex128 where the reference uses eager TensorFlow ops.
Gradients and Hessians come from ``torch.autograd`` (the reference uses ``tf.GradientTape``,
``tf_pwa/model/model.py:57-140``). ``torch.where`` has the same gradient rule as
``tf.where`` (the unselected branch receives an exact zero), so the threshold clamps of
``get_relative_p`` / ``Gamma`` / ``Bprime_q2`` differentiate the same way.

Parity pins (what checks each part independently of this file, see ``tests/``):
  * cg_coef ............ sympy's CG and the reference's integer-j ``cg_table.json`` values.
  * ls_list ............ angular-momentum rules and the reference's LS count test.
  * small_d_matrix ..... textbook d^1/2, d^1, d^3/2 tables, unitarity, sympy Rotation.d.
  * BW / BWR / barrier . the reference's BW golden vector (tests/test_amp.py:294-305),
                         Gamma(m0)=Gamma0, reverse-Bessel |theta_l(iw)|^2 polynomials.
  * clip_log ........... closed form; sum_hessian closed form (tests/test_model_new.py).
  * cal_angle .......... invariants: gamma=0 for helicity angles, reconstruction of the
                         boosted momenta from (m, theta, phi). Aligned angles: the Wigner
                         rotation of 4x4 Lorentz helicity frames built from the momenta is a
                         pure rotation equal to the aligned Euler angles' SO(3) matrix; the
                         SL(2, C) frame product in the reference's step convention equals
                         D^1/2 of the aligned angles sign included; |A|^2 is invariant under a
                         2 pi shift of any vertex's rotation branch and under a change of the
                         proton's reference topology (tests/test_oracle_alignment.py).
  * amplitude / NLL .... |A|^2 >= 0, rotation invariance, single-chain closed forms
                         (Lc -> K*(892) p -> K pi p: a cos^2 th_K + b sin^2 th_K;
                         Lc -> L(1520) pi -> p K pi: d^3/2 squares) with helicity couplings
                         from sympy's CG, four-body single-chain closed forms, finite
                         differences of the gradient.
Parity unpinned (source reading only): the per-vertex alpha folding ranges of
``cal_helicity_angle`` ([-pi, pi) for outs[0], [-2 pi, 0) for outs[1]), a convention restated
from cal_angle.py:291-303 that every pin above takes as given; the pins fix everything that
follows from it (frames, aligned angles with their SU(2) sign, the assembly).
"""

import functools
import math
import re
from fractions import Fraction

import numpy as np
import torch
import yaml

DT = torch.float64
CT = torch.complex128


# --------------------------------------------------------------------------------------
# small helpers (particle.py:43-52)
# --------------------------------------------------------------------------------------
def spin_int(x):
    if isinstance(x, int):
        return x
    return int(x + 0.1)


def spin_range(a, b):
    while a <= b:
        yield a
        a = a + 1


def _num(x):
    if isinstance(x, str):
        return float(Fraction(x))
    return x


# --------------------------------------------------------------------------------------
# LS couplings: particle.py:207-248 (GetA2BC_LS_list)
# --------------------------------------------------------------------------------------
def ls_list(ja, jb, jc, pa=None, pb=None, pc=None, p_break=False, ca=None):
    if pa is None or pb is None or pc is None:
        p_break = True
    if not p_break:
        dl = 0 if pa * pb * pc == 1 else 1
    ret = []
    for s in spin_range(abs(jb - jc), jb + jc):
        for l in spin_range(abs(ja - s), ja + s):
            if 0.3 < l % 1.0 < 0.7:
                break
            l = spin_int(l)
            if ca is not None and ca != (-1) ** (l + s):
                continue
            if not p_break:
                if l % 2 == dl:
                    ret.append((l, s))
            else:
                ret.append((l, s))
    return ret


# --------------------------------------------------------------------------------------
# Clebsch-Gordan: cg.py:18-26 (sympy path of the reference)
# --------------------------------------------------------------------------------------
@functools.lru_cache(maxsize=None)
def cg_coef(jb, jc, mb, mc, ja, ma):
    from sympy import Rational
    from sympy.physics.quantum.cg import CG

    R = lambda x: Rational(int(round(2 * x)), 2)
    return float(CG(R(jb), R(mb), R(jc), R(mc), R(ja), R(ma)).doit().evalf())


# --------------------------------------------------------------------------------------
# Wigner d / D: dfun.py:118-266
# --------------------------------------------------------------------------------------
@functools.lru_cache(maxsize=None)
def small_d_weight(j):
    """dfun.py:118-143; ``j`` is 2J, result indexed [l][(m+j)/2][(n+j)/2]."""
    ret = np.zeros((j + 1, j + 1, j + 1))

    def f(x):
        return math.factorial(x >> 1)

    for m in range(-j, j + 1, 2):
        for n in range(-j, j + 1, 2):
            for k in range(max(0, n - m), min(j - m, j + n) + 1, 2):
                l = (2 * k + (m - n)) // 2
                sign = (-1) ** ((k + m - n) // 2)
                tmp = sign * math.sqrt(1.0 * f(j + m) * f(j - m) * f(j + n) * f(j - n))
                tmp /= f(j - m - k) * f(j + n - k) * f(k + m - n) * f(k)
                ret[l][(m + j) // 2][(n + j) // 2] = tmp
    return ret


def small_d_matrix(theta, j):
    """dfun.py:146-173: d^{j/2}(theta) as (N, j+1, j+1) via the sin/cos power matmul."""
    a = torch.arange(0, j + 1, dtype=DT).reshape(1, -1)
    half = 0.5 * theta
    s = torch.sin(half).reshape(-1, 1)
    c = torch.cos(half).reshape(-1, 1)
    sc = torch.pow(s, a) * torch.pow(c, j - a)
    w = torch.as_tensor(small_d_weight(j), dtype=DT).reshape(j + 1, (j + 1) * (j + 1))
    return torch.matmul(sc, w).reshape(-1, j + 1, j + 1)


def exp_i(theta, mi):
    """dfun.py:176-193"""
    m_theta = torch.as_tensor(mi, dtype=DT) * theta.reshape(-1, 1)
    return torch.exp(torch.complex(torch.zeros_like(m_theta), m_theta))


def D_matrix_conj(alpha, beta, gamma, j):
    """dfun.py:196-218: D^{j/2*}_{m1 m2} = e^{i m1 alpha} d_{m1 m2}(beta) e^{i m2 gamma}."""
    m = np.arange(-j / 2, j / 2 + 1, 1).reshape(1, -1)
    d = small_d_matrix(beta, j)
    ea = exp_i(alpha, m).reshape(-1, j + 1, 1)
    eg = exp_i(gamma, m).reshape(-1, 1, j + 1)
    dc = torch.complex(d, torch.zeros_like(d))
    return (ea * eg) * dc


def get_D_matrix_lambda(angle, ja, la, lb, lc=None):
    """dfun.py:241-266 with Dfun_delta_v2 (98-115): gather D[la, lb - lc]."""
    j2 = spin_int(2 * ja)
    d = D_matrix_conj(angle["alpha"], angle["beta"], angle["gamma"], j2)
    ln = j2 + 1
    lc_t = (0,) if lc is None else tuple(lc)
    idx = []
    for la_i in la:
        for lb_i in lb:
            for lc_i in lc_t:
                delta = lb_i - lc_i
                if abs(delta) <= ja:
                    idx.append(spin_int((la_i + ja) * ln + delta + ja))
                else:
                    idx.append(ln * ln)
    d = d.reshape(-1, ln * ln)
    over_d = torch.cat([d, torch.zeros_like(d[:, :1])], dim=-1)
    ret = over_d[:, idx]
    if lc is None:
        return ret.reshape(-1, len(la), len(lb))
    return ret.reshape(-1, len(la), len(lb), len(lc_t))


# --------------------------------------------------------------------------------------
# Lineshapes and barrier factors: breit_wigner.py:56-91,231-277,290-355; amp/core.py:311-334
# --------------------------------------------------------------------------------------
BPRIME_COEFF = {
    0: [1.0],
    1: [1.0, 1.0],
    2: [1.0, 3.0, 9.0],
    3: [1.0, 6.0, 45.0, 225.0],
    4: [1.0, 10.0, 135.0, 1575.0, 11025.0],
    5: [1.0, 15.0, 315.0, 6300.0, 99225.0, 893025.0],
}


def Bprime_polynomial(l, z):
    """breit_wigner.py:332-355 (tf.math.polyval, highest power first)."""
    cof = BPRIME_COEFF[int(l + 0.01)]
    ret = torch.zeros_like(z) + cof[0]
    for c in cof[1:]:
        ret = ret * z + c
    return ret


def BW(m, m0, g0):
    """breit_wigner.py:56-71"""
    x = m0 * m0 - m * m
    y = m0 * g0
    s = x * x + y * y
    return torch.complex(x / s, y / s)


def Bprime(L, q, q0, d):
    """breit_wigner.py:290-306"""
    num = torch.sqrt(Bprime_polynomial(L, (q0 * d) ** 2))
    den = torch.sqrt(Bprime_polynomial(L, (q * d) ** 2))
    return num / den


def Gamma(m, gamma0, q, q0, L, m0, d):
    """breit_wigner.py:231-245"""
    qq0 = torch.where(q0 > 1e-15, (q / q0) ** (2 * L + 1), torch.ones_like(q / q0))
    mm0 = m0 / m
    bp = Bprime(L, q, q0, d) ** 2
    return gamma0 * qq0 * mm0 * bp


def BWR(m, m0, g0, q, q0, L, d):
    """breit_wigner.py:74-91"""
    gamma = Gamma(m, g0, q, q0, L, m0, d)
    x = m0 * m0 - m * m
    y = m0 * gamma
    s = x * x + y * y
    return torch.complex(x / s, y / s)


def Bprime_q2(L, q2, q02, d):
    """breit_wigner.py:268-277"""
    z0 = q02 * d ** 2
    z = q2 * d ** 2
    bp = Bprime_polynomial(L, z0) / Bprime_polynomial(L, z)
    return torch.sqrt(torch.where(bp > 0, bp, torch.ones_like(bp)))


def get_relative_p(m_0, m_1, m_2):
    """amp/core.py:311-321 (with the m_eff = max(m0, m1+m2) clamp)."""
    m12s = m_1 + m_2
    m12d = m_1 - m_2
    m_0 = torch.as_tensor(m_0, dtype=DT)
    m_eff = torch.where(m_0 > m12s, m_0, torch.zeros_like(m_0) + m12s)
    p = (m_eff - m12s) * (m_eff + m12s) * (m_eff - m12d) * (m_eff + m12d)
    return torch.sqrt(p) / (2 * m_eff)


def get_relative_p2(m_0, m_1, m_2):
    """amp/core.py:324-334 (no clamp), same as cal_angle.py:515-528 Getp2."""
    m12s = m_1 + m_2
    m12d = m_1 - m_2
    m_0 = torch.as_tensor(m_0, dtype=DT)
    p = (m_0 - m12s) * (m_0 + m12s) * (m_0 - m12d) * (m_0 + m12d)
    return p / (2 * m_0) ** 2


def clip_log(x, eps=1e-6):
    """model/model.py:283-292"""
    x_cut = torch.where(x > eps, x, torch.ones_like(x) * eps)
    b_t = torch.log(x_cut)
    delta_x = x - eps
    b_f = np.log(eps) + delta_x / eps - (delta_x / eps) ** 2 / 2.0
    return torch.where(x > eps, b_t, b_f)


# --------------------------------------------------------------------------------------
# Lorentz / SU(2) algebra: angle.py:13-395
# --------------------------------------------------------------------------------------
_EPS = 1.0e-14


def v_dot(a, b):
    return torch.sum(a * b, dim=-1)


def v_norm(a):
    return torch.linalg.vector_norm(a, dim=-1)


def v_unit(a):
    return a / v_norm(a)[..., None]


def v_cross(a, b):
    a0, a1, a2 = a[..., 0], a[..., 1], a[..., 2]
    b0, b1, b2 = b[..., 0], b[..., 1], b[..., 2]
    return torch.stack([a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0], dim=-1)


def v_cross_unit(a, b):
    """angle.py:57-67"""
    cro = v_cross(a, b)
    n = v_norm(cro)[..., None]
    mask = n < _EPS
    bias_other = torch.ones_like(n) + b
    cro = torch.where(mask, v_cross(a, bias_other), cro)
    return v_unit(cro)


def v_angle_from(v, x, y):
    return torch.atan2(v_dot(v, y), v_dot(v, x))


def lv_vect(p):
    return p[..., 1:4]


def lv_boost(p, pb):
    """angle.py:148-163"""
    beta2 = torch.sum(pb * pb, dim=-1)
    gamma = 1.0 / torch.sqrt(1 - beta2)
    bp = v_dot(pb, lv_vect(p))
    gamma2 = torch.where(beta2 > _EPS, (gamma - 1.0) / beta2, torch.zeros_like(beta2))
    p_r = lv_vect(p)
    p_r = p_r + (gamma2 * bp)[..., None] * pb
    p_r = p_r + (gamma * p[..., 0])[..., None] * pb
    t_r = (gamma * (p[..., 0] + bp))[..., None]
    return torch.cat([t_r, p_r], dim=-1)


def lv_rest_vector(p_frame, other):
    """angle.py:140-146"""
    return lv_boost(other, -(p_frame[..., 1:4] / p_frame[..., 0:1]))


def lv_M(p):
    m2 = p[..., 0] ** 2 - p[..., 1] ** 2 - p[..., 2] ** 2 - p[..., 3] ** 2
    return torch.sqrt(torch.abs(m2))


def lv_omega(p):
    """angle.py:198-213"""
    pb = p[..., 1:4] / p[..., 0:1]
    beta2 = torch.sum(pb * pb, dim=-1)
    beta2 = torch.where(beta2 < 1, beta2, torch.zeros_like(beta2))
    gamma = 1.0 / torch.sqrt(1 - beta2)
    return torch.acosh(gamma)


def lv_neg(p):
    return torch.cat([p[..., 0:1], -p[..., 1:4]], dim=-1)


def angle_zx_z_getx(z1, x1, z2):
    """angle.py:285-308"""
    u_z1 = v_unit(z1)
    u_z2 = v_unit(z2)
    u_y1 = v_cross_unit(z1, x1)
    u_x1 = v_cross_unit(u_y1, z1)
    u_yr = v_cross_unit(z1, z2)
    u_xr = v_cross_unit(u_yr, z1)
    alpha = v_angle_from(u_xr, u_x1, u_y1)
    beta = v_angle_from(u_z2, u_z1, u_xr)
    gamma = torch.zeros_like(beta)
    u_x2 = v_cross_unit(u_yr, u_z2)
    return {"alpha": alpha, "beta": beta, "gamma": gamma}, u_x2


class SU2:
    """angle.py:329-392, a 2x2 complex matrix stored as nested lists."""

    def __init__(self, x):
        self.x = x

    @staticmethod
    def boost_z(omega):
        zeros = torch.zeros_like(omega)
        a = torch.complex(torch.exp(omega / 2), zeros)
        zc = torch.complex(zeros, zeros)
        return SU2([[1 / a, zc], [zc, a]])

    @staticmethod
    def boost_z_from_p(p):
        return SU2.boost_z(lv_omega(p))

    @staticmethod
    def rotation_z(alpha):
        zeros = torch.zeros_like(alpha)
        zc = torch.complex(zeros, zeros)
        a = torch.exp(torch.complex(zeros, alpha / 2))
        return SU2([[1 / a, zc], [zc, a]])

    @staticmethod
    def rotation_y(beta):
        zeros = torch.zeros_like(beta)
        s = torch.complex(torch.sin(beta / 2), zeros)
        c = torch.complex(torch.cos(beta / 2), zeros)
        return SU2([[c, -s], [s, c]])

    def __mul__(self, other):
        x, y = self.x, other.x
        aa = x[0][0] * y[0][0] + x[0][1] * y[1][0]
        ab = x[0][0] * y[0][1] + x[0][1] * y[1][1]
        ba = x[1][0] * y[0][0] + x[1][1] * y[1][0]
        bb = x[1][0] * y[0][1] + x[1][1] * y[1][1]
        return SU2([[aa, ab], [ba, bb]])

    def inv(self):
        (aa, ab), (ba, bb) = self.x
        return SU2([[bb, -ab], [-ba, aa]])

    def get_euler_angle(self):
        x = self.x
        cosbeta = torch.real(x[0][0] * x[1][1] + x[0][1] * x[1][0])
        cosbeta = torch.clamp(cosbeta, -1, 1)
        beta = torch.acos(cosbeta)
        a_p_g = torch.angle(x[1][1])
        a_m_g = -torch.angle(x[1][0])
        return {"alpha": a_p_g + a_m_g, "beta": beta, "gamma": a_p_g - a_m_g}


# --------------------------------------------------------------------------------------
# Model structure: config_loader/decay_config.py, amp/core.py, particle.py
# --------------------------------------------------------------------------------------
class Particle:
    def __init__(self, name, J=0, P=-1, spins=None, mass=None, width=None, model="default",
                 **kw):
        self.name = name
        self.J = _num(J)
        self.P = P
        if spins is None:
            spins = tuple(spin_range(-self.J, self.J))
        self.spins = tuple(_num(s) for s in spins)
        self.mass = mass
        self.width = width
        self.model = model
        self.float = kw.get("float", "")
        # options of the default particle model (amp/core.py:366-377)
        self.running_width = kw.get("running_width", True)
        self.bw_l = kw.get("bw_l", None)
        self.width_norm = kw.get("width_norm", False)
        self.d = 3.0  # Particle.init_params (core.py:380); decay_d may replace it
        self.decays = []

    def __repr__(self):
        return self.name


class Decay:
    def __init__(self, core, outs, p_break=False, l_list=None, ls_list=None, **kw):
        """LS couplings as HelicityDecay.get_ls_list (amp/core.py:1212-1233): ``ls_list``
        replaces the allowed list, ``l_list`` keeps the couplings with those l."""
        self.core = core
        self.outs = tuple(outs)
        self.p_break = p_break
        if ls_list is not None:
            self.ls = [(int(l), _num(sv)) for l, sv in ls_list]
        else:
            self.ls = globals()["ls_list"](core.J, outs[0].J, outs[1].J, core.P, outs[0].P,
                                           outs[1].P, p_break=p_break)
            if l_list is not None:
                self.ls = [(l, sv) for l, sv in self.ls if l in set(int(x) for x in l_list)]
        self.d = 3.0

    def __repr__(self):
        return "{}->{}".format(self.core, "+".join(str(i) for i in self.outs))

    @property
    def params_head(self):
        return _get_name(str(self), "")[:-1]

    @functools.cached_property
    def cg_matrix(self):
        """amp/core.py:851-900 (_get_cg_matrix)."""
        ja, jb, jc = self.core.J, self.outs[0].J, self.outs[1].J
        ret = np.zeros((len(self.ls), len(self.outs[0].spins), len(self.outs[1].spins)))
        for i, (l, s) in enumerate(self.ls):
            for i1, lb in enumerate(self.outs[0].spins):
                for i2, lc in enumerate(self.outs[1].spins):
                    ret[i, i1, i2] = (np.sqrt(2 * l + 1) / np.sqrt(2 * ja + 1)
                                      * cg_coef(jb, jc, lb, -lc, s, lb - lc)
                                      * cg_coef(l, s, 0, lb - lc, ja, lb - lc))
        return ret


def _get_name(head, names):
    """amp/core.py:235-246"""
    return ((str(head) + "_" + names).replace(":", "/").replace("+", ".").replace(",", "")
            .replace("[", "").replace("]", "").replace(" ", ""))


class Chain:
    def __init__(self, decays, top):
        self.decays = decays
        self.top = top
        cores = [d.core for d in decays]
        outs = [o for d in decays for o in d.outs]
        self.inner = sorted([p for p in cores if p in outs], key=lambda p: p.name)
        self.outs = sorted([p for p in outs if p not in cores], key=lambda p: p.name)

    def __repr__(self):
        return "[" + ", ".join(str(d) for d in self.decays) + "]"

    def final_sets(self):
        """sorted_table (particle.py:545-569): particle -> frozenset of finals below it."""
        table = {p: frozenset([p.name]) for p in self.outs}
        pending = list(self.decays)
        while pending:
            rest = []
            for d in pending:
                if all(o in table for o in d.outs):
                    table[d.core] = frozenset().union(*[table[o] for o in d.outs])
                else:
                    rest.append(d)
            pending = rest
        return table

    def topology_key(self):
        t = self.final_sets()
        return tuple(sorted(tuple(sorted(v)) for v in t.values()))


class Model:
    """A config.yml-built decay group (config_loader/decay_config.py:49-427) with the
    variable layout of amp/core.py:1650-1671 / config_loader.py:358-537."""

    def __init__(self, config):
        if isinstance(config, str):
            with open(config) as f:
                config = yaml.safe_load(f)
        self.config = config
        pconf = dict(config["particle"])
        top = pconf.pop("$top")
        finals = pconf.pop("$finals")
        props = {}
        props.update(top)
        props.update(finals)
        props.update(pconf)
        self.particles = {}
        for name, prop in props.items():
            kw = {}
            for k, v in prop.items():
                k2 = {"m0": "mass", "g0": "width", "Par": "P", "bw": "model"}.get(k, k)
                kw[k2] = v
            self.particles[name] = Particle(name, **kw)
        self.top = self.particles[list(top.keys())[0]]
        self.finals = sorted([self.particles[n] for n in finals], key=lambda p: p.name)
        decays = []
        for core, outs in config["decay"].items():
            items = outs if all(isinstance(i, list) for i in outs) else [outs]
            for it in items:
                parts = [x for x in it if not isinstance(x, dict)]
                params = {}
                for x in it:
                    if isinstance(x, dict):
                        params.update(x)
                d = Decay(self.particles[core], [self.particles[p] for p in parts], **params)
                self.particles[core].decays.append(d)
                decays.append(d)
        self.decays = decays
        self.chains = [Chain(c, self.top) for c in self._chain_decay(self.top)]
        res = []
        for c in self.chains:
            for r in c.inner:
                if r not in res:
                    res.append(r)
        self.resonances = res
        self._init_params()

    @staticmethod
    def _chain_decay(p):
        """particle.py:171-181 (chain_decay / cross_combine)."""
        ret = []
        for d in p.decays:
            combos = [[d]]
            for o in d.outs:
                sub = Model._chain_decay(o)
                if sub:
                    combos = [a + b for a in combos for b in sub]
            ret += combos
        return ret

    # ---------------- variables (variable.py:82-330, amp/core.py:380-390,764-773,1371-1375)
    def _add(self, name, value, trainable):
        self.values[name] = float(value)
        if trainable:
            self.trainable.append(name)

    def _init_params(self):
        self.values = {}
        self.trainable = []
        self.var_polar = {}
        for r in self.resonances:
            self._add(r.name + "_mass", r.mass, False)
            if r.width is not None:
                self._add(r.name + "_width", r.width, False)
        inited = set()
        for c in self.chains:
            head = _get_name(c, "total")
            c.total_name = head + "_0"
            self._add(head + "_0r", 1.0, True)
            self._add(head + "_0i", 0.0, True)
            for d in c.decays:
                if id(d) in inited:
                    continue
                inited.add(id(d))
                head = _get_name(d.params_head, "g_ls")
                d.gls_names = [head + "_{}".format(i) for i in range(len(d.ls))]
                for i, n in enumerate(d.gls_names):
                    self._add(n + "r", 1.0 if i == 0 else 1.0, i != 0)
                    self._add(n + "i", 0.0, i != 0)
        # add_decay_constraints (config_loader.py:358-389)
        cons = (self.config.get("constrains") or {}).get("decay") or {}
        decay_d = cons.get("decay_d")
        if decay_d is not None:
            for c in self.chains:
                if isinstance(decay_d, (int, float)):
                    pairs = [(float(decay_d), d) for d in c.decays]
                elif isinstance(decay_d, (list, tuple)):
                    pairs = list(zip([float(x) for x in decay_d], c.decays))
                else:
                    pairs = [(float(decay_d[d.core.name]), d) for _, d in zip(decay_d, c.decays)
                             if d.core.name in decay_d]
                for v, d in pairs:
                    d.d = v
                    d.core.d = v
        idx = cons.get("fix_chain_idx", 0)
        val = cons.get("fix_chain_val", 1.0)
        name = self.chains[idx].total_name
        self.values[name + "r"] = float(val)
        self.values[name + "i"] = 0.0
        for n in (name + "r", name + "i"):
            if n in self.trainable:
                self.trainable.remove(n)
        # add_particle_constraints (config_loader.py:407-537): float: mg frees mass/width
        for c in self.chains:
            for r in c.inner:
                fl = str(r.float or "")
                if "m" in fl and r.name + "_mass" not in self.trainable:
                    self.trainable.append(r.name + "_mass")
                if "g" in fl and r.name + "_width" not in self.trainable:
                    self.trainable.append(r.name + "_width")

    def set_params(self, dic):
        for k, v in dic.items():
            if k in self.values:
                self.values[k] = float(v)

    def trainable_values(self):
        return np.array([self.values[n] for n in self.trainable])

    def var_tensors(self, x=None):
        """name -> scalar tensor; trainable names take the entries of ``x``."""
        out = {k: torch.tensor(v, dtype=DT) for k, v in self.values.items()}
        if x is not None:
            for i, n in enumerate(self.trainable):
                out[n] = x[i]
        return out

    # ---------------- topology bookkeeping (particle.py:902-940)
    def topologies(self):
        """Unique topologies in chain order, each represented by its first chain."""
        reps = []
        for c in self.chains:
            if not any(r.topology_key() == c.topology_key() for r in reps):
                reps.append(c)
        return reps


# --------------------------------------------------------------------------------------
# Helicity angles: cal_angle.py:138-566
# --------------------------------------------------------------------------------------
def cal_angle(model, p4, r_boost=True, final_rest=True, random_z=True, alpha_shift=None,
              ref_topology=None):
    """p4: dict final name -> (N, 4) tensor. Returns per-topology angle data, the masses of
    every particle keyed by the frozenset of finals below it, and per-decay |q|^2.

    Follows cal_angle_from_momentum_single (cal_angle.py:825-865) with
    cal_angle_from_particle (408-490), cal_helicity_angle (262-343), cal_chain_boost
    (186-240) and aligned_angle_ref_rule1 (346-378).

    Two switches exist for the invariance pins only (the reference has neither):
    ``alpha_shift`` maps (topology key, core key, out key) to an angle added to that out's
    folded alpha (a shift by 2 pi moves it to the other SU(2) branch, in the vertex D
    function and in the frame products alike); ``ref_topology`` maps a final's name to the
    topology key whose helicity frame is its reference instead of rule 1's choice.
    """
    p4 = {k: torch.as_tensor(v, dtype=DT) for k, v in p4.items()}
    reps = model.topologies()
    # masses/momenta keyed by frozenset of final names
    mom = {frozenset([f.name]): p4[f.name] for f in model.finals}
    top_key = frozenset(f.name for f in model.finals)
    mom[top_key] = sum(p4[f.name] for f in model.finals)
    for c in reps:
        for key in c.final_sets().values():
            if key not in mom:
                mom[key] = sum(p4[n] for n in sorted(key))
    mass = {k: lv_M(v) for k, v in mom.items()}

    p3 = mom[top_key][..., 1:4]
    base_z = torch.tensor([0.0, 0.0, 1.0], dtype=DT) + torch.zeros_like(p3)
    if random_z:
        mask = (v_norm(p3) < 1e-5)[..., None]
        base_z = torch.where(mask, base_z, p3)
    base_x = torch.tensor([1.0, 0.0, 0.0], dtype=DT)

    topo = {}
    for c in reps:
        fs = c.final_sets()
        key = lambda p: fs[p]
        # cal_chain_boost
        part = {}
        core_decay = {}
        dset = list(c.decays)
        pset = set(key(p) for p in c.inner) | set(key(p) for p in c.outs)
        while dset:
            rest = []
            for d in dset:
                if d.core is c.top:
                    p_rest = mom[top_key]
                    rp = {}
                    for j in d.outs:
                        core_decay[key(j)] = d
                        rp[key(j)] = lv_rest_vector(p_rest, mom[key(j)])
                        pset.discard(key(j))
                    for j in pset:
                        rp[j] = lv_rest_vector(p_rest, mom[j])
                    part[id(d)] = rp
                elif key(d.core) in core_decay:
                    up = part[id(core_decay[key(d.core)])]
                    p_rest = up[key(d.core)]
                    rp = {}
                    for j in d.outs:
                        core_decay[key(j)] = d
                        rp[key(j)] = lv_rest_vector(p_rest, up[key(j)])
                        pset.discard(key(j))
                    for j in pset:
                        rp[j] = lv_rest_vector(p_rest, up[j])
                    part[id(d)] = rp
                else:
                    rest.append(d)
            dset = rest
        # cal_helicity_angle
        set_x = {top_key: base_x}
        set_z = {top_key: base_z}
        r_mat, b_mat = {}, {}
        ang = {}
        dset = list(c.decays)
        while dset:
            rest = []
            for d in dset:
                kc = key(d.core)
                if kc in set_x:
                    ang[kc] = {}
                    bias = -np.pi
                    for j in d.outs:
                        p_rest = part[id(d)][key(j)]
                        z2 = lv_vect(p_rest)
                        a, x = angle_zx_z_getx(set_z[kc], set_x[kc], z2)
                        set_x[key(j)] = x
                        set_z[key(j)] = z2
                        a["alpha"] = torch.remainder(a["alpha"] - bias, 2 * np.pi) + bias
                        if alpha_shift:
                            a["alpha"] = a["alpha"] + alpha_shift.get(
                                (c.topology_key(), kc, key(j)), 0.0)
                        bias -= np.pi
                        ang[kc][key(j)] = a
                        b_mat[key(j)] = SU2.boost_z_from_p(p_rest)
                        r = SU2.rotation_y(a["beta"]) * SU2.rotation_z(a["alpha"])
                        if kc in r_mat:
                            r_mat[key(j)] = r * b_mat[kc] * r_mat[kc]
                        else:
                            r_mat[key(j)] = r
                else:
                    rest.append(d)
            dset = rest
        topo[c.topology_key()] = {"ang": ang, "r": r_mat, "b": b_mat, "chain": c}

    # aligned_angle_ref_rule1: reference chain of each final = first topology where it is a
    # direct daughter of the top, else the first topology.
    ref = {}
    for c in reps:
        for d in c.decays:
            if d.core is c.top:
                for o in d.outs:
                    if o in model.finals and o.name not in ref:
                        ref[o.name] = c.topology_key()
    for f in model.finals:
        if f.name not in ref:
            ref[f.name] = reps[0].topology_key()
    if ref_topology:
        ref.update(ref_topology)
    aligned = {}
    for c in reps:
        tk = c.topology_key()
        aligned[tk] = {}
        for f in model.finals:
            if ref[f.name] == tk:
                continue
            fk = frozenset([f.name])
            r_matrix = topo[tk]["r"][fk]
            b_matrix = topo[tk]["b"][fk]
            r_ref = topo[ref[f.name]]["r"][fk]
            b_ref = topo[ref[f.name]]["b"][fk]
            if r_boost:
                R = r_ref * r_matrix.inv()
                if final_rest:
                    R = b_ref * R * b_matrix.inv()
                aligned[tk][f.name] = R.get_euler_angle()
            else:
                raise NotImplementedError("r_boost=False is not restated")
    return {"topo": topo, "aligned": aligned, "mass": mass, "top_key": top_key}


# --------------------------------------------------------------------------------------
# Amplitude: amp/core.py:399-418, 931-1164, 1392-1445, 1563-1604, 1696-1728, 1919-1931
# --------------------------------------------------------------------------------------
def _letters():
    return iter("abcdefghjklmnopqrstuvwxyz")


def particle_amp(r, m, m0, width, q, q0):
    """Particle.get_amp (amp/core.py:399-418): 1 without a width; BW (constant width) when
    running_width is false, else BWR with L = bw_l or the smallest l of the particle's first
    decay and its barrier radius d; times the width when width_norm. ParticleBW (model BW,
    amp/base.py:262-266) is BW alone. Other particle models are not restated."""
    model = r.model or "default"
    if model not in ("default", "BWR", "BW"):
        raise NotImplementedError("particle model {}".format(model))
    if r.width is None:
        return torch.ones_like(m)
    if model == "BW":
        return BW(m, m0, width)
    if not r.running_width:
        ret = BW(m, m0, width)
    else:
        L = r.bw_l if r.bw_l is not None else min(l for l, s in r.decays[0].ls)
        ret = BWR(m, m0, width, q, q0, L, r.d)
    if r.width_norm:
        ret = torch.complex(width, torch.zeros_like(width)) * ret
    return ret


def aligned_D(al, f):
    """The aligned-angle factor D^{J_f*}_{lambda, lambda'}(alpha, beta, gamma) of a spin>0
    final (amp/core.py:1414-1429), indexed [event][lambda in the chain's frame][lambda' in
    the reference frame]."""
    return get_D_matrix_lambda(al, f.J, f.spins, f.spins)


def chain_amplitude(model, chain, angles, var):
    """DecayChain.get_amp for one chain: (N, n_top, n_final_0, ...) complex, finals in
    sorted-name order."""
    tk = chain.topology_key()
    rep = angles["topo"][tk]["chain"]
    fs = chain.final_sets()
    mass = angles["mass"]
    ang_t = angles["topo"][tk]["ang"]
    # einsum letters: top, sorted finals, then inner (amp/core.py:1652-1666, 2032-2041)
    gen = _letters()
    letter = {model.top.name: next(gen)}
    for f in model.finals:
        letter[f.name] = next(gen)
    for r in chain.inner:
        letter[r.name] = next(gen)
    amps, idxs = [], []
    for d in chain.decays:
        a, b, c = d.core, d.outs[0], d.outs[1]
        ka = fs[a] if a is not chain.top else angles["top_key"]
        ang = ang_t[ka][fs[b]]
        D = get_D_matrix_lambda(ang, a.J, a.spins, b.spins, c.spins)
        # get_ls_amp (1025-1043) + get_barrier_factor2 (1072-1097)
        g = [torch.complex(var[n + "r"] * torch.cos(var[n + "i"]),
                           var[n + "r"] * torch.sin(var[n + "i"])) for n in d.gls_names]
        m0 = mass[ka] if a is not chain.top else mass[angles["top_key"]]
        q2 = get_relative_p2(m0, mass[fs[b]], mass[fs[c]])

        def pmass(p):
            if p in chain.inner:
                return var[p.name + "_mass"]
            return torch.tensor(p.mass, dtype=DT)

        q02 = get_relative_p2(pmass(a), pmass(b), pmass(c))
        bfs = []
        for l, s in d.ls:
            bp = Bprime_q2(l, q2, q02, d.d)
            bfs.append((q2 ** (l / 2) * bp).reshape(-1, 1))
        bf = torch.cat(bfs, dim=-1)
        m_dep = torch.stack(g) * torch.complex(bf, torch.zeros_like(bf))
        cg = torch.as_tensor(d.cg_matrix, dtype=DT)
        cg = torch.complex(cg, torch.zeros_like(cg))
        H = torch.sum(m_dep.reshape(-1, len(d.ls), 1, 1) * cg, dim=1)
        amps.append(H[:, None] * D)
        idxs.append(letter[a.name] + letter[b.name] + letter[c.name])
    # get_amp_particle (1563-1604): BWR of every inner particle
    rs = None
    for r in chain.inner:
        dec = [d for d in chain.decays if d.core is r][0]
        y, z = dec.outs
        m = mass[fs[r]]
        q = get_relative_p(m, mass[fs[y]], mass[fs[z]])
        m0 = var[r.name + "_mass"]
        q0 = get_relative_p(m0, torch.tensor(y.mass, dtype=DT), torch.tensor(z.mass, dtype=DT))
        amp_r = particle_amp(r, m, m0, var.get(r.name + "_width"), q, q0)
        if not torch.is_complex(amp_r):
            amp_r = torch.complex(amp_r, torch.zeros_like(amp_r))
        rs = amp_r if rs is None else rs * amp_r
    n = chain.total_name
    total = torch.complex(var[n + "r"] * torch.cos(var[n + "i"]),
                          var[n + "r"] * torch.sin(var[n + "i"]))
    total = total * rs if rs is not None else total
    amps.append(total)
    idxs.append("")
    final_idx = letter[model.top.name] + "".join(letter[f.name] for f in model.finals)
    # aligned D for spin>0 finals (1414-1429)
    for f in model.finals:
        if f.J == 0:
            continue
        al = angles["aligned"][tk].get(f.name)
        if al is None:
            continue
        dt = aligned_D(al, f)
        amps.append(dt)
        lo = letter[f.name]
        idxs.append(lo + lo.upper())
        final_idx = final_idx.replace(lo, lo.upper())
    expr = ",".join("..." + i for i in idxs) + "->..." + final_idx
    return torch.einsum(expr, *amps)


def amplitude(model, angles, var):
    """DecayGroup.get_amp (amp/core.py:1696-1728): coherent sum over chains."""
    ret = None
    for c in model.chains:
        a = chain_amplitude(model, c, angles, var)
        ret = a if ret is None else ret + a
    return ret


def sum_amp(model, angles, var):
    """DecayGroup.sum_amp (amp/core.py:1919-1931): sum over helicities of |A|^2."""
    amp = amplitude(model, angles, var)
    amp2 = torch.real(amp * torch.conj(amp))
    return torch.sum(amp2, dim=list(range(1, amp2.dim())))


# --------------------------------------------------------------------------------------
# Likelihood: model/model.py:41-94, 407-449, 603-641, 1126-1229
# --------------------------------------------------------------------------------------
def loaded_weight(idx, n, weight_option):
    """The weight a config-loaded sample carries (SimpleData.load_extra_var / load_data,
    config_loader/data.py:216-251): ``<idx>_weight`` as one number for every event or the
    values of weight files, default 1, times -1 for names matching "bg*" (get_weight_sign,
    149-155; re.match of "bg*" is any name that starts with b)."""
    if weight_option is None:
        w = np.ones(n)
    elif isinstance(weight_option, (int, float)):
        w = np.ones(n) * weight_option
    else:
        files = [weight_option] if isinstance(weight_option, str) else weight_option
        w = np.concatenate([np.loadtxt(f).reshape((-1,)) for f in files])[:n]
    sign = -1 if re.match("bg*", idx) else 1
    return sign * w


def get_weight_data(n_data, weight=None, bg_n=None, bg_weight=None, w_bkg=0.0, alpha=True):
    """Model.get_weight_data (model/model.py:603-641): data weights (1 when None), then the
    background's own weights, or -w_bkg for each background event when it has none, and
    the alpha = sum w / sum w^2 rescale."""
    w = np.ones(n_data) if weight is None else np.asarray(weight, dtype=np.float64)
    if bg_n is not None:
        bw = np.full(bg_n, -w_bkg) if bg_weight is None else np.asarray(bg_weight, float)
        w = np.concatenate([w, bw])
    if alpha:
        w = w * (w.sum() / (w * w).sum())
    return torch.as_tensor(w, dtype=DT)


def fcn_weights(n_data, weight=None, n_mc=None, mc_weight=None):
    """FCN.__init__ (model.py:1126-1164) + get_weight_data (603-641)."""
    w = np.ones(n_data) if weight is None else np.asarray(weight, dtype=np.float64)
    alpha = w.sum() / (w * w).sum()
    w = alpha * w
    wm = np.ones(n_mc) if mc_weight is None else np.asarray(mc_weight, dtype=np.float64)
    wm = wm / wm.sum()
    return torch.as_tensor(w, dtype=DT), torch.as_tensor(wm, dtype=DT)


def nll_fn(model, ang_data, ang_mc, w, wm):
    """Returns x -> NLL following BaseModel.nll_grad_batch (model.py:407-449) with
    _batch_sum (41-54) at resolution_size = 1."""

    def f(x):
        var = model.var_tensors(x)
        fd = sum_amp(model, ang_data, var)
        fm = sum_amp(model, ang_mc, var)
        dom = torch.where(w == 0, torch.ones_like(w), w)
        ln = torch.sum(w * clip_log(fd * w / dom))
        int_mc = torch.sum(wm * fm)
        sw = torch.sum(w)
        return -ln + sw * torch.log(int_mc)

    return f


def nll_grad(model, ang_data, ang_mc, w, wm, x=None):
    x = torch.tensor(model.trainable_values() if x is None else x, dtype=DT,
                     requires_grad=True)
    nll = nll_fn(model, ang_data, ang_mc, w, wm)(x)
    (g,) = torch.autograd.grad(nll, x)
    return float(nll.detach()), g.numpy()


def nll_grad_hessian(model, ang_data, ang_mc, w, wm, x=None):
    x0 = torch.tensor(model.trainable_values() if x is None else x, dtype=DT)
    f = nll_fn(model, ang_data, ang_mc, w, wm)
    x = x0.clone().requires_grad_(True)
    nll = f(x)
    (g,) = torch.autograd.grad(nll, x, create_graph=True)
    rows = []
    for i in range(len(g)):
        (h,) = torch.autograd.grad(g[i], x, retain_graph=True)
        rows.append(h)
    return float(nll.detach()), g.detach().numpy(), torch.stack(rows).numpy()


# --------------------------------------------------------------------------------------
# Phase space and toys: phasespace.py:31-253, applications.py:394-510
# --------------------------------------------------------------------------------------
def _get_p(M, ma, mb):
    p2 = (M * M - (ma + mb) ** 2) * (M * M - (ma - mb) ** 2)
    p2 = np.where(p2 <= 0, 0.0, p2)
    return np.sqrt(p2) / (2.0 * M)


def _np_rest_vector(p_frame, other):
    return lv_rest_vector(torch.as_tensor(p_frame), torch.as_tensor(other)).numpy()


def phsp_3body(m0, masses, n, rng):
    """PhaseSpaceGenerator.generate (phasespace.py:93-202) for three final particles:
    m12 uniform, accept with the GENBOD weight, then sequential two-body decays."""
    m1, m2, m3 = masses
    a, b = m3 + m2, m0 - m1
    wt_max = _get_p(m0, m1 + 0.0, b) * _get_p(b, m2, m3)
    # max weight of q(m0; m1, m12) q(m12; m2, m3) over m12 found numerically
    grid = np.linspace(a, b, 20001)
    wt_max = (_get_p(m0, m1, grid) * _get_p(grid, m2, m3)).max() * 1.001
    out = []
    got = 0
    while got < n:
        k = int((n - got) * 2.2) + 1000
        m12 = rng.uniform(a, b, k)
        w = _get_p(m0, m1, m12) * _get_p(m12, m2, m3) / wt_max
        m12 = m12[w > rng.uniform(0, 1, k)]
        out.append(m12)
        got += len(m12)
    m12 = np.concatenate(out)[:n]

    def two_body(M, ma, mb, k):
        cos_t = 2 * rng.uniform(0, 1, k) - 1
        sin_t = np.sqrt(1 - cos_t * cos_t)
        phi = 2 * np.pi * rng.uniform(0, 1, k)
        q = _get_p(M, ma, mb)
        pa = np.stack([np.sqrt(q * q + ma * ma), q * sin_t * np.cos(phi),
                       q * sin_t * np.sin(phi), q * cos_t], axis=-1)
        pb = np.stack([np.sqrt(q * q + mb * mb), -pa[:, 1], -pa[:, 2], -pa[:, 3]], axis=-1)
        return pa, pb

    # m12 -> m2 + m3 in the m12 rest frame, then m0 -> m1 + m12
    p2, p3 = two_body(m12, m2, m3, n)
    p1, p12 = two_body(m0, m1, m12, n)
    p2 = _np_rest_vector(lv_neg(torch.as_tensor(p12)).numpy(), p2)
    p3 = _np_rest_vector(lv_neg(torch.as_tensor(p12)).numpy(), p3)
    return [p1, p2, p3]


def gen_mc(model, n, seed):
    rng = np.random.default_rng(seed)
    masses = [model.particles[n_].mass for n_ in model.config["data"]["dat_order"]]
    ps = phsp_3body(model.top.mass, masses, n, rng)
    return dict(zip(model.config["data"]["dat_order"], ps))


def random_couplings(model, seed):
    """|g| ~ U(0.5, 2), arg g ~ U(-pi, pi) for every free complex coupling."""
    rng = np.random.default_rng(seed)
    vals = {}
    for n in model.trainable:
        if n.endswith("r") and n[:-1] + "i" in model.trainable:
            vals[n] = rng.uniform(0.5, 2.0)
            vals[n[:-1] + "i"] = rng.uniform(-np.pi, np.pi)
    return vals


def gen_data(model, n, seed, var=None, batch=200000):
    """Accept-reject toy sampling on |A|^2 (applications.py:394-487)."""
    rng = np.random.default_rng(seed)
    var = model.var_tensors() if var is None else var
    probe = gen_mc(model, 20000, seed + 1)
    fmax = float(sum_amp(model, cal_angle(model, probe), var).max()) * 1.2
    keep = {k: [] for k in model.config["data"]["dat_order"]}
    got = 0
    it = 0
    while got < n:
        cand = gen_mc(model, batch, seed + 2 + it)
        it += 1
        f = sum_amp(model, cal_angle(model, cand), var).numpy()
        sel = f > rng.uniform(0, fmax, len(f))
        for k in keep:
            keep[k].append(cand[k][sel])
        got += int(sel.sum())
    return {k: np.concatenate(v)[:n] for k, v in keep.items()}
