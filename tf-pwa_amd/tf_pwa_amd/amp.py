"""Amplitude model: the config-built decay group, its parameters, and its lowering to the
engine's model description.

The amplitude of tf-pwa (tf_pwa/amp/core.py:1392-1445, 1696-1728, 1919-1931) is

    A_lambda(x) = sum_chains total_c * BW_c(m) * sum_{ls1, ls2} g1_ls1 g2_ls2 *
                  [q^l B'_l CG D]_vertex1 [q^l B'_l CG D]_vertex2 (contracted over the
                  resonance helicity, rotated by the aligned D of spin>0 finals).

The engine stores, per event, the parameter-free partial-wave amplitudes h_k,lambda with
k = (chain, ls1, ls2) and evaluates A_lambda = sum_k c_k h_k,lambda with host coefficients
c_k = total_c g1_ls1 g2_ls2 sqrt(P_l1(z0_1)) sqrt(P_l2(z0_2)). The barrier factor
q^l sqrt(P_l(z0)/P_l(z)) of Bprime_q2 (tf_pwa/breit_wigner.py:268-277) splits into the
per-event q^l / sqrt(P_l(z)) and the parameter constant sqrt(P_l(z0)); when P_l(z0) <= 0 the
reference clamps the ratio to 1 and the per-event factor becomes q^l (flag ``act`` = 0).
"""

import ctypes

import numpy as np

from . import _lib
from .particle import ls_couplings  # noqa: F401  (re-exported for callers)


def rel_p2(m0, m1, m2):
    """get_relative_p2 (tf_pwa/amp/core.py:324-334)."""
    s, d = m1 + m2, m1 - m2
    return (m0 - s) * (m0 + s) * (m0 - d) * (m0 + d) / (2 * m0) ** 2


def rel_p(m0, m1, m2):
    """get_relative_p (tf_pwa/amp/core.py:311-321) with m_eff = max(m0, m1+m2)."""
    s, d = m1 + m2, m1 - m2
    me = m0 if m0 > s else s
    return np.sqrt((me - s) * (me + s) * (me - d) * (me + d)) / (2 * me)


_BPRIME = {0: [1.0], 1: [1.0, 1.0], 2: [1.0, 3.0, 9.0], 3: [1.0, 6.0, 45.0, 225.0],
           4: [1.0, 10.0, 135.0, 1575.0, 11025.0],
           5: [1.0, 15.0, 315.0, 6300.0, 99225.0, 893025.0]}


def bprime_poly(l, z):
    """|theta_l(i w)|^2 of the reverse Bessel polynomial (tf_pwa/breit_wigner.py:332-355)."""
    r = 0.0
    for c in _BPRIME[l]:
        r = r * z + c
    return r


def bprime_poly_deriv(l, z):
    cof = _BPRIME[l]
    n = len(cof) - 1
    r = 0.0
    for i, c in enumerate(cof[:-1]):
        r = r * z + c * (n - i)
    return r


def bprime_poly_deriv2(l, z):
    c = _BPRIME[l]
    n = len(c) - 1
    return sum(c[i] * (n - i) * (n - i - 1) * z ** (n - i - 2) for i in range(n - 1))


def _rel_p2_d2_parent(M, m1, m2):
    """d2 q^2 / dM^2 of rel_p2(M, m1, m2) = M^2/4 - (s^2 + d^2)/4 + s^2 d^2 / (4 M^2)."""
    s, d = m1 + m2, m1 - m2
    return 0.5 + 1.5 * s * s * d * d / M ** 4


def _rel_p2_d2_daughter(M, m1, m2, which):
    """d2 q^2 / dm_which^2 of rel_p2(M, m1, m2) (which 0: m1, 1: m2)."""
    s, d = m1 + m2, m1 - m2
    sig = 1.0 if which == 0 else -1.0
    return (-2.0 * (M * M - d * d) - 2.0 * (M * M - s * s) + 8.0 * sig * s * d) / (4.0 * M * M)


def _rel_p2_d_parent(M, m1, m2):
    """d rel_p2 / d M."""
    s, d = m1 + m2, m1 - m2
    F = (M * M - s * s) * (M * M - d * d)
    dF = 2 * M * (M * M - d * d) + 2 * M * (M * M - s * s)
    return dF / (4 * M * M) - F / (2 * M ** 3)


def _rel_p2_d_daughter(M, m1, m2, which):
    """d rel_p2 / d m1 (which = 0) or d m2 (which = 1)."""
    s, d = m1 + m2, m1 - m2
    sd = 1.0 if which == 0 else -1.0
    dF = -2 * s * (M * M - d * d) - 2 * sd * d * (M * M - s * s)
    return dF / (4 * M * M)


# Lineshape codes of the device description (TpaChain.shape, Tpa4Chain.r_shape): bits 0-1 the
# propagator (1 BW with constant width, 2 BWR with running width), bit 2 the width_norm factor.
SHAPE_NONE, SHAPE_BW, SHAPE_BWR, SHAPE_WIDTH_NORM = 0, 1, 2, 4
LINESHAPE_MODELS = ("default", "BWR", "BW")


def lineshape_code(R):
    """The lineshape of resonance R as the default particle model computes it
    (Particle.get_amp, tf_pwa/amp/core.py:399-418): 1 without a width; BW when
    ``running_width`` is false, else BWR with the running width; times the width when
    ``width_norm``. ``model: BW`` (ParticleBW, tf_pwa/amp/base.py:245-274) is BW and ignores
    both options. Other particle models are not implemented and raise."""
    model = R.model or "default"
    if model not in LINESHAPE_MODELS:
        raise NotImplementedError("particle model {!r} of {} (lineshapes implemented: {})".format(
            model, R.name, ", ".join(LINESHAPE_MODELS)))
    if R.width is None:
        return SHAPE_NONE
    if model == "BW":
        return SHAPE_BW
    code = SHAPE_BWR if R.kwargs.get("running_width", True) else SHAPE_BW
    if R.kwargs.get("width_norm", False):
        code |= SHAPE_WIDTH_NORM
    return code


def running_width_l(R):
    """Orbital L of the running width: ``bw_l`` when given, else the smallest l of R's first
    decay (core.py:409-411)."""
    if R.kwargs.get("bw_l") is not None:
        return int(R.kwargs["bw_l"])
    return min(R.decays[0].l_list) if R.decays else 0


class PartialWave:
    """One partial wave: a chain and the LS index at each of its vertices."""

    __slots__ = ("chain", "ls")

    def __init__(self, chain, *ls):
        self.chain, self.ls = chain, tuple(ls)

    @property
    def ls1(self):
        return self.ls[0]

    @property
    def ls2(self):
        return self.ls[1]


class AmplitudeModel:
    """A decay group plus its VarsManager, lowered for the engine (three-body finals)."""

    def __init__(self, decay_group, vm, random_z=True):
        self.decay_group = decay_group
        self.vm = vm
        self.random_z = bool(random_z)
        self._dev = None
        self._lower()

    # ---------------------------------------------------------------- lowering
    def _lower(self):
        g = self.decay_group
        self.n_final = len(g.outs)
        if self.n_final == 4:
            return self._lower4()
        if self.n_final != 3:
            raise NotImplementedError("the engine supports three- and four-body final states")
        self.finals = list(g.outs)
        fidx = {p.name: i for i, p in enumerate(self.finals)}
        reps = g.topologies()
        self.reps = reps
        topo = []
        for rep in reps:
            d1 = rep.decay_of(g.top)
            R = [o for o in d1.outs if o in rep.inner]
            if len(R) != 1 or len(rep) != 2:
                raise NotImplementedError("only A -> R X, R -> Y Z chains are supported")
            R = R[0]
            X = d1.outs[1] if d1.outs[0] is R else d1.outs[0]
            d2 = rep.decay_of(R)
            topo.append(dict(x=fidx[X.name], y=fidx[d2.outs[0].name], z=fidx[d2.outs[1].name],
                             r_first=int(d1.outs[0] is R)))
        ref_topo = []
        for f in range(3):
            cand = [t for t, T in enumerate(topo) if T["x"] == f]
            ref_topo.append(cand[0] if cand else 0)
        self.topo = topo
        self.ref_topo = ref_topo
        chains = []
        cg = []
        k0 = 0
        waves = []
        for ci, c in enumerate(g.chains):
            key = c.topology_key()
            t = [i for i, r in enumerate(reps) if r.topology_key() == key][0]
            d1 = c.decay_of(g.top)
            R = [o for o in d1.outs if o in c.inner][0]
            d2 = c.decay_of(R)
            shape = lineshape_code(R)
            cg1 = d1.cg_matrix()
            cg2 = d2.cg_matrix()
            ch = dict(topo=t, top_r_first=int(d1.outs[0] is R),
                      r_y_first=int(fidx[d2.outs[0].name] == topo[t]["y"]),
                      R=R, d1=d1, d2=d2, k0=k0, shape=shape, bw_l=running_width_l(R),
                      cg1_off=len(cg), cg2_off=len(cg) + cg1.size)
            cg.extend(cg1.ravel().tolist())
            cg.extend(cg2.ravel().tolist())
            for s1 in range(len(d1.ls)):
                for s2 in range(len(d2.ls)):
                    waves.append(PartialWave(ci, s1, s2))
            k0 += len(d1.ls) * len(d2.ls)
            chains.append(ch)
        self.chains = chains
        self.waves = waves
        self.P = k0
        self.cg = np.array(cg, dtype=np.float64)
        lam = len(g.top.spins)
        for p in self.finals:
            lam *= len(p.spins)
        self.lam = lam
        lp = 1
        while lp < lam:
            lp <<= 1
        self.lam_pad = lp

    # ---------------------------------------------------------------- four-body lowering
    def _lower4(self):
        """Trees of three two-body vertices over four spin-0 finals: vertex 0 the top decay,
        vertices 1 and 2 the decays of the two resonances, in the representative chain's
        order (quasi-two-body A -> R1 R2 or sequential A -> R1 f, R1 -> R2 f)."""
        g = self.decay_group
        self.finals = list(g.outs)
        for p in self.finals:
            if p.J != 0:
                raise NotImplementedError("four-body final states must be spin 0")
        fidx = {p.name: i for i, p in enumerate(self.finals)}

        def masks(chain):
            return {p: sum(1 << fidx[n] for n in names)
                    for p, names in chain.finals_below().items()}

        reps = g.topologies()
        self.reps = reps
        topo = []
        for rep in reps:
            if len(rep) != 3:
                raise NotImplementedError("four-body chains need three two-body vertices")
            mk = masks(rep)
            d0 = rep.decay_of(g.top)
            inner0 = [o for o in d0.outs if o in rep.inner]
            if len(inner0) == 2:
                v1, v2 = d0.outs
                parent, pout = [-1, 0, 0], [0, 0, 1]
            elif len(inner0) == 1:
                v1 = inner0[0]
                d1 = rep.decay_of(v1)
                inner1 = [o for o in d1.outs if o in rep.inner]
                if len(inner1) != 1:
                    raise NotImplementedError("unsupported four-body chain")
                v2 = inner1[0]
                parent, pout = [-1, 0, 1], [0, d0.outs.index(v1), d1.outs.index(v2)]
            else:
                raise NotImplementedError("unsupported four-body chain")
            decs = [d0, rep.decay_of(v1), rep.decay_of(v2)]
            topo.append(dict(core=[mk[d.core] for d in decs],
                             out=[[mk[o] for o in d.outs] for d in decs],
                             parent=parent, parent_out=pout))
        self.topo = topo
        chains, cg, waves, k0 = [], [], [], 0
        for ci, c in enumerate(g.chains):
            t = [i for i, r in enumerate(reps) if r.topology_key() == c.topology_key()][0]
            T = topo[t]
            mk = masks(c)
            by_core = {mk[d.core]: d for d in c.decays}
            decs = [by_core[m] for m in T["core"]]
            R = [decs[1].core, decs[2].core]

            def slot(p):
                if p is R[0]:
                    return 1
                if p is R[1]:
                    return 2
                return 3 + fidx[p.name]

            cg_off = []
            for d in decs:
                cg_off.append(len(cg))
                cg.extend(d.cg_matrix().ravel().tolist())

            def nominal(p):
                return p.mass

            res = []
            for r in range(2):
                Rr, dr = R[r], decs[1 + r]
                res.append(dict(R=Rr, shape=lineshape_code(Rr), bw_l=running_width_l(Rr),
                                q0m=[nominal(o) for o in dr.outs]))
            n_ls = [len(d.ls) for d in decs]
            chains.append(dict(topo=t, decs=decs, R=R, res=res, k0=k0, cg_off=cg_off,
                               swap=[int(mk[decs[v].outs[0]] != T["out"][v][0])
                                     for v in range(3)],
                               out_slot=[[slot(o) for o in d.outs] for d in decs],
                               n_ls=n_ls))
            for s0 in range(n_ls[0]):
                for s1 in range(n_ls[1]):
                    for s2 in range(n_ls[2]):
                        waves.append(PartialWave(ci, s0, s1, s2))
            k0 += n_ls[0] * n_ls[1] * n_ls[2]
        self.chains = chains
        self.waves = waves
        self.P = k0
        self.cg = np.array(cg, dtype=np.float64)
        self.lam = len(g.top.spins)
        lp = 1
        while lp < self.lam:
            lp <<= 1
        self.lam_pad = lp

    def _describe4(self):
        M = _lib.TpaModel4()
        g = self.decay_group
        M.n_final = 4
        M.n_topo = len(self.topo)
        M.n_chain = len(self.chains)
        M.P = self.P
        M.lam = self.lam
        M.lam_pad = self.lam_pad
        M.top_j2 = int(round(2 * g.top.J))
        M.n_top = len(g.top.spins)
        for i, sp in enumerate(g.top.spins):
            M.hel_top[i] = int(round(2 * sp))
        M.random_z = int(self.random_z)
        M.n_cg = len(self.cg)
        for t, T in enumerate(self.topo):
            for v in range(3):
                M.topo[t].core[v] = T["core"][v]
                M.topo[t].out[v][0], M.topo[t].out[v][1] = T["out"][v]
                M.topo[t].parent[v] = T["parent"][v]
                M.topo[t].parent_out[v] = T["parent_out"][v]
        bc = self.barrier_constants()
        shapes = self.shape_params()
        for c, ch in enumerate(self.chains):
            C = M.chain[c]
            C.topo = ch["topo"]
            C.k0 = ch["k0"]
            for v in range(3):
                C.swap[v] = ch["swap"][v]
                C.out_slot[v][0], C.out_slot[v][1] = ch["out_slot"][v]
                C.n_ls[v] = ch["n_ls"][v]
                for i, l in enumerate(ch["decs"][v].l_list):
                    C.l[v][i] = l
                    C.act[v][i] = bc[c]["act"][v][i]
                C.cg_off[v] = ch["cg_off"][v]
            for r, rr in enumerate(ch["res"]):
                R = rr["R"]
                C.r_j2[r] = int(round(2 * R.J))
                C.r_n[r] = len(R.spins)
                for i, sp in enumerate(R.spins):
                    C.r_hel[r][i] = int(round(2 * sp))
                C.r_shape[r] = rr["shape"]
                C.r_bw_l[r] = rr["bw_l"]
                C.r_m0[r], C.r_g0[r] = shapes[c][2 * r], shapes[c][2 * r + 1]
                C.r_q0m[r][0], C.r_q0m[r][1] = rr["q0m"]
            C.d = ch["decs"][0].d
        return M

    def _barrier_constants4(self):
        """Per chain and vertex: sqrt(P_l(z0)) at the nominal masses (resonances at their
        mass parameters), or 1 with act = 0 when P_l(z0) <= 0."""
        top = self.decay_group.top
        out = []
        for ch, sh in zip(self.chains, self.shape_params()):
            m0 = {id(ch["R"][0]): sh[0], id(ch["R"][1]): sh[2], id(top): top.mass}

            def mass(p):
                return m0.get(id(p), p.mass)

            cs, acts = [], []
            for d in ch["decs"]:
                q02 = rel_p2(mass(d.core), mass(d.outs[0]), mass(d.outs[1]))
                c_v, a_v = [], []
                for l in d.l_list:
                    p0 = bprime_poly(l, q02 * d.d * d.d)
                    a_v.append(int(p0 > 0))
                    c_v.append(np.sqrt(p0) if p0 > 0 else 1.0)
                cs.append(np.array(c_v))
                acts.append(a_v)
            out.append(dict(c=cs, act=acts, dlog=[np.zeros(len(x)) for x in cs]))
        return out

    def _names(self):
        """Variable names of every chain: total, the g_ls of both vertices, mass and width
        (fixed by the decay structure, so built once)."""
        if getattr(self, "_names_cache", None) is None:
            self._names_cache = self._names_now()
        return self._names_cache

    def _names_now(self):
        if self.n_final == 4:
            out = []
            for c, ch in zip(self.decay_group.chains, self.chains):
                g = [[d.params_head() + "_g_ls_{}".format(i) for i in range(len(d.ls))]
                     for d in ch["decs"]]
                R1, R2 = ch["R"]
                out.append(dict(total=c.params_head() + "_total_0", g=g, g1=g[0], g2=g[1],
                                mass=R1.name + "_mass", width=R1.name + "_width",
                                mass2=R2.name + "_mass", width2=R2.name + "_width"))
            return out
        out = []
        for c, ch in zip(self.decay_group.chains, self.chains):
            g1 = [ch["d1"].params_head() + "_g_ls_{}".format(i) for i in range(len(ch["d1"].ls))]
            g2 = [ch["d2"].params_head() + "_g_ls_{}".format(i) for i in range(len(ch["d2"].ls))]
            out.append(dict(total=c.params_head() + "_total_0", g1=g1, g2=g2, g=[g1, g2],
                            mass=ch["R"].name + "_mass", width=ch["R"].name + "_width"))
        return out

    def shape_params(self):
        """(m0, g0) of every chain's resonance from the parameter store (four-body chains:
        (m0, g0, m0', g0') of their two resonances), cached by the store's version."""
        vm = self.vm
        if getattr(self, "_sp_version", None) == vm.version:
            return self._sp
        self._sp = self._shape_params_now()
        self._sp_version = vm.version
        return self._sp

    def _shape_params_now(self):
        vm = self.vm
        if self.n_final == 4:
            out = []
            for ch in self.chains:
                t = ()
                for R in ch["R"]:
                    m0 = vm.variables.get(R.name + "_mass", R.mass)
                    g0 = vm.variables.get(R.name + "_width", R.width)
                    t += (float(m0), float(g0) if g0 is not None else 0.0)
                out.append(t)
            return out
        out = []
        for ch in self.chains:
            R = ch["R"]
            m0 = vm.variables.get(R.name + "_mass", R.mass)
            g0 = vm.variables.get(R.name + "_width", R.width)
            out.append((float(m0), float(g0) if g0 is not None else 0.0))
        return out

    def barrier_constants(self):
        """Per chain: sqrt(P_l(z0)) (or 1 when clamped) and the act flags of both vertices,
        cached by the lineshape parameters they depend on."""
        key = tuple(self.shape_params())
        if getattr(self, "_bc_key", None) != key:
            self._bc = self._barrier_constants()
            self._bc_key = key
        return self._bc

    def _barrier_constants(self):
        if self.n_final == 4:
            return self._barrier_constants4()
        top = self.decay_group.top
        out = []
        for ch, (m0, g0) in zip(self.chains, self.shape_params()):
            d1, d2 = ch["d1"], ch["d2"]

            def pm(p):
                return m0 if p is ch["R"] else p.mass

            d = d1.d
            q01 = rel_p2(top.mass, pm(d1.outs[0]), pm(d1.outs[1]))
            q02 = rel_p2(m0, d2.outs[0].mass, d2.outs[1].mass)
            # d q0^2 / d m0: R is a daughter of vertex 1 and the parent of vertex 2
            r_pos = 0 if d1.outs[0] is ch["R"] else 1
            dq01 = _rel_p2_d_daughter(top.mass, pm(d1.outs[0]), pm(d1.outs[1]), r_pos)
            dq02 = _rel_p2_d_parent(m0, d2.outs[0].mass, d2.outs[1].mass)
            d2q01 = _rel_p2_d2_daughter(top.mass, pm(d1.outs[0]), pm(d1.outs[1]), r_pos)
            d2q02 = _rel_p2_d2_parent(m0, d2.outs[0].mass, d2.outs[1].mass)

            def vertex(ls, q0, dq0, d2q0):
                # sqrt(P_l(z0)), z0 = q0^2 d^2, and the first two m0 derivatives of its log
                c, a, g, h = [], [], [], []
                z0, z1, z2 = q0 * d * d, dq0 * d * d, d2q0 * d * d
                for l in ls:
                    p0 = bprime_poly(l, z0)
                    a.append(int(p0 > 0))
                    c.append(np.sqrt(p0) if p0 > 0 else 1.0)
                    if p0 > 0:
                        p1, p2 = bprime_poly_deriv(l, z0), bprime_poly_deriv2(l, z0)
                        g.append(0.5 * p1 * z1 / p0)
                        h.append(0.5 * ((p2 * z1 * z1 + p1 * z2) / p0 - (p1 * z1 / p0) ** 2))
                    else:
                        g.append(0.0)
                        h.append(0.0)
                return np.array(c), a, np.array(g), np.array(h)

            c1, a1, g1, h1 = vertex(d1.l_list, q01, dq01, d2q01)
            c2, a2, g2, h2 = vertex(d2.l_list, q02, dq02, d2q02)
            out.append(dict(c1=c1, c2=c2, act1=a1, act2=a2, q01=q01, q02=q02, dlog1=g1,
                            dlog2=g2, c=[c1, c2], act=[a1, a2], dlog=[g1, g2],
                            d2log=[h1, h2]))
        return out

    def describe(self, fold=True):
        """The TpaModel structure for the current parameter values. The structure is cached:
        when only masses and widths changed since the last call, their fields are updated."""
        bc = self.barrier_constants()
        flags = tuple(tuple(tuple(a) for a in b["act"]) for b in bc)
        cached = getattr(self, "_desc_cache", None)
        if cached is not None and cached[0] == flags:
            M = cached[1]
            for c, sh in enumerate(self.shape_params()):
                if self.n_final == 4:
                    for r in range(2):
                        M.chain[c].r_m0[r], M.chain[c].r_g0[r] = sh[2 * r], sh[2 * r + 1]
                else:
                    M.chain[c].m0, M.chain[c].g0 = sh
            return M
        M = self._describe4() if self.n_final == 4 else self._describe()
        self._desc_cache = (flags, M)
        return M

    def _describe(self):
        M = _lib.TpaModel()
        g = self.decay_group
        M.n_final = 3
        M.n_topo = len(self.topo)
        M.n_chain = len(self.chains)
        M.P = self.P
        M.lam = self.lam
        M.lam_pad = self.lam_pad
        M.top_j2 = int(round(2 * g.top.J))
        M.n_top = len(g.top.spins)
        for i, s in enumerate(g.top.spins):
            M.hel_top[i] = int(round(2 * s))
        for f, p in enumerate(self.finals):
            M.final_j2[f] = int(round(2 * p.J))
            M.n_fh[f] = len(p.spins)
            for i, s in enumerate(p.spins):
                M.hel_f[f][i] = int(round(2 * s))
            M.ref_topo[f] = self.ref_topo[f]
        M.random_z = int(self.random_z)
        M.n_cg = len(self.cg)
        for t, T in enumerate(self.topo):
            M.topo[t].x, M.topo[t].y, M.topo[t].z = T["x"], T["y"], T["z"]
            M.topo[t].r_first = T["r_first"]
            for f in range(3):
                M.topo[t].aligned[f] = int(self.ref_topo[f] != t)
        bc = self.barrier_constants()
        for c, (ch, (m0, g0), b) in enumerate(zip(self.chains, self.shape_params(), bc)):
            C = M.chain[c]
            R = ch["R"]
            C.topo = ch["topo"]
            C.top_r_first = ch["top_r_first"]
            C.r_y_first = ch["r_y_first"]
            C.jr2 = int(round(2 * R.J))
            C.n_r = len(R.spins)
            for i, s in enumerate(R.spins):
                C.hel_r[i] = int(round(2 * s))
            C.n_ls1, C.n_ls2 = len(ch["d1"].ls), len(ch["d2"].ls)
            for i, l in enumerate(ch["d1"].l_list):
                C.l1[i] = l
                C.act1[i] = b["act1"][i]
            for i, l in enumerate(ch["d2"].l_list):
                C.l2[i] = l
                C.act2[i] = b["act2"][i]
            C.k0 = ch["k0"]
            C.shape = ch["shape"]
            C.bw_l = ch["bw_l"]
            C.cg1_off, C.cg2_off = ch["cg1_off"], ch["cg2_off"]
            C.m0, C.g0 = m0, g0
            C.m_y, C.m_z = ch["d2"].outs[0].mass, ch["d2"].outs[1].mass
            C.d = ch["d1"].d
        return M

    @property
    def shape_mode(self):
        """True when a mass or width is trainable: the cache is then built without the
        lineshapes, which are evaluated per event with their derivatives."""
        key = self.vm.struct_version
        if getattr(self, "_sm_key", None) != key:
            tv = set(self.vm.index_map())
            self._sm = any(nm.get(k) in tv for nm in self._names()
                           for k in ("mass", "width", "mass2", "width2"))
            self._sm_key = key
        return self._sm

    def build_state(self):
        """(cache state, description state), cached by the parameter store's version.
        The cache depends on the barrier clamp flags and, when lineshapes are folded, on the
        lineshape parameters; the device description always carries the latter."""
        # with fixed lineshapes only fixed values matter; trainable lineshapes follow every change
        key = ("v", self.vm.version) if self.shape_mode else ("s", self.vm.struct_version)
        if getattr(self, "_bs_version", None) == key:
            return self._bs
        bc = self.barrier_constants()
        flags = tuple(tuple(tuple(a) for a in b["act"]) for b in bc)
        shapes = tuple(self.shape_params())
        fold = not self.shape_mode
        cache = (fold, flags, shapes if fold else None)
        self._bs = (cache, (flags, shapes))
        self._bs_version = key
        return self._bs

    def device(self):
        """The engine model handle (one per amplitude model); its description is refreshed
        when the lineshape parameters or barrier clamp flags change."""
        cache, desc = self.build_state()
        L = _lib.lib()
        four = self.n_final == 4
        if self._dev is None:
            M = self.describe()
            h = ctypes.c_void_p()
            create = L.tpa_model4_create if four else L.tpa_model_create
            _lib.check(create(ctypes.byref(M), ctypes.sizeof(M), _lib.dptr(self.cg),
                              len(self.cg), ctypes.byref(h)))
            self._dev = h
            self._desc_state = desc
        elif self._desc_state != desc:
            M = self.describe()
            update = L.tpa_model4_update if four else L.tpa_model_update
            _lib.check(update(self._dev, ctypes.byref(M), ctypes.sizeof(M)))
            self._desc_state = desc
        self._dev_state = cache
        return self._dev

    def fused_source(self):
        """The HIP source of this model's fused pass (tpa_fused_source; three-body)."""
        if self.n_final != 3:
            raise NotImplementedError("the fused pass covers three-body models")
        M = self.describe()
        n = ctypes.c_int64(0)
        L = _lib.lib()
        _lib.check(L.tpa_fused_source(ctypes.byref(M), ctypes.sizeof(M), _lib.dptr(self.cg),
                                      len(self.cg), None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        _lib.check(L.tpa_fused_source(ctypes.byref(M), ctypes.sizeof(M), _lib.dptr(self.cg),
                                      len(self.cg), buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def __del__(self):
        h = getattr(self, "_dev", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tpa_model_destroy(h)
            self._dev = None

    # ---------------------------------------------------------------- coefficients
    def _factor_index(self):
        """Unique complex factors (chain totals and the g_ls of every vertex) and, per partial
        wave, the index of each of its factors: slots[0] the chain total, slots[1 + v] the g_ls
        of vertex v (three slots for three-body chains, four for four-body ones)."""
        if getattr(self, "_fidx", None) is None:
            names = self._names()
            cnames, pos = [], {}

            def ix(n):
                if n not in pos:
                    pos[n] = len(cnames)
                    cnames.append(n)
                return pos[n]

            nv = len(self.waves[0].ls)
            slots = [[] for _ in range(1 + nv)]
            ch = []
            for w in self.waves:
                nm = names[w.chain]
                slots[0].append(ix(nm["total"]))
                for v in range(nv):
                    slots[1 + v].append(ix(nm["g"][v][w.ls[v]]))
                ch.append(w.chain)
            slots = [np.array(x) for x in slots]
            ls = np.array([w.ls for w in self.waves]).reshape(len(self.waves), nv)
            self._fidx = dict(cnames=cnames, slots=slots, ch=np.array(ch), ls=ls,
                              it=slots[0], i1=slots[1], i2=slots[2], s1=ls[:, 0], s2=ls[:, 1])
        return self._fidx

    def _factor_values(self):
        """Complex value of every unique factor and its derivatives d/dr, d/di (cached by the
        parameter store's version)."""
        if getattr(self, "_fv_version", None) == self.vm.version:
            return self._fv
        self._fv = self._factor_values_now()
        self._fv_version = self.vm.version
        return self._fv

    def _factor_values_now(self):
        fx = self._factor_index()
        vm = self.vm
        r = np.array([vm.variables[n + "r"] for n in fx["cnames"]])
        i = np.array([vm.variables[n + "i"] for n in fx["cnames"]])
        polar = np.array([vm.complex_vars.get(n, vm.polar) for n in fx["cnames"]])
        e = np.exp(1j * i)
        val = np.where(polar, r * e, r + 1j * i)
        d_r = np.where(polar, e, 1.0 + 0j)
        d_i = np.where(polar, 1j * r * e, 1j)
        return val, d_r, d_i

    def _wave_barrier_index(self):
        """Per partial wave and vertex, the position of its barrier constant in the flat list
        of every chain's vertex constants (chain-major, then vertex, then LS)."""
        if getattr(self, "_wbi", None) is None:
            fx = self._factor_index()
            bc = self.barrier_constants()
            start, pos = {}, 0
            for c, b in enumerate(bc):
                for v, arr in enumerate(b["c"]):
                    start[(c, v)] = pos
                    pos += len(arr)
            self._wbi = np.array([[start[(c, v)] + l for v, l in enumerate(ls)]
                                  for c, ls in zip(fx["ch"], fx["ls"])], dtype=np.int64)
        return self._wbi

    def _barrier_flat(self, field):
        bc = self.barrier_constants()
        return np.concatenate([np.asarray(a, dtype=np.float64) for b in bc for a in b[field]])

    def _barrier_per_wave(self):
        key = tuple(self.shape_params())
        if getattr(self, "_bpw_key", None) != key:
            idx = self._wave_barrier_index()
            self._bpw = np.prod(self._barrier_flat("c")[idx], axis=1)
            self._bpw_key = key
        return self._bpw

    def coefficients(self):
        """c_k for the current parameters (complex array of length P)."""
        fx = self._factor_index()
        val, _, _ = self._factor_values()
        c = self._barrier_per_wave().astype(np.complex128)
        for sl in fx["slots"]:
            c = c * val[sl]
        return c

    @staticmethod
    def _others(val, slots, skip, base):
        out = base
        for b, sl in enumerate(slots):
            if b not in skip:
                out = out * val[sl]
        return out

    def jacobian(self, couplings_only=False):
        """J[i, k] = d c_k / d trainable_vars[i] (complex) for the coupling parameters.
        Raises if a lineshape parameter is trainable (needs the shape-gradient path), unless
        ``couplings_only``: then the rows of masses and widths are left zero."""
        vm = self.vm
        index = vm.index_map()
        for nm in self._names():
            for key in ("mass", "width"):
                if nm[key] in index and not couplings_only:
                    raise NotImplementedError(
                        "gradients with respect to floating masses/widths use the shape path")
        fx = self._factor_index()
        val, d_r, d_i = self._factor_values()
        C = self._barrier_per_wave()
        slots = fx["slots"]
        J = np.zeros((len(vm.trainable_vars), self.P), dtype=np.complex128)
        ks = np.arange(self.P)
        row_r = np.array([index.get(n + "r", -1) for n in fx["cnames"]])
        row_i = np.array([index.get(n + "i", -1) for n in fx["cnames"]])
        for a in range(len(slots)):
            rest = self._others(val, slots, (a,), C.astype(np.complex128))
            f = slots[a]
            for rows, d in ((row_r, d_r), (row_i, d_i)):
                rr = rows[f]
                m = rr >= 0
                np.add.at(J, (rr[m], ks[m]), d[f][m] * rest[m])
        return J

    def _grad_rows(self):
        vm = self.vm
        key = tuple(vm.trainable_vars)
        if getattr(self, "_rows_key", None) != key:
            index = vm.index_map()
            fx = self._factor_index()
            self._rows = (np.array([index.get(n + "r", -1) for n in fx["cnames"]]),
                          np.array([index.get(n + "i", -1) for n in fx["cnames"]]))
            self._rows_key = key
        return self._rows

    def grad_from_adjoint(self, G):
        """2 Re sum_k (d c_k / d theta) G_k for every trainable theta, without forming the
        Jacobian (the chain rule through the three complex factors of each c_k)."""
        row_r, row_i = self._grad_rows()
        fx = self._factor_index()
        val, d_r, d_i = self._factor_values()
        C = self._barrier_per_wave()
        slots = fx["slots"]
        n = len(self.vm.trainable_vars)
        out = np.zeros(n)
        CG = C * np.asarray(G)
        for a in range(len(slots)):
            rest = self._others(val, slots, (a,), CG)
            f = slots[a]
            for rows, d in ((row_r, d_r), (row_i, d_i)):
                rr = rows[f]
                m = rr >= 0
                if m.any():
                    out += np.bincount(rr[m], weights=2.0 * np.real(d[f][m] * rest[m]),
                                       minlength=n)
        return out

    def hess_from_adjoint(self, F1, Fcc, Fcb):
        """Hessian in the trainable couplings of a real function F of the coefficients, from
        its Wirtinger derivatives F1_k = dF/dc_k, Fcc_kl = d2F/dc_k dc_l and
        Fcb_kl = d2F/dc_k dconj(c_l):
        H = 2 Re[J Fcc J^T + J Fcb J^H + sum_k F1_k d2c_k/dtheta dtheta] (J = jacobian())."""
        J = self.jacobian()
        H = 2.0 * np.real(J @ Fcc @ J.T + J @ Fcb @ J.conj().T)
        return H + 2.0 * np.real(self._coupling_second(F1, H.shape[0]))

    def _coupling_second(self, F1, n):
        """sum_k F1_k d2 c_k / dtheta dphi over the coupling parameters (complex [n, n]):
        the second derivatives of each c_k's polar factors and the products of the first
        derivatives of two of its factors."""
        row_r, row_i = self._grad_rows()
        fx = self._factor_index()
        val, d_r, d_i = self._factor_values()
        polar = np.array([self.vm.complex_vars.get(n, self.vm.polar) for n in fx["cnames"]])
        C = self._barrier_per_wave() * np.asarray(F1)
        slots = fx["slots"]
        M = np.zeros((n, n), dtype=np.complex128)
        i_unit = 1j
        for a in range(len(slots)):
            f = slots[a]
            rest = self._others(val, slots, (a,), C.astype(np.complex128))
            # second derivatives of one polar factor: d2/dr dphi = i e^{i phi},
            # d2/dphi2 = -r e^{i phi} (cartesian factors are linear)
            e = d_r[f]
            rr, ri = row_r[f], row_i[f]
            m = polar[f] & (rr >= 0) & (ri >= 0)
            np.add.at(M, (rr[m], ri[m]), i_unit * e[m] * rest[m])
            np.add.at(M, (ri[m], rr[m]), i_unit * e[m] * rest[m])
            m = polar[f] & (ri >= 0)
            np.add.at(M, (ri[m], ri[m]), -val[f][m] * rest[m])
            # products of first derivatives of two different factors
            for b in range(len(slots)):
                if b == a:
                    continue
                g = slots[b]
                rest2 = self._others(val, slots, (a, b), C.astype(np.complex128))
                for rows_a, da in ((row_r, d_r), (row_i, d_i)):
                    for rows_b, db in ((row_r, d_r), (row_i, d_i)):
                        ra, rb = rows_a[f], rows_b[g]
                        m = (ra >= 0) & (rb >= 0)
                        np.add.at(M, (ra[m], rb[m]), da[f][m] * db[g][m] * rest2[m])
        return M

    def _shape_rows(self):
        """Per partial wave: the trainable row of its chain's mass and width (-1 if fixed),
        the m0 derivative of its log barrier constant and the second derivative."""
        fx = self._factor_index()
        index = self.vm.index_map()
        names = self._names()
        mrow = np.array([index.get(nm["mass"], -1) for nm in names])[fx["ch"]]
        wrow = np.array([index.get(nm["width"], -1) for nm in names])[fx["ch"]]
        idx = self._wave_barrier_index()
        dlog = self._barrier_flat("dlog")[idx].sum(axis=1)
        d2log = self._barrier_flat("d2log")[idx].sum(axis=1)
        return mrow, wrow, dlog, d2log

    def hess_shape_from_adjoint(self, F1, Fcc, Fcb):
        """Hessian in every trainable parameter (couplings, masses, widths) of a real function
        S of the amplitude when the lineshapes float, from its Wirtinger derivatives in the
        extended coefficients alpha of A = sum_j alpha_j X_j, X_(t,k) = b^t h_k with the
        lineshape factors b = (B, dB/dm0, dB/dg0, d2B/dm0^2, d2B/dm0dg0, d2B/dg0^2): F1 (6P),
        Fcc and Fcb ([3P, 3P], the first three blocks; the others enter A at second order
        only). At the current point alpha = (c, 0, ...) and
          H = 2 Re[J Fcc J^T + J Fcb J^H + sum_j F1_j d2 alpha_j / dtheta dphi],
        with J the first derivatives: d alpha_(0,k) = d c_k (couplings, and c_k dlog_k for the
        chain's mass through the barrier constants), d alpha_(1,k) / dm0 = d alpha_(2,k) / dg0
        = c_k; second derivatives: d2 c_k in the couplings, dlog_k dc_k / dtheta (mass x
        coupling), c_k (dlog_k^2 + d2log_k) (mass x mass), and for the lineshape blocks the
        products of their factors (tf_pwa/model/model.py:97-150 computes the same with nested
        tapes)."""
        P = self.P
        n = len(self.vm.trainable_vars)
        F1 = np.asarray(F1)
        F0, Fm, Fg, Fmm, Fmg, Fgg = (F1[i * P:(i + 1) * P] for i in range(6))
        c = self.coefficients()
        mrow, wrow, dlog, d2log = self._shape_rows()
        Jc = self.jacobian(couplings_only=True)  # coupling rows of d c_k
        Jf = Jc.copy()                           # every row of d c_k (mass rows added)
        ks = np.arange(P)
        m = mrow >= 0
        np.add.at(Jf, (mrow[m], ks[m]), c[m] * dlog[m])
        J = np.zeros((n, 3 * P), dtype=np.complex128)
        J[:, :P] = Jf
        np.add.at(J, (mrow[m], P + ks[m]), c[m])
        mw = wrow >= 0
        np.add.at(J, (wrow[mw], 2 * P + ks[mw]), c[mw])
        H = 2.0 * np.real(J @ Fcc @ J.T + J @ Fcb @ J.conj().T)
        M = self._coupling_second(F0, n)
        # block 0: the barrier constants' m0 dependence
        for k in np.nonzero(m)[0]:
            r = mrow[k]
            x = F0[k] * dlog[k] * Jc[:, k]
            M[r, :] += x
            M[:, r] += x
            M[r, r] += F0[k] * c[k] * (dlog[k] ** 2 + d2log[k])
        # blocks dB/dm0 and dB/dg0: d2 (c_k delta) = dc_k/dphi [theta = m0] + (theta <-> phi)
        for k in range(P):
            if mrow[k] >= 0:
                x = Fm[k] * Jf[:, k]
                M[mrow[k], :] += x
                M[:, mrow[k]] += x
                M[mrow[k], mrow[k]] += Fmm[k] * c[k]
            if wrow[k] >= 0:
                x = Fg[k] * Jf[:, k]
                M[wrow[k], :] += x
                M[:, wrow[k]] += x
                M[wrow[k], wrow[k]] += Fgg[k] * c[k]
            if mrow[k] >= 0 and wrow[k] >= 0:
                M[mrow[k], wrow[k]] += Fmg[k] * c[k]
                M[wrow[k], mrow[k]] += Fmg[k] * c[k]
        H = H + 2.0 * np.real(M)
        return 0.5 * (H + H.T)

    def shape_grad(self, G, Dm, Dg):
        """Gradient of the lineshape parameters: for the resonance of chain c,
        d/dm0 = 2 Re sum_{k in c} [(d c_k/d m0) G_k + c_k Dm_k] (the barrier constants of both
        vertices depend on m0) and d/dg0 = 2 Re sum_{k in c} c_k Dg_k."""
        vm = self.vm
        n = len(vm.trainable_vars)
        key = (tuple(vm.trainable_vars), tuple(self.shape_params()))
        if getattr(self, "_sg_key", None) != key:
            rkey = tuple(vm.trainable_vars)
            if getattr(self, "_sg_rows_key", None) != rkey:
                index = vm.index_map()
                fx = self._factor_index()
                names = self._names()
                self._sg_rows = (np.array([index.get(nm["mass"], -1) for nm in names])[fx["ch"]],
                                 np.array([index.get(nm["width"], -1) for nm in names])[fx["ch"]])
                self._sg_rows_key = rkey
            dlog = self._barrier_flat("dlog")[self._wave_barrier_index()].sum(axis=1)
            self._sg = (dlog,) + self._sg_rows
            self._sg_key = key
        dlog, mrow, wrow = self._sg
        c = self.coefficients()
        out = np.zeros(n)
        tm = 2.0 * np.real(c * (dlog * G + Dm))
        tw = 2.0 * np.real(c * Dg)
        m = mrow >= 0
        if m.any():
            out += np.bincount(mrow[m], weights=tm[m], minlength=n)
        m = wrow >= 0
        if m.any():
            out += np.bincount(wrow[m], weights=tw[m], minlength=n)
        return out

    # ---------------------------------------------------------------- parameters
    def set_params(self, var):
        if isinstance(var, dict):
            self.vm.set_all(var)
        else:
            self.vm.set_all(var)

    def get_params(self, trainable_only=False):
        return self.vm.get_all_dic(trainable_only)

    @property
    def trainable_variables(self):
        return list(self.vm.trainable_vars)

    def __call__(self, events):
        """Per-event f = sum_lambda |A_lambda|^2 (AbsPDF.__call__, tf_pwa/amp/amp.py:101-110)."""
        return events.amp2(self)
