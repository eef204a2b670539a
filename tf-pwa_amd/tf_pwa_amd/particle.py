"""Decay topology of a config.yml model: particles, two-body decays, chains and groups.

Mirrors the structure of tf_pwa/particle.py (BaseParticle, Decay, DecayChain, DecayGroup)
and tf_pwa/config_loader/decay_config.py: chains are enumerated from the top particle in
config order (particle.py:171-181), finals and inner particles are kept sorted by name
(particle.py:528-534), and LS couplings follow GetA2BC_LS_list (particle.py:207-248).
"""

from fractions import Fraction

import numpy as np

from .cg import cg_coef


def spin_value(x):
    """Spin given as number or string such as '3/2'."""
    if isinstance(x, str):
        return float(Fraction(x))
    return float(x)


def spin_range(j):
    """-j, -j+1, ..., j"""
    n = int(round(2 * j)) + 1
    return tuple(-j + i for i in range(n))


def ls_couplings(ja, jb, jc, pa=None, pb=None, pc=None, p_break=False, ca=None):
    """(l, s) couplings of A -> B C: |jb-jc| <= s <= jb+jc, |ja-s| <= l <= ja+s with integer
    l, parity P_A = P_B P_C (-1)^l unless p_break, optional C-parity (-1)^(l+s)
    (tf_pwa/particle.py:207-248)."""
    if pa is None or pb is None or pc is None:
        p_break = True
    ret = []
    n_s = int(round(2 * min(jb, jc))) + 1
    for i in range(n_s):
        s = abs(jb - jc) + i
        n_l = int(round(2 * min(ja, s))) + 1
        for k in range(n_l):
            l2 = round(2 * (abs(ja - s) + k))
            if l2 % 2:
                break
            l = l2 // 2
            if ca is not None and ca != (-1) ** int(round(l + s)):
                continue
            if not p_break and (-1) ** l * pb * pc != pa:
                continue
            ret.append((l, s))
    return ret


class Particle:
    """A particle of the config (name, J, P, helicities, mass, width, lineshape)."""

    def __init__(self, name, J=0, P=-1, C=None, spins=None, mass=None, width=None,
                 model="default", **kwargs):
        self.name = str(name)
        self.J = spin_value(J)
        self.P = P
        self.C = C
        if spins is None:
            spins = spin_range(self.J)
        self.spins = tuple(spin_value(s) for s in spins)
        self.mass = None if mass is None else float(mass)
        self.width = None if width is None else float(width)
        self.model = model
        self.kwargs = kwargs
        self.decays = []

    def __repr__(self):
        return self.name

    def __lt__(self, other):
        return self.name < other.name


class Decay:
    """A two-body vertex core -> outs[0] + outs[1] (tf_pwa/amp/core.py:650-1233)."""

    def __init__(self, core, outs, p_break=False, c_break=True, model="default", l_list=None,
                 ls_list=None, **kwargs):
        if len(outs) != 2:
            raise NotImplementedError("only two-body vertices are supported")
        self.core = core
        self.outs = tuple(outs)
        self.p_break = bool(p_break)
        self.c_break = c_break
        self.model = model
        self.kwargs = kwargs
        self.d = 3.0
        ca = core.C if not c_break else None
        # ls_list replaces the allowed couplings, l_list keeps those with the listed l
        # (HelicityDecay.get_ls_list, tf_pwa/amp/core.py:1212-1233)
        if ls_list is not None:
            self.ls = [(int(l), spin_value(sv)) for l, sv in ls_list]
        else:
            self.ls = ls_couplings(core.J, outs[0].J, outs[1].J, core.P, outs[0].P, outs[1].P,
                                   p_break=self.p_break, ca=ca)
            if l_list is not None:
                self.ls = [(l, sv) for l, sv in self.ls if l in set(int(x) for x in l_list)]
        if not self.ls:
            raise ValueError("invalid spin parity for {}{}".format(
                self, "" if self.p_break else ", maybe you should set `p_break: True`"))

    def __repr__(self):
        return "{}->{}".format(self.core, "+".join(str(o) for o in self.outs))

    @property
    def l_list(self):
        return tuple(l for l, s in self.ls)

    def cg_matrix(self):
        """[ls][lambda_b][lambda_c] = sqrt((2l+1)/(2J_a+1)) <J_b l_b; J_c -l_c | s d>
        <l 0; s d | J_a d> with d = l_b - l_c (tf_pwa/amp/core.py:851-900)."""
        a, b, c = self.core, self.outs[0], self.outs[1]
        ret = np.zeros((len(self.ls), len(b.spins), len(c.spins)))
        for i, (l, s) in enumerate(self.ls):
            for i1, lb in enumerate(b.spins):
                for i2, lc in enumerate(c.spins):
                    ret[i, i1, i2] = (np.sqrt(2 * l + 1) / np.sqrt(2 * a.J + 1)
                                      * cg_coef(b.J, c.J, lb, -lc, s, lb - lc)
                                      * cg_coef(l, s, 0, lb - lc, a.J, lb - lc))
        return ret

    def params_head(self):
        return _var_name("{}->{}".format(self.core, "+".join(str(o) for o in self.outs)))


def _var_name(s):
    """Variable-name sanitiser of tf_pwa/amp/core.py:235-246."""
    return (s.replace(":", "/").replace("+", ".").replace(",", "").replace("[", "")
            .replace("]", "").replace(" ", ""))


class DecayChain:
    """An ordered list of decays from the top particle to the finals."""

    def __init__(self, decays):
        self.decays = list(decays)
        cores = [d.core for d in self.decays]
        outs = [o for d in self.decays for o in d.outs]
        tops = [p for p in cores if p not in outs]
        assert len(tops) == 1, "a chain needs exactly one top particle"
        self.top = tops[0]
        self.inner = sorted([p for p in cores if p in outs])
        self.outs = sorted([p for p in outs if p not in cores])

    def __iter__(self):
        return iter(self.decays)

    def __len__(self):
        return len(self.decays)

    def __repr__(self):
        return "[" + ", ".join(str(d) for d in self.decays) + "]"

    def params_head(self):
        return _var_name(str(self))

    def finals_below(self):
        """particle -> frozenset of final names below it (sorted_table, particle.py:545-569)."""
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
        return tuple(sorted(tuple(sorted(v)) for v in self.finals_below().values()))

    def decay_of(self, particle):
        for d in self.decays:
            if d.core is particle:
                return d
        raise KeyError(particle)


class DecayGroup:
    """Chains with the same top and finals, summed coherently (tf_pwa/amp/core.py:1637)."""

    def __init__(self, chains):
        self.chains = list(chains)
        self.top = self.chains[0].top
        self.outs = sorted(self.chains[0].outs)
        for c in self.chains:
            assert c.top is self.top and [p.name for p in c.outs] == [p.name for p in self.outs]
        res = []
        for c in self.chains:
            for r in c.inner:
                if r not in res:
                    res.append(r)
        self.resonances = res

    def __iter__(self):
        return iter(self.chains)

    def __len__(self):
        return len(self.chains)

    def __getitem__(self, i):
        return self.chains[i]

    def topologies(self):
        """Unique topologies in chain order, each represented by its first chain
        (DecayGroup.topology_structure, particle.py:902-920)."""
        reps = []
        for c in self.chains:
            if all(r.topology_key() != c.topology_key() for r in reps):
                reps.append(c)
        return reps

    def get_particle(self, name):
        for c in self.chains:
            for d in c:
                for p in (d.core,) + d.outs:
                    if p.name == name:
                        return p
        raise ValueError("no particle {}".format(name))


def chain_decays(particle):
    """All decay chains below ``particle`` in config order (particle.py:171-181)."""
    ret = []
    for d in particle.decays:
        combos = [[d]]
        for o in d.outs:
            sub = chain_decays(o)
            if sub:
                combos = [a + b for a in combos for b in sub]
        ret += combos
    return ret
