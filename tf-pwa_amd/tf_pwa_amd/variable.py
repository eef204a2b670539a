"""Parameter store with tf-pwa's names, ordering and complex/bound conventions.

Follows tf_pwa/variable.py: every parameter is a named real; complex couplings are a pair
``<name>r``/``<name>i`` in polar form (r, phi) by default (variable.py:171-197, 1715-1765);
``trainable_vars`` is the order of the gradient vector (638-646); bounds transform the fit
variable x to the physical y with dy/dx applied to the gradient (797-821, 1029-1135).
"""

import numpy as np


class Bound:
    """Bound transform y(x) of tf_pwa/variable.py:1029-1135 ("asin" style)."""

    def __init__(self, a=None, b=None):
        if a is not None and b is not None and a > b:
            raise ValueError("bound lower {} above upper {}".format(a, b))
        self.lower = a
        self.upper = b

    def get_x2y(self, x):
        a, b = self.lower, self.upper
        if a is not None and b is not None:
            return (b - a) * (np.sin(x) + 1) / 2 + a
        if a is not None:
            return a + np.sqrt(x * x + 1) - 1
        if b is not None:
            return b + 1 - np.sqrt(x * x + 1)
        return x

    def get_y2x(self, y):
        """Inverse transform; y is clamped into the bound first, as the reference does, so a
        start value outside a one-sided range maps to its edge instead of NaN."""
        a, b = self.lower, self.upper
        if a is not None and b is not None:
            return np.arcsin(np.clip(2 * (y - a) / (b - a) - 1, -1, 1))
        if a is not None:
            y = max(y, a)
            return np.sqrt((y - a + 1) ** 2 - 1)
        if b is not None:
            y = min(y, b)
            return np.sqrt((b - y + 1) ** 2 - 1)
        return y

    def get_dydx(self, x):
        a, b = self.lower, self.upper
        if a is not None and b is not None:
            return (b - a) * np.cos(x) / 2
        if a is not None:
            return x / np.sqrt(x * x + 1)
        if b is not None:
            return -x / np.sqrt(x * x + 1)
        return 1.0


class VarsManager:
    """Named real parameters, the trainable subset and complex pairs."""

    def __init__(self, polar=True):
        self.polar = polar
        self.variables = {}
        self.trainable_vars = []
        self.complex_vars = {}
        self.bnd_dic = {}
        self.init_val = {}
        # shared variables (set_same): every alias name reads and writes its canonical name's
        # value, and only the canonical one can be trainable
        self.same = {}
        # bumped on every change of a value or of the trainable set, so derived quantities
        # (coefficients, lineshape state) can be cached by version
        self.version = 0
        # bumped on every change except new values of the trainable variables in order
        # (set_all with a sequence): what only depends on fixed values and on the trainable
        # set is cached by this one
        self.struct_version = 0

    # -------------------------------------------------------------- creation
    def add_real_var(self, name, value=None, range_=None, trainable=True, rng=None):
        if name in self.trainable_vars:
            self.trainable_vars.remove(name)
        if value is None:
            rng = rng or np.random
            lo, hi = (0.0, 1.0) if range_ is None else range_
            value = rng.uniform(lo, hi)
        self.variables[name] = float(value)
        self.init_val[name] = float(value)
        if trainable:
            self.trainable_vars.append(name)
        self.version += 1
        self.struct_version += 1

    def add_complex_var(self, name, polar=None, trainable=True, fix_vals=(1.0, 0.0), rng=None):
        polar = self.polar if polar is None else polar
        if trainable:
            if polar:
                self.add_real_var(name + "r", range_=(0.0, 2.0), rng=rng)
                self.add_real_var(name + "i", range_=(-np.pi, np.pi), rng=rng)
            else:
                self.add_real_var(name + "r", range_=(-1.0, 1.0), rng=rng)
                self.add_real_var(name + "i", range_=(-1.0, 1.0), rng=rng)
        else:
            self.add_real_var(name + "r", value=fix_vals[0], trainable=False)
            self.add_real_var(name + "i", value=fix_vals[1], trainable=False)
        self.complex_vars[name] = polar

    def set_fix(self, name, value=None, unfix=False):
        """variable.py:415-440: fixing removes from trainable_vars, freeing appends."""
        self.version += 1
        self.struct_version += 1
        name = self.same.get(name, name)
        if value is not None:
            self.variables[name] = float(value)
            self._mirror()
        if unfix:
            if name not in self.trainable_vars:
                self.trainable_vars.append(name)
        elif name in self.trainable_vars:
            self.trainable_vars.remove(name)

    def set_same(self, names):
        """Make the listed real variables one (VarsManager.set_same, variable.py:500-546):
        the first name is canonical and takes its current value; the others leave
        trainable_vars, and if any of them is fixed the canonical one is fixed too. A name
        already shared joins with its whole group."""
        names = [n for n in names if n in self.variables]
        if len(names) < 2:
            return
        group = []
        for n in names:
            c = self.same.get(n, n)
            for m in [c] + [a for a, cc in self.same.items() if cc == c]:
                if m not in group:
                    group.append(m)
        canon = self.same.get(names[0], names[0])
        group.remove(canon)
        group.insert(0, canon)
        fixed = any(n not in self.trainable_vars for n in group)
        for n in group[1:]:
            self.same[n] = canon
            if n in self.trainable_vars:
                self.trainable_vars.remove(n)
        if fixed and canon in self.trainable_vars:
            self.trainable_vars.remove(canon)
        self._mirror()
        self.version += 1
        self.struct_version += 1

    def _mirror(self):
        for a, c in self.same.items():
            self.variables[a] = self.variables[c]

    def index_map(self):
        """Name -> position in trainable_vars for every name that moves with a trainable
        variable, shared aliases included (their derivatives add into the canonical row)."""
        index = {n: i for i, n in enumerate(self.trainable_vars)}
        for a, c in self.same.items():
            if c in index:
                index[a] = index[c]
        return index

    def set_bound(self, bound_dic):
        self.struct_version += 1
        for name, (a, b) in bound_dic.items():
            self.bnd_dic[name] = Bound(a, b)

    # -------------------------------------------------------------- access
    def get(self, name):
        return self.variables[name]

    def set(self, name, value):
        name = self.same.get(name, name)
        self.variables[name] = float(value)
        if self.same:
            self._mirror()
        self.version += 1
        if name not in self.trainable_vars:
            self.struct_version += 1

    def get_all_val(self):
        return [self.variables[n] for n in self.trainable_vars]

    def get_all_dic(self, trainable_only=False):
        if trainable_only:
            return {n: self.variables[n] for n in self.trainable_vars}
        return dict(self.variables)

    def set_all(self, vals):
        """dict: by name; sequence: in trainable_vars order (variable.py:684-697)."""
        self.version += 1
        if isinstance(vals, dict):
            self.struct_version += 1
            for k, v in vals.items():
                if k in self.variables:
                    self.variables[self.same.get(k, k)] = float(v)
        else:
            # ndarray.tolist gives Python floats in one call (float() per numpy scalar costs
            # ~0.2 us each, on every likelihood call)
            vals = vals.tolist() if isinstance(vals, np.ndarray) else [float(v) for v in vals]
            if len(vals) != len(self.trainable_vars):
                raise ValueError("expected {} values, got {}".format(
                    len(self.trainable_vars), len(vals)))
            self.variables.update(zip(self.trainable_vars, vals))
        if self.same:
            self._mirror()

    def complex_value(self, name):
        """Variable.__call__ (variable.py:1715-1765)."""
        r, i = self.variables[name + "r"], self.variables[name + "i"]
        if self.complex_vars.get(name, self.polar):
            return complex(r * np.cos(i), r * np.sin(i))
        return complex(r, i)

    def complex_derivs(self, name):
        """d value / d (name+'r'), d value / d (name+'i')."""
        r, i = self.variables[name + "r"], self.variables[name + "i"]
        if self.complex_vars.get(name, self.polar):
            e = complex(np.cos(i), np.sin(i))
            return e, 1j * r * e
        return 1.0 + 0j, 1j

    # -------------------------------------------------------------- bounds
    def trans_fcn_grad(self, fcn_grad):
        """F(x) = F(y(x)), dF/dx = dF/dy dy/dx (variable.py:797-821)."""

        def fcn_t(xvals):
            xvals = np.asarray(xvals, dtype=np.float64)
            yvals = xvals.copy()
            dydx = np.ones_like(xvals)
            for i, name in enumerate(self.trainable_vars):
                if name in self.bnd_dic:
                    yvals[i] = self.bnd_dic[name].get_x2y(xvals[i])
                    dydx[i] = self.bnd_dic[name].get_dydx(xvals[i])
            f, g = fcn_grad(yvals)
            return f, np.asarray(g) * dydx

        return fcn_t

    def get_x_from_y(self):
        x = []
        for name in self.trainable_vars:
            y = self.variables[name]
            x.append(self.bnd_dic[name].get_y2x(y) if name in self.bnd_dic else y)
        return np.array(x)

    def std_polar_all(self):
        """r >= 0 and phi in [-pi, pi) for every polar pair (variable.py:716-741)."""
        self.version += 1
        self.struct_version += 1
        for name, polar in self.complex_vars.items():
            if not polar:
                continue
            r, p = self.variables[name + "r"], self.variables[name + "i"]
            if r < 0:
                r, p = -r, p + np.pi
            self.variables[name + "r"] = r
            self.variables[name + "i"] = (p + np.pi) % (2 * np.pi) - np.pi
