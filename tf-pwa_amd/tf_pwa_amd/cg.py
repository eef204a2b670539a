"""Clebsch-Gordan coefficients <j1 m1; j2 m2 | J M> from Racah's closed formula.

The reference obtains them from sympy or a JSON table (tf_pwa/cg.py:18-64); here the host
evaluates Racah's sum exactly in rational arithmetic and takes one square root at the end.
"""

import functools
import math
from fractions import Fraction


def _twice(x):
    t = round(2 * float(x))
    if abs(2 * float(x) - t) > 1e-9:
        raise ValueError("spin {} is not a multiple of 1/2".format(x))
    return int(t)


@functools.lru_cache(maxsize=None)
def _cg2(j1, m1, j2, m2, J, M):
    """All arguments doubled (2j, 2m)."""
    if m1 + m2 != M:
        return 0.0
    for j, m in ((j1, m1), (j2, m2), (J, M)):
        if abs(m) > j or (j + m) % 2:
            return 0.0
    if J < abs(j1 - j2) or J > j1 + j2 or (j1 + j2 + J) % 2:
        return 0.0
    f = lambda x2: math.factorial(x2 // 2)
    pre = Fraction((J + 1) * f(J + j1 - j2) * f(J - j1 + j2) * f(j1 + j2 - J), f(j1 + j2 + J + 2))
    pre *= f(J + M) * f(J - M) * f(j1 - m1) * f(j1 + m1) * f(j2 - m2) * f(j2 + m2)
    s = Fraction(0)
    for k in range(0, j1 + j2 + 1, 2):
        args = (j1 + j2 - J - k, j1 - m1 - k, j2 + m2 - k, J - j2 + m1 + k, J - j1 - m2 + k)
        if min(args) < 0:
            continue
        term = Fraction(1, f(k) * f(args[0]) * f(args[1]) * f(args[2]) * f(args[3]) * f(args[4]))
        s += -term if (k // 2) % 2 else term
    if s == 0:
        return 0.0
    val = math.sqrt(pre * s * s)
    return val if s > 0 else -val


def cg_coef(jb, jc, mb, mc, ja, ma):
    """<jb mb; jc mc | ja ma>, argument order of tf_pwa/cg.py:18-26."""
    return _cg2(_twice(jb), _twice(mb), _twice(jc), _twice(mc), _twice(ja), _twice(ma))
