"""Phase-space generators (tf_pwa/phasespace.py:31-253): three bodies by the m23 weight,
four and more by GENBOD's sorted-random intermediate masses with the product-of-momenta weight
and its maximum (F. James, CERN 68-15).

Host NumPy utility that produces the flat MC samples and the candidate pools of toy
generation; the likelihood never runs here.
"""

import numpy as np


def _q(M, ma, mb):
    p2 = (M * M - (ma + mb) ** 2) * (M * M - (ma - mb) ** 2)
    return np.sqrt(np.where(p2 > 0, p2, 0.0)) / (2.0 * M)


def _boost(p, b):
    """Boost four-vectors p (n,4) by velocity b (n,3)."""
    b2 = np.sum(b * b, axis=-1)
    g = 1.0 / np.sqrt(1.0 - b2)
    bp = np.sum(b * p[:, 1:], axis=-1)
    g2 = np.where(b2 > 1e-14, (g - 1.0) / np.where(b2 > 0, b2, 1.0), 0.0)
    out = np.empty_like(p)
    out[:, 1:] = p[:, 1:] + (g2 * bp)[:, None] * b + (g * p[:, 0])[:, None] * b
    out[:, 0] = g * (p[:, 0] + bp)
    return out


def _iso(M, ma, mb, rng):
    """Back-to-back pair of masses ma, mb in the rest frame of M, isotropic direction."""
    n = M.shape[0]
    c = 2.0 * rng.uniform(0.0, 1.0, n) - 1.0
    s = np.sqrt(1.0 - c * c)
    phi = 2.0 * np.pi * rng.uniform(0.0, 1.0, n)
    q = _q(M, ma, mb)
    pa = np.stack([np.sqrt(q * q + ma * ma), q * s * np.cos(phi), q * s * np.sin(phi), q * c], -1)
    pb = np.stack([np.sqrt(q * q + mb * mb), -pa[:, 1], -pa[:, 2], -pa[:, 3]], -1)
    return pa, pb


def _genbod(m0, masses, n, rng):
    N = len(masses)
    tecm = m0 - sum(masses)
    emmax, emmin, wtmax = tecm + masses[0], 0.0, 1.0
    for i in range(1, N):
        emmin += masses[i - 1]
        emmax += masses[i]
        wtmax *= float(_q(np.array(emmax), np.array(emmin), masses[i]))
    csum = np.cumsum(masses)
    keep, got = [], 0
    while got < n:
        k = int((n - got) * 3) + 1024
        r = np.sort(rng.uniform(0.0, 1.0, (k, N - 2)), axis=1)
        inv = np.empty((k, N))
        inv[:, 0] = masses[0]
        inv[:, 1:N - 1] = r * tecm + csum[1:N - 1]
        inv[:, N - 1] = m0
        w = np.ones(k)
        for i in range(1, N):
            w = w * _q(inv[:, i], inv[:, i - 1], masses[i])
        inv = inv[w > rng.uniform(0.0, 1.0, k) * wtmax]
        keep.append(inv)
        got += inv.shape[0]
    inv = np.concatenate(keep)[:n]
    # particles 0 and 1 in the rest frame of inv[1]; then each next particle recoils against
    # the system of the previous ones in the rest frame of inv[i], which is boosted along
    p = list(_iso(inv[:, 1], masses[0], masses[1], rng))
    for i in range(2, N):
        sys_p, pi = _iso(inv[:, i], inv[:, i - 1], masses[i], rng)
        b = sys_p[:, 1:] / sys_p[:, :1]
        p = [_boost(x, b) for x in p] + [pi]
    return p


def generate_phsp(m0, masses, n, rng):
    """n flat phase-space events of m0 -> masses in the m0 rest frame; returns one (n, 4)
    array per final in the order of ``masses``."""
    if len(masses) > 3:
        return _genbod(float(m0), [float(x) for x in masses], n, rng)
    m1, m2, m3 = masses
    lo, hi = m2 + m3, m0 - m1
    grid = np.linspace(lo, hi, 20001)
    wmax = (_q(m0, m1, grid) * _q(grid, m2, m3)).max() * 1.001
    keep, got = [], 0
    while got < n:
        k = int((n - got) * 2.5) + 1024
        m23 = rng.uniform(lo, hi, k)
        w = _q(m0, m1, m23) * _q(m23, m2, m3) / wmax
        m23 = m23[w > rng.uniform(0.0, 1.0, k)]
        keep.append(m23)
        got += m23.size
    m23 = np.concatenate(keep)[:n]

    def iso(M, ma, mb):
        c = 2.0 * rng.uniform(0.0, 1.0, n) - 1.0
        s = np.sqrt(1.0 - c * c)
        phi = 2.0 * np.pi * rng.uniform(0.0, 1.0, n)
        q = _q(M, ma, mb)
        pa = np.stack([np.sqrt(q * q + ma * ma), q * s * np.cos(phi), q * s * np.sin(phi), q * c], -1)
        pb = np.stack([np.sqrt(q * q + mb * mb), -pa[:, 1], -pa[:, 2], -pa[:, 3]], -1)
        return pa, pb

    p2, p3 = iso(m23, m2, m3)
    p1, p23 = iso(np.full(n, float(m0)), m1, m23)
    b = p23[:, 1:] / p23[:, :1]
    return [p1, _boost(p2, b), _boost(p3, b)]
