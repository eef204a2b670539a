"""The host chain rule of the exact Hessian with floating masses and widths
(AmplitudeModel.hess_shape_from_adjoint) on the CPU: a numpy stand-in for the device pass
builds the moments of tpa_hess_shape from random partial waves and an analytic lineshape per
chain (BW: B = 1/(m0^2 - m^2 - i m0 g0), derivatives in closed form), and the Hessian of the
same sum as a function of every trainable parameter (couplings through c_k, masses through
c_k's barrier constants and B, widths through B) is compared with central second differences
of the sum itself. The GPU test holds the device moments to the oracle's autograd Hessian."""

import numpy as np
import pytest
import yaml

from conftest import config_path


def _bw(m, m0, g0):
    """B and its first and second derivatives in (m0, g0) per event."""
    D = m0 * m0 - m * m - 1j * m0 * g0
    B = 1.0 / D
    Dm, Dg = 2 * m0 - 1j * g0, -1j * m0
    Dmm, Dmg, Dgg = 2.0, -1j, 0.0

    def d1(Dx):
        return -B * B * Dx

    def d2(Dx, Dy, Dxy):
        return 2 * B ** 3 * Dx * Dy - B * B * Dxy

    return [B, d1(Dm), d1(Dg), d2(Dm, Dm, Dmm), d2(Dm, Dg, Dmg), d2(Dg, Dg, Dgg)]


class _Stand:
    """Random waves h[k, e, lam] of a model, per-event pair masses per chain, weights."""

    def __init__(self, amp, n, seed, mode):
        rng = np.random.default_rng(seed)
        self.amp = amp
        self.mode = mode
        self.h = rng.normal(size=(amp.P, n, amp.lam)) + 1j * rng.normal(size=(amp.P, n, amp.lam))
        self.w = rng.uniform(0.5, 1.5, n)
        shapes = amp.shape_params()
        self.m = np.array([sh[0] + rng.uniform(-1.5, 1.5, n) * sh[1] for sh in shapes])
        self.kchain = np.array([w.chain for w in amp.waves])

    def factors(self):
        shapes = self.amp.shape_params()
        b = [_bw(self.m[c], m0, g0) for c, (m0, g0) in enumerate(shapes)]
        return [np.array([b[c][t] for c in self.kchain]) for t in range(6)]  # [t][k, e]

    def value(self):
        c = self.amp.coefficients()
        B = self.factors()[0]
        A = np.einsum("k,ke,kel->el", c, B, self.h)
        f = np.sum(np.abs(A) ** 2, axis=1)
        return np.sum(self.w * (np.log(f) if self.mode == 0 else f))

    def moments(self):
        c = self.amp.coefficients()
        b = self.factors()
        P = self.amp.P
        X = np.concatenate([b[t][:, :, None] * self.h for t in range(3)])  # [3P, e, lam]
        A = np.einsum("k,kel->el", c, X[:P])
        f = np.sum(np.abs(A) ** 2, axis=1)
        u = self.w / f if self.mode == 0 else self.w
        v = -self.w / f ** 2 if self.mode == 0 else 0.0 * self.w
        z = np.einsum("el,jel->je", np.conj(A), X)
        q = np.einsum("el,kel->ke", np.conj(A), self.h)
        F1 = np.concatenate([z @ u] + [(b[3 + t] * q) @ u for t in range(3)])
        Fcc = (z * v) @ z.T
        Fcb = (z * v) @ np.conj(z).T + np.einsum("e,iel,jel->ij", u, X, np.conj(X))
        return F1, Fcc, Fcb


@pytest.mark.parametrize("mode", [0, 1])
def test_shape_hessian_chain_rule_matches_second_differences(mode):
    from tf_pwa_amd import ConfigLoader

    cfg = yaml.safe_load(open(config_path("lc_pkpi_2res.yml")))
    for r in ("Kst", "L1520"):
        cfg["particle"][r]["float"] = "mg"
        cfg["particle"][r]["model"] = "BW"
    c = ConfigLoader(cfg, seed=3)
    amp = c.get_amplitude()
    assert amp.shape_mode
    st = _Stand(amp, 40, 5, mode)
    vm = c.vm
    x0 = np.array(vm.get_all_val())
    F1, Fcc, Fcb = st.moments()
    H = amp.hess_shape_from_adjoint(F1, Fcc, Fcb)
    g = amp.grad_from_adjoint(F1[:amp.P]) + amp.shape_grad(
        F1[:amp.P], F1[amp.P:2 * amp.P], F1[2 * amp.P:3 * amp.P])
    n = len(x0)
    Hn = np.zeros((n, n))
    gn = np.zeros(n)

    def S(dx):
        vm.set_all(x0 + dx)
        return st.value()

    # central differences with one Richardson step (error O(h^4)): the narrow L(1520) makes
    # plain differences of the mass and width rows too coarse at any step that keeps roundoff
    # small
    for i in range(n):
        ei = np.eye(n)[i] * 1e-4 * max(1.0, abs(x0[i]))
        d = [(S(ei / s) - S(-ei / s)) / (2 * ei[i] / s) for s in (1, 2)]
        gn[i] = (4 * d[1] - d[0]) / 3
        for j in range(i, n):
            ej = np.eye(n)[j] * 1e-4 * max(1.0, abs(x0[j]))
            d = [(S((ei + ej) / s) - S((ei - ej) / s) - S((ej - ei) / s) + S(-(ei + ej) / s))
                 / (4 * ei[i] * ej[j] / s ** 2) for s in (1, 2)]
            Hn[i, j] = Hn[j, i] = (4 * d[1] - d[0]) / 3
    vm.set_all(x0)
    np.testing.assert_allclose(g, gn, rtol=1e-6, atol=1e-9 * np.abs(gn).max())
    scale = np.abs(Hn).max()
    assert np.abs(H - Hn).max() <= 1e-6 * scale, (np.abs(H - Hn).max() / scale)
    # the mass rows see the barrier constants (P_l(z0) with q0 from m0) and the lineshape
    names = vm.trainable_vars
    for r in ("Kst_mass", "L1520_mass", "Kst_width"):
        i = names.index(r)
        assert np.abs(H[i]).max() > 1e-6 * scale
