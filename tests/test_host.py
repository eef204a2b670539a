"""Host logic of the product (no GPU): CG and LS tables, the parameter layout, bounds, the
partial-wave coefficient Jacobian, the C ABI library's exported symbols and the phase-space
generator."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, config_path
from oracle import tfpwa_oracle as O


def test_racah_cg_matches_sympy():
    from tf_pwa_amd.cg import cg_coef

    spins = [0, 0.5, 1, 1.5, 2, 2.5]
    rng = np.random.default_rng(0)
    n = 0
    for j1 in spins:
        for j2 in spins:
            for J in np.arange(abs(j1 - j2), j1 + j2 + 0.5, 1.0):
                for m1 in np.arange(-j1, j1 + 0.5, 1.0):
                    for m2 in np.arange(-j2, j2 + 0.5, 1.0):
                        if abs(m1 + m2) > J or rng.uniform() > 0.3:
                            continue
                        a = cg_coef(j1, j2, m1, m2, J, m1 + m2)
                        b = O.cg_coef(j1, j2, m1, m2, J, m1 + m2)
                        assert abs(a - b) < 1e-14, (j1, j2, m1, m2, J, a, b)
                        n += 1
    assert n > 200
    assert cg_coef(1, 1, 1, 1, 1, 1) == 0.0  # m1 + m2 != M is not passed; M = 2 > J
    assert cg_coef(0.5, 0.5, 0.5, 0.5, 0, 1) == 0.0


def test_ls_lists_match_oracle():
    from tf_pwa_amd.particle import ls_couplings

    cases = [(0.5, 1, 0.5, 1, -1, 1), (0.5, 1.5, 0, 1, -1, -1), (1.5, 0.5, 0, -1, 1, -1),
             (1, 1, 1, -1, -1, -1), (2.5, 0.5, 0, 1, 1, -1), (0, 0, 0, 1, -1, -1)]
    for ja, jb, jc, pa, pb, pc in cases:
        for pbrk in (False, True):
            a = ls_couplings(ja, jb, jc, pa, pb, pc, p_break=pbrk)
            b = O.ls_list(ja, jb, jc, pa, pb, pc, p_break=pbrk)
            assert [(l, float(s)) for l, s in a] == [(l, float(s)) for l, s in b]
    assert len(ls_couplings(1, 1, 1, -1, -1, -1)) == 4


@pytest.mark.parametrize("cfg", ["lc_pkpi_2res.yml", "lc_pkpi_6res.yml"])
def test_parameter_layout_matches_oracle(cfg):
    from tf_pwa_amd import ConfigLoader

    c = ConfigLoader(config_path(cfg), seed=0)
    m = O.Model(config_path(cfg))
    assert c.vm.trainable_vars == m.trainable
    amp = c.get_amplitude()
    assert amp.P == {"lc_pkpi_2res.yml": 8, "lc_pkpi_6res.yml": 22}[cfg]
    for ch in c.decay_group:
        for d in ch:
            od = [x for x in m.decays if str(x) == str(d)][0]
            np.testing.assert_allclose(d.cg_matrix(), od.cg_matrix, atol=1e-15)


def test_floating_shapes_append_in_constraint_order():
    import yaml

    from tf_pwa_amd import ConfigLoader

    cfg = yaml.safe_load(open(config_path("lc_pkpi_2res.yml")))
    cfg["particle"]["Kst"]["float"] = "mg"
    cfg["particle"]["L1520"]["float"] = "m"
    c = ConfigLoader(cfg, seed=0)
    assert c.vm.trainable_vars[-3:] == ["Kst_mass", "Kst_width", "L1520_mass"]


def test_bound_transform():
    from tf_pwa_amd.variable import Bound

    for a, b in ((0.5, 1.5), (0.2, None), (None, 3.0)):
        bd = Bound(a, b)
        for x in (-0.7, 0.1, 0.9):
            y = bd.get_x2y(x)
            if a is not None:
                assert y >= a - 1e-12
            if b is not None:
                assert y <= b + 1e-12
            h = 1e-6
            fd = (bd.get_x2y(x + h) - bd.get_x2y(x - h)) / (2 * h)
            assert abs(fd - bd.get_dydx(x)) < 1e-8
            assert abs(bd.get_x2y(bd.get_y2x(y)) - y) < 1e-12


def test_coefficient_jacobian_finite_difference():
    from tf_pwa_amd import ConfigLoader

    c = ConfigLoader(config_path("lc_pkpi_6res.yml"), seed=4)
    amp = c.get_amplitude()
    x0 = np.array(c.vm.get_all_val())
    J = amp.jacobian()
    for i in range(len(x0)):
        h = 1e-6
        xp, xm = x0.copy(), x0.copy()
        xp[i] += h
        xm[i] -= h
        c.vm.set_all(xp)
        cp = amp.coefficients()
        c.vm.set_all(xm)
        cm = amp.coefficients()
        np.testing.assert_allclose((cp - cm) / (2 * h), J[i], atol=1e-7)
    c.vm.set_all(x0)


def test_library_exports_header_symbols():
    from tf_pwa_amd import _lib

    header = open(os.path.join(ROOT, "include", "tfpwa_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(tpa_\w+)\(", header, re.M))
    assert declared == set(_lib.exported_symbols())
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert _lib.lib().tpa_abi_version() == 1
    # the description layout seen by ctypes is the one the library validates
    rc = _lib.lib().tpa_model_create(ctypes.byref(_lib.TpaModel()), 1, None, 0,
                                     ctypes.byref(ctypes.c_void_p()))
    assert rc != 0 and b"size mismatch" in _lib.lib().tpa_last_error()


def test_header_flags_match_the_binding():
    """Every TPA_FLAG_* and TPA_MODE_* value of the header is the binding's."""
    from tf_pwa_amd import _lib

    header = open(os.path.join(ROOT, "include", "tfpwa_amd.h")).read()
    defs = dict(re.findall(r"^#define TPA_((?:FLAG|MODE)_\w+) (\d+)", header, re.M))
    assert "FLAG_F32H" in defs
    for name, val in defs.items():
        assert getattr(_lib, name) == int(val), name


def test_phase_space_generator_invariants():
    from tf_pwa_amd.phasespace import generate_phsp

    ps = generate_phsp(2.286, [0.938, 0.494, 0.140], 20000, np.random.default_rng(1))
    tot = ps[0] + ps[1] + ps[2]
    np.testing.assert_allclose(tot[:, 0], 2.286, atol=1e-12)
    np.testing.assert_allclose(tot[:, 1:], 0.0, atol=1e-12)
    for p, m in zip(ps, [0.938, 0.494, 0.140]):
        np.testing.assert_allclose(np.sqrt(np.abs(p[:, 0] ** 2 - (p[:, 1:] ** 2).sum(1))), m,
                                   atol=1e-12)
    # flat in the Dalitz plane: m^2(K pi) histogram density is uniform inside the boundary
    m23sq = (ps[1][:, 0] + ps[2][:, 0]) ** 2 - ((ps[1][:, 1:] + ps[2][:, 1:]) ** 2).sum(1)
    assert m23sq.min() > (0.494 + 0.140) ** 2 - 1e-12
    assert m23sq.max() < (2.286 - 0.938) ** 2 + 1e-12


def test_grad_from_adjoint_matches_jacobian():
    from tf_pwa_amd import ConfigLoader

    c = ConfigLoader(config_path("lc_pkpi_6res.yml"), seed=9)
    amp = c.get_amplitude()
    rng = np.random.default_rng(2)
    G = rng.normal(size=amp.P) + 1j * rng.normal(size=amp.P)
    J = amp.jacobian()
    np.testing.assert_allclose(amp.grad_from_adjoint(G), 2 * np.real(J @ G), rtol=1e-12,
                               atol=1e-12)


def test_cached_fg_patches_nan_gradient():
    from tf_pwa_amd.fit import CachedFG

    def fg(x):
        g = 2 * x
        g[1] = np.nan
        return float(np.sum(x * x)), g

    c = CachedFG(fg)
    f, g = c(np.array([1.0, 2.0]))
    assert abs(g[1] - 4.0) < 1e-5 and g[0] == 2.0
    assert c(np.array([1.0, 2.0])) is c.fg  # cached


def test_fit_scipy_on_quadratic_with_bounds():
    from tf_pwa_amd.fit import fit_scipy
    from tf_pwa_amd.variable import VarsManager

    vm = VarsManager()
    vm.add_real_var("a", 0.3)
    vm.add_real_var("b", 0.8)
    vm.set_bound({"b": (0.0, 1.0)})

    class F:
        def __init__(self):
            self.vm = vm

        def nll_grad(self, y):
            a, b = y
            return (a - 0.7) ** 2 + (b - 0.25) ** 2, np.array([2 * (a - 0.7), 2 * (b - 0.25)])

    r = fit_scipy(F())
    assert abs(r.params["a"] - 0.7) < 1e-5 and abs(r.params["b"] - 0.25) < 1e-5
    assert r.min_nll < 1e-9


def test_barrier_constant_mass_derivatives():
    import yaml

    from tf_pwa_amd import ConfigLoader

    cfg = yaml.safe_load(open(config_path("lc_pkpi_6res.yml")))
    for r in ("Kst", "K0st", "L1520", "D1232", "L1405", "L1820"):
        cfg["particle"][r]["float"] = "mg"
    c = ConfigLoader(cfg, seed=1)
    amp = c.get_amplitude()
    assert amp.shape_mode
    bc = amp.barrier_constants()
    for ci, ch in enumerate(amp.chains):
        name = ch["R"].name + "_mass"
        m0 = c.vm.get(name)
        h = 1e-6
        c.vm.set(name, m0 + h)
        bp = amp.barrier_constants()[ci]
        c.vm.set(name, m0 - h)
        bm = amp.barrier_constants()[ci]
        c.vm.set(name, m0)
        for key, dkey in (("c1", "dlog1"), ("c2", "dlog2")):
            fd = (np.log(bp[key]) - np.log(bm[key])) / (2 * h)
            np.testing.assert_allclose(bc[ci][dkey], fd, atol=1e-6, rtol=1e-6)


@pytest.mark.parametrize("cfg", ["lc_pkpi_2res.yml", "lc_pkpi_6res.yml"])
def test_serving_coupling_map_reproduces_coefficients(cfg):
    """The native server rebuilds c_k from the coupling map; the map must give the model's
    coefficients at any parameters, polar and cartesian factors alike."""
    import yaml

    from tf_pwa_amd import ConfigLoader
    from tf_pwa_amd.serving import coupling_map, map_coefficients

    for polar in (True, False):
        conf = yaml.safe_load(open(config_path(cfg)))
        conf["polar"] = polar
        c = ConfigLoader(conf, seed=5)
        amp = c.get_amplitude()
        _, keep = coupling_map(amp)
        rng = np.random.default_rng(3)
        for _ in range(3):
            y = np.array(c.vm.get_all_val()) + rng.normal(scale=0.3, size=len(c.vm.trainable_vars))
            c.vm.set_all(y)
            np.testing.assert_allclose(map_coefficients(keep, y), amp.coefficients(), rtol=1e-14,
                                       atol=1e-14)
        assert keep["factor_var"].max() < len(c.vm.trainable_vars)
        # the fixed chain total of fix_chain_idx appears as a fixed factor
        assert (keep["factor_var"] < 0).any()


def test_hessian_chain_rule_on_a_known_function():
    """hess_from_adjoint / grad_from_adjoint against central differences of
    F(theta) = 2 Re(a^T c) + 2 Re(c^T B c) + c^H C c, whose Wirtinger derivatives are
    F1 = a + 2 B c + C^T conj(c), Fcc = 2 B, Fcb = C^T."""
    import yaml

    from tf_pwa_amd import ConfigLoader

    for polar in (True, False):
        conf = yaml.safe_load(open(config_path("lc_pkpi_6res.yml")))
        conf["polar"] = polar
        c = ConfigLoader(conf, seed=8)
        amp = c.get_amplitude()
        P = amp.P
        rng = np.random.default_rng(4)
        a = rng.normal(size=P) + 1j * rng.normal(size=P)
        B = rng.normal(size=(P, P)) + 1j * rng.normal(size=(P, P))
        B = 0.5 * (B + B.T)
        C = rng.normal(size=(P, P)) + 1j * rng.normal(size=(P, P))
        C = 0.5 * (C + C.conj().T)

        def F(x):
            c.vm.set_all(x)
            cc = amp.coefficients()
            return float(2 * np.real(a @ cc) + 2 * np.real(cc @ B @ cc)
                         + np.real(cc.conj() @ C @ cc))

        x0 = np.array(c.vm.get_all_val())
        c.vm.set_all(x0)
        cc = amp.coefficients()
        F1 = a + 2 * B @ cc + C.T @ cc.conj()
        g = amp.grad_from_adjoint(F1)
        H = amp.hess_from_adjoint(F1, 2 * B, C.T)
        n = len(x0)
        h = 1e-4
        for i in range(n):
            e = np.zeros(n)
            e[i] = h
            gd = (F(x0 + e) - F(x0 - e)) / (2 * h)
            assert abs(gd - g[i]) < 1e-6 * max(1, abs(g).max())
        for i in range(0, n, 3):
            for j in range(0, n, 2):
                ei, ej = np.zeros(n), np.zeros(n)
                ei[i], ej[j] = h, h
                fd = (F(x0 + ei + ej) - F(x0 + ei - ej) - F(x0 - ei + ej) + F(x0 - ei - ej)) / (
                    4 * h * h)
                assert abs(fd - H[i, j]) < 1e-5 * max(1, np.abs(H).max()), (i, j, fd, H[i, j])
        c.vm.set_all(x0)


def test_second_moment_decode_layout():
    """decode_second_moments against explicit sums over synthetic q and h vectors packed the
    way the kernel packs them (upper triangles of O1 = sum v z z^T, O2 = sum u y y^T)."""
    from tf_pwa_amd.data import decode_second_moments

    P, n, L = 5, 40, 3
    rng = np.random.default_rng(0)
    h = rng.normal(size=(n, L, P)) + 1j * rng.normal(size=(n, L, P))
    q = rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))
    u, v = rng.normal(size=n), rng.normal(size=n)
    z = np.concatenate([q.real, q.imag], 1)
    y = np.concatenate([h.real, h.imag], 2).reshape(n * L, 2 * P)
    O1 = np.einsum("e,ei,ej->ij", v, z, z)
    O2 = np.einsum("e,ei,ej->ij", np.repeat(u, L), y, y)
    iu = np.triu_indices(2 * P)
    G = (u[:, None] * q).sum(0)
    out = np.concatenate([[1.5, 0.0], G.view(np.float64), O1[iu], O2[iu]])
    val, G2, Dcc, Dcb = decode_second_moments(out, P)
    assert val == 1.5
    np.testing.assert_allclose(G2, G, rtol=1e-13)
    np.testing.assert_allclose(Dcc, np.einsum("e,ek,el->kl", v, q, q), rtol=1e-12, atol=1e-12)
    ref = np.einsum("e,ek,el->kl", v, q, q.conj()) + np.einsum("e,elk,elm->km", u, h, h.conj())
    np.testing.assert_allclose(Dcb, ref, rtol=1e-12, atol=1e-12)


def test_genbod_against_the_three_body_generator():
    """The n-body GENBOD generator, run for three bodies, against the independent m23-weight
    three-body generator: the same m23^2 distribution (two-sample KS), exact four-momentum
    conservation and masses; and for four bodies conservation, masses and kinematic limits."""
    from scipy.stats import ks_2samp

    from tf_pwa_amd.phasespace import _genbod, generate_phsp

    m0, ms = 2.286, [0.938, 0.494, 0.140]
    a = _genbod(m0, ms, 40000, np.random.default_rng(1))
    b = generate_phsp(m0, ms, 40000, np.random.default_rng(2))

    def m2(p, q):
        s = p + q
        return s[:, 0] ** 2 - (s[:, 1:] ** 2).sum(1)

    assert ks_2samp(m2(a[1], a[2]), m2(b[1], b[2])).pvalue > 1e-3
    assert ks_2samp(m2(a[0], a[1]), m2(b[0], b[1])).pvalue > 1e-3
    ms4 = [0.493677, 0.493677, 0.13957, 0.13957]
    p = generate_phsp(3.0969, ms4, 20000, np.random.default_rng(3))
    tot = sum(p)
    np.testing.assert_allclose(tot[:, 0], 3.0969, atol=1e-12)
    np.testing.assert_allclose(tot[:, 1:], 0.0, atol=1e-12)
    for x, m in zip(p, ms4):
        np.testing.assert_allclose(np.sqrt(np.abs(x[:, 0] ** 2 - (x[:, 1:] ** 2).sum(1))), m,
                                   atol=1e-12)
    m01 = np.sqrt(m2(p[0], p[1]))
    assert m01.min() > ms4[0] + ms4[1] - 1e-12 and m01.max() < 3.0969 - ms4[2] - ms4[3] + 1e-12


def test_four_body_layout_matches_oracle():
    from tf_pwa_amd import ConfigLoader
    from tf_pwa_amd.serving import coupling_map, map_coefficients

    c = ConfigLoader(config_path("jpsi_kkpipi_12ch.yml"), seed=0)
    m = O.Model(config_path("jpsi_kkpipi_12ch.yml"))
    assert c.vm.trainable_vars == m.trainable
    amp = c.get_amplitude()
    assert (amp.n_final, amp.P, amp.lam, len(amp.chains), len(amp.topo)) == (4, 25, 2, 12, 6)
    _, keep = coupling_map(amp)
    y = np.array(c.vm.get_all_val())
    np.testing.assert_allclose(map_coefficients(keep, y), amp.coefficients(), rtol=1e-14)
    M = amp.describe()
    assert M.P == 25 and M.n_chain == 12


@pytest.mark.parametrize("cfg", ["lc_pkpi_2res.yml", "lc_pkpi_6res.yml"])
def test_fused_source_indices_stay_inside_their_arrays(cfg):
    """The generated fused pass (tpa_fused_source, host code): its constants match the model
    and every literal index it emits lies inside the array it indexes (angle-cache fields, the
    coefficient, lineshape and clamp-flag arguments, the amplitude and adjoint arrays)."""
    from tf_pwa_amd import ConfigLoader

    amp = ConfigLoader(config_path(cfg), seed=1).get_amplitude()
    src = amp.fused_source()
    M = amp.describe()
    define = dict(re.findall(r"#define (TPA_\w+) (\d+)", src))
    assert int(define["TPA_P"]) == amp.P and int(define["TPA_NC"]) == M.n_chain
    ns = int(define["TPA_NSLOT"])
    assert ns == amp.lam

    def top(pat):
        return max(int(x) for x in re.findall(pat, src))

    assert top(r"ang\[(\d+)LL \* S") == 18 * M.n_topo + 3  # the last of the four masses
    assert top(r"a\.c\[(\d+)\]") == 2 * amp.P - 1
    for arr in ("m0", "g0", "q0", "act"):
        assert top(r"a\.%s\[(\d+)\]" % arr) == M.n_chain - 1
    assert top(r"gr\[(\d+)\]") == amp.P - 1
    assert top(r"[^_]A\[(\d+)\] =") == ns - 1
    assert top(r"_(?:Ap|B)\[(\d+)\] =") == ns - 1
    assert "extern \"C\" __global__ __launch_bounds__(256) void tpa_fused_pass" in src


def test_fused_pass_size_limit():
    """Models above TPA_FUSED_MAX_P partial waves are refused before any device work; the
    limit in the header and in the bindings agree."""
    from tf_pwa_amd import ConfigLoader, Model, _lib

    header = open(os.path.join(ROOT, "include", "tfpwa_amd.h")).read()
    assert int(re.search(r"#define TPA_FUSED_MAX_P (\d+)", header).group(1)) == _lib.FUSED_MAX_P
    small = Model(ConfigLoader(config_path("lc_pkpi_2res.yml"), seed=1).get_amplitude(),
                  evaluation="fused")
    assert small.evaluation == "fused"
    big = ConfigLoader(config_path("lc_pkpi_6res.yml"), seed=1).get_model()
    with pytest.raises(NotImplementedError, match="disabled above 8 partial waves"):
        big.evaluation = "fused"
    assert big.evaluation == "stream"
