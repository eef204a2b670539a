"""Fit fractions from the partial-wave Gram matrix of the MC sample.

tf-pwa integrates |A_R|^2 and |A_R + A_S|^2 over MC once per resonance and pair with tape
gradients (cal_fitfractions, tf_pwa/fitfractions.py:178-258). With the Gram matrix
I_kl = sum_mc w conj(h_k) h_l every such integral is a quadratic form of the coefficients:
int |sum_{k in R} c_k h_k|^2 = c_R^dagger I c_R, with c_R the coefficients of R's partial waves.
So all fractions and their gradients are P x P algebra on the host after one pass.

FF_R = c_R^H I c_R / c^H I c, and FF_RS = (c_RS^H I c_RS) / c^H I c - FF_R - FF_S for the
interference of R and S; the gradient of a quadratic form N = c^H M c in the parameters is
2 Re sum_k (dc_k/dtheta) conj((M c)_k) (the chain rule of grad_from_adjoint). With an error
matrix V the error of a fraction is sqrt(g^T V g) (FitFractions.get_frac, fitfractions.py:137-146).
"""

import numpy as np


def gram_fp64(mc):
    """FP64 Gram matrix I_kl = sum w conj(h_k) h_l of an event set (FP64 MFMA, tpa_gram_f64)."""
    return mc.gram("fp64")


def _groups(amp, res):
    names = [ch["R"].name for ch in amp.chains]
    if res is None:
        res = []
        for n in names:
            if n not in res:
                res.append(n)
    idx = {}
    for r in res:
        ks = [k for k, w in enumerate(amp.waves) if names[w.chain] == str(r)]
        if not ks:
            raise ValueError("no partial wave belongs to {}".format(r))
        idx[str(r)] = np.array(ks)
    return [str(r) for r in res], idx


def cal_fitfractions(amp, mc=None, res=None, gram=None, error_matrix=None, sum_diag=True):
    """(fractions, errors or gradients) keyed like tf-pwa: resonance name for FF_R and
    (name_i, name_j) for the interference of the i-th and j-th resonance (j < i), plus
    "sum_diag". ``gram``: the MC Gram matrix (default: FP64 Gram of ``mc``, unit-normalised
    weights are not needed since fractions are ratios). With ``error_matrix`` (trainable_vars
    order) the second dict holds errors, otherwise gradients."""
    if amp.shape_mode:
        raise NotImplementedError("fit fractions from the Gram matrix need fixed lineshapes")
    if gram is None:
        gram = gram_fp64(mc)
    I = np.asarray(gram)
    c = amp.coefficients()
    names, idx = _groups(amp, res)

    def quad(sel):
        cs = np.zeros_like(c)
        cs[sel] = c[sel]
        Mc = I @ cs
        # only the selected rows contribute: N = cs^H I cs, dN/dc_k = conj((I cs)_k) for k in sel
        adj = np.zeros_like(c)
        adj[sel] = np.conj(Mc[sel])
        return float(np.real(np.vdot(cs, Mc))), amp.grad_from_adjoint(adj)

    tot, g_tot = quad(np.arange(amp.P))
    frac, grad = {}, {}
    for i, ri in enumerate(names):
        n, g = quad(idx[ri])
        frac[ri] = n / tot
        grad[ri] = g / tot - n * g_tot / tot ** 2
        for j in range(i - 1, -1, -1):
            rj = names[j]
            n, g = quad(np.concatenate([idx[ri], idx[rj]]))
            key = (ri, rj)
            frac[key] = n / tot - frac[ri] - frac[rj]
            grad[key] = g / tot - n * g_tot / tot ** 2 - grad[ri] - grad[rj]
    if sum_diag:
        frac["sum_diag"] = sum(frac[r] for r in names)
        grad["sum_diag"] = sum(grad[r] for r in names)
    if error_matrix is None:
        return frac, grad
    V = np.asarray(error_matrix)
    return frac, {k: float(np.sqrt(abs(g @ V @ g))) for k, g in grad.items()}
