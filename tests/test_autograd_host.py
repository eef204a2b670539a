"""The torch operation around the likelihood (tf_pwa_amd.autograd) on the CPU, with a stand-in
FCN whose value and gradient are known in closed form: backward returns the FCN's gradient
scaled by the incoming gradient, composes with other torch terms, and keeps the tensor's
dtype. The engine's gradient itself is held to central differences in the GPU tests."""

import numpy as np
import pytest
import torch


class _Vars:
    def __init__(self, n):
        self.trainable_vars = ["x{}".format(i) for i in range(n)]


class _QuadraticFCN:
    """NLL(x) = 1/2 x^T A x + b^T x with gradient A x + b."""

    def __init__(self, n, seed=0):
        rng = np.random.default_rng(seed)
        M = rng.normal(size=(n, n))
        self.A = M @ M.T + n * np.eye(n)
        self.b = rng.normal(size=n)
        self.vm = _Vars(n)
        self.calls = 0

    def nll_grad(self, x):
        self.calls += 1
        return 0.5 * x @ self.A @ x + self.b @ x, self.A @ x + self.b

    def nll_grad_hessian(self, x):
        v, g = self.nll_grad(x)
        return v, g, self.A


def test_backward_is_the_fcn_gradient():
    from tf_pwa_amd.autograd import hessian, nll

    f = _QuadraticFCN(7)
    x0 = np.random.default_rng(1).normal(size=7)
    x = torch.tensor(x0, requires_grad=True)
    y = nll(f, x)
    assert y.dim() == 0 and abs(y.item() - f.nll_grad(x0)[0]) < 1e-12
    (3.0 * y + (x * x).sum()).backward()
    np.testing.assert_allclose(x.grad.numpy(), 3.0 * (f.A @ x0 + f.b) + 2.0 * x0, rtol=1e-13)
    assert f.calls == 2  # one engine call for the value and the gradient together (+1 above)
    np.testing.assert_allclose(hessian(f, x).numpy(), f.A)


def test_dtype_and_shape_checks():
    from tf_pwa_amd.autograd import nll

    f = _QuadraticFCN(4)
    x = torch.zeros(4, dtype=torch.float32, requires_grad=True)
    y = nll(f, x)
    assert y.dtype == torch.float32
    y.backward()
    assert x.grad.dtype == torch.float32
    np.testing.assert_allclose(x.grad.numpy(), f.b.astype(np.float32), rtol=1e-6)
    with pytest.raises(ValueError):
        nll(f, torch.zeros(5))


def test_device_tensors_are_refused():
    """Parameter vectors stay on the host: a tensor on another device is refused, not moved
    (the engine alone puts work on the GPU of its process)."""
    from tf_pwa_amd.autograd import hessian, nll

    f = _QuadraticFCN(3)
    x = torch.zeros(3, dtype=torch.float64, device="meta")
    with pytest.raises(ValueError, match="CPU tensor"):
        nll(f, x)
    with pytest.raises(ValueError, match="CPU tensor"):
        hessian(f, x)
    assert f.calls == 0
