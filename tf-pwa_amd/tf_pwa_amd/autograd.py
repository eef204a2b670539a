"""The likelihood as a torch operation: ``nll(fcn, x)`` returns NLL(x) as a 0-d tensor whose
backward pass is the engine's adjoint gradient.

The forward pass is one ``FCN.nll_grad`` call (tf_pwa/model/model.py:1207-1229): the data and
MC passes on the device return the value and the adjoint sums G_k (and the lineshape adjoints
when masses or widths float), which the host chain-rules to d NLL / d x. There is no tape: the
gradient is computed with the value and held for ``backward``, which scales it by the
incoming gradient. PyTorch carries the parameter vector and the autograd graph around the
call, so the NLL composes with other torch terms (penalties, priors) and torch optimisers.
``x`` is the vector of trainable values in ``fcn.vm.trainable_vars`` order.

``x`` must be a CPU tensor. The parameters are host values (the engine takes them as a host
array per call), and the engine owns the GPU work of its process: a torch tensor on the GPU
would put torch's own kernels (copies, fills) on the device beside the engine's, so such
tensors are refused rather than moved.
"""

import numpy as np
import torch


class NLLFunction(torch.autograd.Function):
    """forward(x, fcn) -> NLL(x); backward -> grad_output * dNLL/dx from the engine."""

    @staticmethod
    def forward(ctx, x, fcn):
        xs = x.detach().to(torch.float64).numpy()
        value, grad = fcn.nll_grad(np.ascontiguousarray(xs))
        ctx.save_for_backward(torch.as_tensor(np.asarray(grad, dtype=np.float64)).to(
            dtype=x.dtype))
        return torch.tensor(float(value), dtype=x.dtype)

    @staticmethod
    def backward(ctx, grad_output):
        (grad,) = ctx.saved_tensors
        return grad_output * grad, None


def _check_host(x):
    if x.device.type != "cpu":
        raise ValueError("x must be a CPU tensor (got one on {}): the parameters are host "
                         "values and the engine alone runs work on the GPU".format(x.device))


def nll(fcn, x):
    """NLL of ``fcn`` at the trainable values ``x`` (a 1-d CPU tensor) as a differentiable
    torch scalar."""
    _check_host(x)
    if x.dim() != 1 or x.numel() != len(fcn.vm.trainable_vars):
        raise ValueError("x must hold one value per trainable variable ({})".format(
            len(fcn.vm.trainable_vars)))
    return NLLFunction.apply(x, fcn)


def hessian(fcn, x):
    """The exact Hessian of the NLL at ``x`` as a tensor (FCN.nll_grad_hessian: second-moment
    passes for couplings and lineshape parameters)."""
    _check_host(x)
    xs = x.detach().to(torch.float64).numpy()
    _, _, H = fcn.nll_grad_hessian(np.ascontiguousarray(xs))
    return torch.as_tensor(np.asarray(H, dtype=np.float64)).to(dtype=x.dtype)
