"""Leave-one-out cross-validation MSE as a differentiable function of the
variogram parameters, for PyTorch optimisers.

``loo_mse(k, theta, base)`` returns a 0-d tensor whose backward pass uses the
gradient from the library's adjoint kernel (``Kriging.loo_grad``): no
autograd tape is recorded through the kriging systems.  ``theta`` holds the
parameters named in ``names`` (default: nugget, sill, a_major, a_minor1,
azimuth); the other parameters of the structure come from ``base``.
"""
import torch

from .kriging import Kriging

DEFAULT_NAMES = ("nugget", "sill", "a_major", "a_minor1", "azimuth")


class _LOOMSE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, k, base, names):
        vals = dict(base)
        for n, v in zip(names, theta.detach().cpu().double().tolist()):
            vals[n] = v
        k.set_vgm(1, 1, append=False, **vals)
        mse, grad = k.loo_grad()
        ctx.save_for_backward(torch.tensor([grad[n] for n in names], dtype=theta.dtype,
                                           device=theta.device))
        return theta.new_tensor(mse)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return gout * g, None, None, None


def loo_mse(k: Kriging, theta: torch.Tensor, base: dict, names=DEFAULT_NAMES) -> torch.Tensor:
    """``base``: set_vgm keywords of the single structure (vtype, nugget, sill,
    a_major, ...); the entries named in ``names`` are taken from ``theta``."""
    return _LOOMSE.apply(theta, k, base, tuple(names))
