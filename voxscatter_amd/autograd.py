"""torch.autograd harness over the device adjoint (CPU tensors only: the device work runs through the
library's own HIP runtime, so torch never touches the GPU here).

``Extinction(scat)(eps_re, eps_im)`` returns C_ext as a 0-d float64 tensor whose backward pass is one
adjoint solve; ``ForwardIntensity`` does the same for |F(k_hat)|^2.
"""
from __future__ import annotations

import numpy as np
import torch

from .adjoint import PerVoxelDDA
from .api import Operator
from .geometry import Scatterer


class _AdjointFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_re, eps_im, owner, which):
        eps = eps_re.detach().numpy() + 1j * eps_im.detach().numpy()
        sysd = PerVoxelDDA.build(owner.scat, eps, owner.levels, op=owner.op, pc=owner.pc)
        owner.pc = sysd.pc
        if which == "cext":
            val, gre, gim = sysd.cext_and_grad(owner.tol)
        else:
            val, gre, gim = sysd.fwd_intensity_and_grad(owner.tol)
        ctx.save_for_backward(torch.from_numpy(gre), torch.from_numpy(gim))
        return torch.tensor(val, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        gre, gim = ctx.saved_tensors
        return grad_out * gre, grad_out * gim, None, None


class _Objective:
    which = "cext"

    def __init__(self, scat: Scatterer, tol: float = 1e-10, levels: int = 2):
        self.scat = scat
        self.tol = tol
        self.levels = levels
        self.op = Operator.dda(scat.r, scat.k0)
        self.pc = None

    def __call__(self, eps_re: torch.Tensor, eps_im: torch.Tensor) -> torch.Tensor:
        return _AdjointFn.apply(eps_re, eps_im, self, self.which)


class Extinction(_Objective):
    which = "cext"


class ForwardIntensity(_Objective):
    which = "fwd"
