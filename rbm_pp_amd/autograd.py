"""ln|psi| as a differentiable function of the parameter vector, for checks
and for callers that compose the VMC estimators with torch.

The forward pass evaluates psi(x; gamma) of a batch of configurations on the
GPU (rbmpp_local); the backward pass uses the library's hand-written adjoint,
the log-derivatives O_k = d psi / d gamma_k of calc_phys_quantity.f90:114-222
(RBM part) and slater.f90:68-170 (pair-product part), so
d ln|psi| / d gamma_k = Re(O_k / psi).  Nothing here differentiates through
the kernels; torch only carries the tensors.
"""
from __future__ import annotations

import numpy as np
import torch


class LogAbsPsi(torch.autograd.Function):
    """y_s = ln|psi(x_s; gamma)| for the configurations xs (int32 +-1,
    shape [n, N]); gamma is a float64 tensor of the packed parameters in the
    order of main.f90:263-294."""

    @staticmethod
    def forward(ctx, gamma: torch.Tensor, engine, xs: np.ndarray) -> torch.Tensor:
        engine.set_gamma(gamma.detach().to("cpu", torch.float64).numpy())
        out = engine.local(xs, want_o=True)
        psi = out["psi"]
        dlog = (out["oloc"] / psi[:, None]).real  # d ln|psi| / d gamma
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(dlog)).to(gamma.device))
        return torch.from_numpy(np.log(np.abs(psi))).to(gamma.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (dlog,) = ctx.saved_tensors
        return grad_out.to(dlog.dtype) @ dlog, None, None


def log_abs_psi(gamma: torch.Tensor, engine, xs) -> torch.Tensor:
    return LogAbsPsi.apply(gamma, engine, np.ascontiguousarray(xs, np.int32))
