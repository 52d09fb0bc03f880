"""torch.autograd wrapper of the reverse sweep, for gradient checks.

    J = SweepFunctional.apply(coef, sigma, solver, weights)
    J.backward()            # coef.grad, sigma.grad from moc_adjoint

`coef` holds (c0, c1, c2) per FSR and group ([nsr*fai*G, 3]), `sigma`
holds sigma_t per region and group ([nsr*G]); both are handed to the solver
(moc_set_coefficients) and the functional is the weighted scalar flux of
the next sweep.  The forward pass already runs the adjoint (one reverse
sweep gives every gradient), so backward only scales by dL/dJ.  The solver
state is not changed.  torch is imported on use only.
"""
from __future__ import annotations

import numpy as np


def _function_class():
    import torch

    class SweepFunctional(torch.autograd.Function):
        @staticmethod
        def forward(ctx, coef, sigma, solver, weights):
            c = coef.detach().to("cpu", torch.float32).numpy()
            s = sigma.detach().to("cpu", torch.float32).numpy()
            solver.set_coefficients(c, s)
            try:
                J, gs, gc = solver.adjoint(np.asarray(weights, np.float32))
            finally:
                solver.set_coefficients(None, None)
            ctx.save_for_backward(torch.from_numpy(gc).to(coef.device),
                                  torch.from_numpy(gs).to(sigma.device))
            ctx.dtypes = (coef.dtype, sigma.dtype)
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            gc, gs = ctx.saved_tensors
            g = grad_out.to(torch.float64)
            return (g * gc).to(ctx.dtypes[0]), (g * gs).to(ctx.dtypes[1]), None, None

    return SweepFunctional


_cls = None


def SweepFunctional():
    """The torch.autograd.Function class (built on first use)."""
    global _cls
    if _cls is None:
        _cls = _function_class()
    return _cls
