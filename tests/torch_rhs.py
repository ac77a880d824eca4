"""torch.autograd.Function over the C ABI (test harness only).

forward: q -> R(q), Region%computeRhs(FORWARD) on the GPU.
backward: the vector-Jacobian product from the hand-written discrete adjoint rhs.
magudi's adjoint satisfies <R_adj(lambda), dq>_H = -<lambda, R'(q) dq>_H with the
SBP norm H = diag(norm) (test/adjoint_relation/full_rhs.f90), so the Euclidean VJP
is  v -> -norm * R_adj(v / norm)  (adjoint forcing off, so R_adj is exactly linear
in lambda)."""
import numpy as np
import torch


class MagudiRhs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, region, time):
        qn = q.detach().cpu().numpy().astype(np.float64)
        region.states[0].conservedVariables = qn
        region.computeRhs(1, time=time)
        ctx.region, ctx.q = region, qn
        return torch.from_numpy(region.states[0].rightHandSide.copy())

    @staticmethod
    def backward(ctx, grad_out):
        R = ctx.region
        norm = R.grid(2)[0]
        v = grad_out.detach().cpu().numpy().astype(np.float64)
        R.states[0].conservedVariables = ctx.q
        R.states[0].adjointVariables = v / norm
        R.updateAdjointForcing(no_adjoint_forcing=True)
        R.computeRhs(-1, forcing_factor=0.0)
        vjp = -norm * R.states[0].rightHandSide
        return torch.from_numpy(vjp), None, None
