"""Differentiable PMM negative log-likelihood.

``PMMNegLogLik`` is a torch.autograd.Function whose forward is the inside
pass (negative log-likelihood of every problem of a forest) and whose backward
is the hand-written adjoint: the outside pass over the inside state the
forward left on the device, followed by the Fisher identity on the expected
transition counts (laml_grad in the C ABI).  With a forest in level-checkpoint
mode (``Forest.set_checkpoint(16)``) the backward recomputes the inside
vectors of each 16-level block instead of storing them.

Inputs are device FP64 tensors in the forest's layout: branch lengths in level
order [num_nodes], nu and phi per problem [num_problems].  Gradients are
defined for nu > 0 and 0 < phi < 1.
"""
import torch

__all__ = ["PMMNegLogLik", "neg_log_likelihood"]


class PMMNegLogLik(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, nu, phi, forest):
        d = d.detach().contiguous()
        nu = nu.detach().contiguous()
        phi = phi.detach().contiguous()
        nll = forest.nll(d, nu, phi).clone()
        ctx.forest = forest
        ctx.token = forest.state_token
        ctx.save_for_backward(d, nu, phi)
        return nll

    @staticmethod
    def backward(ctx, g):
        d, nu, phi = ctx.saved_tensors
        F = ctx.forest
        if F.state_token != ctx.token:
            # another call replaced the resident inside state: recompute it
            F.nll(d, nu, phi)
            ctx.token = F.state_token
        gd, gnu, gphi = F.grad(d, nu, phi)
        g = g.to(gd.dtype)
        return gd * g[F.node_problem()], gnu * g, gphi * g, None


def neg_log_likelihood(forest, d, nu, phi):
    """Per-problem negative log-likelihood, differentiable w.r.t. d, nu, phi."""
    return PMMNegLogLik.apply(d, nu, phi, forest)
