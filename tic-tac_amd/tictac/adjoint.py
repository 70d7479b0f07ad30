"""dU/dtheta of one channel's on-shell U-matrix as a torch.autograd.Function.

The forward and the vector-Jacobian product both come from the C-ABI (tt_solve_gradient):
the library returns U(theta) and dU/dtheta from one forward row solve and one adjoint column
solve, so backward is a contraction with the stored dU/dtheta (n_theta x nd^2 nE complex
numbers) and no Krylov iterate is kept.  torch is only the autograd plumbing here.
"""
import numpy as np
import torch


class ChannelGradient:
    """U(theta), dU/dtheta of channel `channel` of `prob` at the on-shell bins of T_lab."""

    def __init__(self, prob, ctx, channel, P, T_lab, basis=None, n_theta=12, tol=1e-12):
        self.prob, self.ctx, self.channel, self.P, self.T_lab, self.tol = prob, ctx, channel, P, list(T_lab), tol
        self.basis = basis if basis is not None else prob.basis_matrices(n_theta)
        self.last_stats = None

    def __call__(self, theta, grad=True):
        U, dU, st = self.prob.gradient(self.ctx, self.channel, self.P, theta, self.T_lab, basis=self.basis,
                                       grad=grad, tol=self.tol)
        self.last_stats = st
        return U, dU


class _UMatrix(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, solver):
        U, dU = solver(theta.detach().cpu().numpy().astype(np.float64), True)
        fctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(dU)))
        return torch.view_as_real(torch.from_numpy(np.ascontiguousarray(U))).to(theta.device)

    @staticmethod
    def backward(fctx, g):
        (dU,) = fctx.saved_tensors
        gc = torch.view_as_complex(g.detach().cpu().contiguous())
        # L depends on Re U and Im U: dL/dtheta_i = sum g_re Re dU_i + g_im Im dU_i = Re(conj(g) . dU_i)
        grad = (torch.conj(gc)[None] * dU).real.reshape(dU.shape[0], -1).sum(1)
        return grad.to(g.device), None


def u_matrix(theta, solver):
    """U(theta) as a real tensor [..., 2] (view_as_real of U[d_row, d_col, j]), differentiable in theta."""
    return _UMatrix.apply(theta, solver)
