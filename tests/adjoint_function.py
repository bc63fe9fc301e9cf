"""torch.autograd.Function over the library's differentiated solver (test harness only).

forward calls fr_adjoint_forward (K unrolled FP64 iterations and the objective), backward calls
fr_adjoint_backward (the hand-written adjoint kernels); the tensors are CPU float64 tensors, copied
to and from the device by the fastreg_amd binding."""
import numpy as np
import torch


class SolverObjective(torch.autograd.Function):
    """E(v0, w_reg, tau) of an fastreg_amd.Adjoint whose images are set; label_weight fixed."""

    @staticmethod
    def forward(ctx, v0, w_reg, tau, adj, label_weight):
        args = (v0.detach().cpu().numpy().astype(np.float64), float(w_reg), float(tau), label_weight)
        e = adj.forward(*args)
        ctx.adj, ctx.args, ctx.call = adj, args, adj.forward_calls
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e):
        if ctx.adj.forward_calls != ctx.call:  # another forward pass ran since: redo this one
            ctx.adj.forward(*ctx.args)
            ctx.call = ctx.adj.forward_calls
        g, gw, gt = ctx.adj.backward()
        s = grad_e.item()
        return (torch.from_numpy(g * s), torch.tensor(gw * s, dtype=torch.float64),
                torch.tensor(gt * s, dtype=torch.float64), None, None)
