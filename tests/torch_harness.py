"""torch.autograd.Function around the GPU discrete adjoint, used by the tests only: forward runs the
steps and returns the loss, backward calls spruce_gpu_adjoint for the gradient."""
import numpy as np
import torch


class PeriodicWaveLoss(torch.autograd.Function):
    last_ratio = None

    @staticmethod
    def forward(ctx, temp, v0, problem):
        loss, rho, sec = problem.forward(temp.detach().numpy(), float(v0))
        ctx.problem = problem
        ctx.rho = rho
        ctx.forward_sec = sec
        ctx.save_for_backward(temp, v0)
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        temp, v0 = ctx.saved_tensors
        g_temp, g_v0, sec = ctx.problem.gradient(temp.detach().numpy(), float(v0), ctx.rho)
        PeriodicWaveLoss.last_ratio = sec / ctx.forward_sec
        return grad_out * torch.from_numpy(np.ascontiguousarray(g_temp)), grad_out * torch.tensor(g_v0), None
