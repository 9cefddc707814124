"""torch.autograd wrapper of the path-replay adjoint, for the test harness only.

ImageLoss.apply(rho, engine, W, npackets, seed, precision) runs
mcrt_adjoint_image at the densities `rho` (a CPU float64 tensor in the
engine model's cell shape) and returns the loss sum(W * image_I); its
backward scales the adjoint's dL/drho by the incoming gradient.
"""
import numpy as np
import torch


class ImageLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, engine, W, npackets, seed, precision):
        engine.set_density(rho.detach().cpu().numpy())
        loss, grad = engine.adjoint_image(W, npackets, seed, precision=precision, grad=True)
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(grad)))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g_out):
        (grad,) = ctx.saved_tensors
        return g_out * grad, None, None, None, None, None
