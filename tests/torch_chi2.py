"""Test harness: chi^2 of a device batch as a torch.autograd.Function whose
backward is the library's adjoint gradient (pk_batch_evaluate_grad)."""
import numpy as np
import torch


class Chi2Function(torch.autograd.Function):
    """chi2 = Chi2Function.apply(theta, batch, fit_model): theta (float64, in
    fit_model.theta() order) -> chi^2 of the batch's single model."""

    @staticmethod
    def forward(ctx, theta, batch, fm):
        spec = fm.spec(theta.detach().cpu().numpy())
        batch.set_components([spec.m.components[c] for c in range(len(fm.kinds))])
        batch.evaluate_grad()
        chi2, grad = batch.chi2_grad()
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(fm.flat_grad(grad))).to(theta.dtype))
        return theta.new_tensor(float(chi2[0]))

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None
