"""torch.autograd wrappers of the pathwise batch derivative.

The device computes everything (tangent push, th_trace_tangent); torch only
carries the result through autograd:

* ``IonIncrement.apply(theta, batch, rho_s)``: the face-smoothed
  surface-charge increment of ``batch`` (a tangent.PathwiseBatch) at the ion
  distribution theta = (angular sigma [deg], energy shift [eV]); backward
  (vector-Jacobian product) and forward-mode JVPs through the 2-column
  Jacobian of two tangent pushes.
* ``ChargeIncrement.apply(rho_s, batch)``: the increment as a function of the
  surface charge; forward-mode JVPs by one tangent push along the tangent's
  charge direction; backward (vector-Jacobian product) by one adjoint push
  (batch.vjp: the transpose of the tangent, for all surface cells at once).
* ``EnergyIncrement.apply(mean_energy, batch, rho_s)``: the same for the mean
  ion energy, along the exponentially tilted family of the energy
  distribution (batch.energy_score_gradient, a likelihood ratio with a
  bounded score).
* ``SigmaIncrement.apply(sigma, batch, rho_s)``: the increment as a function
  of the ion angular sigma [deg], differentiated by the likelihood ratio
  (batch.sigma_score_gradient): an unbiased estimate of the derivative of
  the expected increment, from the same trace as the forward value.
"""
import numpy as np
import torch


class IonIncrement(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, batch, rho_s=None):
        th = theta.detach().cpu().numpy().astype(float)
        rho = None if rho_s is None else np.asarray(rho_s, dtype=float)
        inc, jac = batch.ion_jacobian(rho, float(th[0]), float(th[1]))
        jac_t = torch.as_tensor(jac, dtype=theta.dtype, device=theta.device)
        ctx.save_for_backward(jac_t)
        ctx.save_for_forward(jac_t)
        return torch.as_tensor(inc, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad):
        jac, = ctx.saved_tensors
        return jac.T @ grad, None, None

    @staticmethod
    def jvp(ctx, dtheta, *_):
        jac, = ctx.saved_tensors
        return jac @ dtheta


class ChargeIncrement(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho_s, batch):
        ctx.batch = batch
        ctx.rho = rho_s.detach().cpu().numpy().astype(float)
        inc = batch.increment(ctx.rho)
        return torch.as_tensor(inc, dtype=rho_s.dtype, device=rho_s.device)

    @staticmethod
    def jvp(ctx, drho, _):
        d = drho.detach().cpu().numpy().astype(float)
        dinc = ctx.batch.charge_jvp(ctx.rho, d)[1]
        return torch.as_tensor(dinc, dtype=drho.dtype, device=drho.device)

    @staticmethod
    def backward(ctx, grad):
        g = grad.detach().cpu().numpy().astype(float)
        out = ctx.batch.vjp(ctx.rho, g)[1]
        return torch.as_tensor(out, dtype=grad.dtype, device=grad.device), None


class EnergyIncrement(torch.autograd.Function):
    """The increment as a function of the mean ion energy [eV] (the tilted
    family, batch.energy_score_gradient); the value is the batch's own at the
    untilted distribution, and the input only selects where the gradient is
    reported."""

    @staticmethod
    def forward(ctx, mean_energy, batch, rho_s=None):
        rho = None if rho_s is None else np.asarray(rho_s, dtype=float)
        inc, grad = batch.energy_score_gradient(rho)
        g = torch.as_tensor(grad, dtype=mean_energy.dtype, device=mean_energy.device)
        ctx.save_for_backward(g)
        ctx.save_for_forward(g)
        ctx.shape = mean_energy.shape
        return torch.as_tensor(inc, dtype=mean_energy.dtype, device=mean_energy.device)

    @staticmethod
    def backward(ctx, grad):
        g, = ctx.saved_tensors
        return (g @ grad).reshape(ctx.shape), None, None

    @staticmethod
    def jvp(ctx, dmean, *_):
        g, = ctx.saved_tensors
        return g * dmean.reshape(-1)[0]


class SigmaIncrement(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma, batch, rho_s=None):
        rho = None if rho_s is None else np.asarray(rho_s, dtype=float)
        inc, grad = batch.sigma_score_gradient(rho, float(sigma.detach().cpu().reshape(-1)[0]), 0.0)
        g = torch.as_tensor(grad, dtype=sigma.dtype, device=sigma.device)
        ctx.save_for_backward(g)
        ctx.save_for_forward(g)
        ctx.shape = sigma.shape
        return torch.as_tensor(inc, dtype=sigma.dtype, device=sigma.device)

    @staticmethod
    def backward(ctx, grad):
        g, = ctx.saved_tensors
        return (g @ grad).reshape(ctx.shape), None, None

    @staticmethod
    def jvp(ctx, dsigma, *_):
        g, = ctx.saved_tensors
        return g * dsigma.reshape(-1)[0]
