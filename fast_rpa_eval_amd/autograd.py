"""torch.autograd.Function wrappers over the C-ABI forward/adjoint pairs.

* ``Rho1Function``: P1 -> first_order_rho; backward is the first-order H integration
  d L / d P1(mu, nu, t) = sum_p dL/drho1(map(p), t) phi_mu(p) phi_nu(p)
  (frpa_rho1_adjoint; evaluate_first_order_H.cpp:144-215 with the GGA terms zero), in FP64 or the
  INT8 split.
* ``FarFieldFunction``: multipole_moments -> delta_v_hartree (far field); backward
  d L / d M(lm, a, t) = sum_p dL/dV(p, t) d V(p, t) / d M(lm, a, t) (frpa_far_adjoint).

Both forwards are linear in their differentiable input (the far field up to the reference's
|M| > 1e-10 mask), so central finite differences check them exactly up to rounding.
"""
import torch

from . import api


class Rho1Function(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density_matrix, plan, wave, T=1, precision=api.FP64, adjoint_precision=api.FP64):
        rho = plan.forward(wave, density_matrix.detach().contiguous(), T=T, precision=precision)
        ctx.plan, ctx.wave, ctx.T, ctx.adjoint_precision = plan, wave, T, adjoint_precision
        return rho

    @staticmethod
    def backward(ctx, grad_rho):
        H = ctx.plan.adjoint(ctx.wave, grad_rho.contiguous(), T=ctx.T, precision=ctx.adjoint_precision)
        return H, None, None, None, None, None


class FarFieldFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, multipole_moments, plan, inputs, variant=api.FAR_HORNER):
        M = multipole_moments.detach().contiguous()
        dv = torch.zeros(inputs.n_full_points * inputs.T, dtype=torch.float64, device=M.device)
        plan.forward(dv, inputs.coord_points, inputs.map, inputs.lmfd, inputs.mrsq, inputs.opr, M,
                     inputs.n_full_points, T=inputs.T, variant=variant)
        ctx.plan, ctx.inputs = plan, inputs
        ctx.save_for_backward(M)
        return dv

    @staticmethod
    def backward(ctx, grad_v):
        (M,) = ctx.saved_tensors
        inp = ctx.inputs
        gM = ctx.plan.adjoint(grad_v.contiguous(), inp.coord_points, inp.map, inp.lmfd, inp.mrsq, inp.opr, M,
                              inp.n_full_points, T=inp.T)
        return gM, None, None, None


def rho1(density_matrix, plan, wave, T=1, precision=api.FP64, adjoint_precision=api.FP64):
    """first_order_rho as a differentiable function of the density matrix; the backward runs the
    adjoint in `adjoint_precision` (FP64 or the INT8 split)."""
    return Rho1Function.apply(density_matrix, plan, wave, T, precision, adjoint_precision)


def far_field(multipole_moments, plan, inputs, variant=api.FAR_HORNER):
    return FarFieldFunction.apply(multipole_moments, plan, inputs, variant)
