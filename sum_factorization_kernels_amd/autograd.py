"""torch.autograd bindings of the operators whose adjoints and parameter
gradients are hand-written HIP kernels (no autodiff of the kernels):

  advect(u, vel_cell, vel_face, op)   y = M^{-1} L(a) u, the DG upwind
      advection operator with per-point velocity a
      (legacy/include/evaluation_dg_advect.h:285-676 with a variable a);
      backward: u-bar = A^T y-bar (sfk_apply_transpose),
      a-bar = d(y-bar^T A(a) u)/da (sfk_grad_velocity).
  laplace_coef(u, k, op)   y = A(k) u, -div(k grad u) on the continuous
      Q_p mesh (n_rhs vectors); backward: u-bar = A(k) y-bar (symmetric),
      k-bar = d(y-bar^T A(k) u)/dk (sfk_grad_coefficient).

Inputs are contiguous CUDA tensors in the operator's precision.  The
operator handle holds the parameter last set; every call sets it again."""
import torch


class _Advect(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, vel_cell, vel_face, op):
        op.set_velocity_device(vel_cell, vel_face)
        y = op.apply(u)
        ctx.save_for_backward(u, vel_cell, vel_face)
        ctx.op = op
        return y

    @staticmethod
    def backward(ctx, gy):
        u, vc, vf = ctx.saved_tensors
        op = ctx.op
        gy = gy.contiguous()
        op.set_velocity_device(vc, vf)
        gu = op.apply_transpose(gy) if ctx.needs_input_grad[0] else None
        gc = gf = None
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            gc, gf = op.grad_velocity(u, gy)
        return gu, gc, gf, None


class _LaplaceCoef(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, k, op):
        op.set_coefficient_device(k)
        y = op.apply(u)
        ctx.save_for_backward(u, k)
        ctx.op = op
        return y

    @staticmethod
    def backward(ctx, gy):
        u, k = ctx.saved_tensors
        op = ctx.op
        gy = gy.contiguous()
        op.set_coefficient_device(k)
        gu = op.apply(gy) if ctx.needs_input_grad[0] else None
        gk = op.grad_coefficient(u, gy) if ctx.needs_input_grad[1] else None
        return gu, gk, None


def advect(u, vel_cell, vel_face, op):
    return _Advect.apply(u, vel_cell, vel_face, op)


def laplace_coef(u, k, op):
    return _LaplaceCoef.apply(u, k, op)
