"""torch.autograd bindings of the hand-written adjoint/gradient kernels:
gradcheck (central differences) on small meshes, and the DG advection
objective J(a) = 0.5 ||A(a) u - f||^2 of the degree-3, 48^3-cell periodic
workload with a(x) = (1, 0.5 + 0.25 sin 2 pi x, 0.25): the kernel gradient
against central finite differences (eps = 1e-6, FP64) on 64 random
directions, and forward/adjoint/gradient against the CPU oracle."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP32, TOL_FP64, advect_velocity_field, rel_err, to_dev, to_np

pytestmark = pytest.mark.gpu

sfk = pytest.importorskip("sum_factorization_kernels_amd")


def _small_velocity(degree, nc, rng):
    n = degree + 1
    ncell = int(np.prod(nc))
    cv = rng.uniform(0.2, 1.0, (ncell, 3, n ** 3))
    # face a.n bounded away from 0: the upwind side does not switch under
    # the finite-difference perturbation
    fv = rng.uniform(0.3, 1.0, 3 * ncell * n * n) * rng.choice([-1.0, 1.0], 3 * ncell * n * n)
    return cv.ravel(), fv


@pytest.mark.parametrize("degree", [1, 2])
def test_advect_gradcheck(cuda, degree):
    import torch
    from sum_factorization_kernels_amd.autograd import advect
    nc = (2, 2, 2)
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, inv_jac=(2.0, 2.0, 2.0), flags=sfk.FLAG_VARIABLE_VELOCITY)
    rng = np.random.default_rng(5000 + degree)
    cv, fv = _small_velocity(degree, nc, rng)
    u = to_dev(rng.standard_normal(op.vector_size)).requires_grad_()
    vc = to_dev(cv).requires_grad_()
    vf = to_dev(fv).requires_grad_()
    assert torch.autograd.gradcheck(lambda a, b, c: advect(a, b, c, op), (u, vc, vf), eps=1e-6, atol=1e-7,
                                    rtol=1e-6)


@pytest.mark.parametrize("degree", [1, 2])
def test_laplace_coef_gradcheck(cuda, degree):
    import torch
    from sum_factorization_kernels_amd.autograd import laplace_coef
    nc = (2, 2, 1)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=(2.0, 2.0, 1.0), flags=sfk.FLAG_VARIABLE_COEF,
                      n_rhs=2)
    rng = np.random.default_rng(5100 + degree)
    k = to_dev(np.exp(0.5 * rng.standard_normal(op.n_cell_q))).requires_grad_()
    u = to_dev(rng.standard_normal(2 * op.vector_size)).requires_grad_()
    assert torch.autograd.gradcheck(lambda a, b: laplace_coef(a, b, op), (u, k), eps=1e-6, atol=1e-7, rtol=1e-6)


def test_advect_objective_gradient_48cube(cuda):
    """J(a) = 0.5 ||A(a) u - f||^2: <dJ/da, d> equals (J(a + eps d) - J(a - eps d)) / (2 eps)
    for 64 random directions d (J is quadratic in a while no a.n changes sign,
    so the central difference is exact up to rounding: relative tolerance 1e-6)"""
    import torch
    from sum_factorization_kernels_amd.autograd import advect
    degree, nc = 3, (48, 48, 48)
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, inv_jac=(48.0, 48.0, 48.0), flags=sfk.FLAG_VARIABLE_VELOCITY)
    cv, fv = advect_velocity_field(degree, nc)
    rng = np.random.default_rng(4)
    u = to_dev(rng.standard_normal(op.vector_size))
    f = to_dev(rng.standard_normal(op.vector_size))
    vc = to_dev(cv).requires_grad_()
    vf = to_dev(fv).requires_grad_()
    r = advect(u, vc, vf, op) - f
    J = 0.5 * (r * r).sum()
    J.backward()
    gc, gf = vc.grad.clone(), vf.grad.clone()
    # the same gradient through the C ABI directly
    op.set_velocity(cv, fv)
    lam = op.apply(u) - f
    dc, df = op.grad_velocity(u, lam)
    assert torch.equal(dc, gc) and torch.equal(df, gf)

    def objective(a_c, a_f):
        op.set_velocity_device(a_c, a_f)
        rr = op.apply(u) - f
        return 0.5 * float((rr * rr).sum())

    drng = np.random.default_rng(40)
    eps = 1e-6
    worst = 0.0
    with torch.no_grad():
        for _ in range(64):
            d_c = to_dev(drng.standard_normal(cv.size))
            d_f = to_dev(drng.standard_normal(fv.size))
            fd = (objective(vc + eps * d_c, vf + eps * d_f) - objective(vc - eps * d_c, vf - eps * d_f)) / (2 * eps)
            an = float((gc * d_c).sum() + (gf * d_f).sum())
            worst = max(worst, abs(fd - an) / abs(an))
    assert worst < 1e-6, worst


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
def test_advect_48cube_matches_oracle(cuda, prec):
    """forward, adjoint and velocity gradient at 48^3 cells, degree 3,
    against the FP64 CPU oracle on the same inputs (seed 4)"""
    import torch
    degree, nc = 3, (48, 48, 48)
    ij = (48.0, 48.0, 48.0)
    dt = torch.float64 if prec == sfk.FP64 else torch.float32
    tol = TOL_FP64 if prec == sfk.FP64 else TOL_FP32
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, precision=prec, inv_jac=ij, flags=sfk.FLAG_VARIABLE_VELOCITY)
    cv, fv = advect_velocity_field(degree, nc)
    op.set_velocity(cv, fv)
    rng = np.random.default_rng(4)
    u, f = rng.standard_normal(op.vector_size), rng.standard_normal(op.vector_size)
    au = ob.dg_advect_var(degree, nc, ij, cv, fv, u)
    assert rel_err(to_np(op.apply(to_dev(u, dt))), au) < tol
    lam = au - f
    assert rel_err(to_np(op.apply_transpose(to_dev(lam, dt))), ob.dg_advect_var_transpose(degree, nc, ij, cv, fv, lam)) < tol
    # velocity gradient in the operator's precision against the FP64 oracle
    # (relative to the largest entry: 1e-12 in FP64, 2e-5 in FP32)
    gc, gf = op.grad_velocity(to_dev(u, dt), to_dev(lam, dt))
    rc, rf = ob.dg_advect_velocity_grad(degree, nc, ij, fv, u, lam)
    assert rel_err(to_np(gc), rc) < tol
    assert rel_err(to_np(gf), rf) < tol
