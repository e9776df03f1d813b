"""Degrees 9..15 (the reference's drivers run the DG Laplacian to degree 12,
advection to 14 and the cell Laplacian beyond, legacy/src/test_dg_integrate.cc
:22-23, test_dg_advect.cc:18-19, test_program.cc:24-25): every kernel family
against the oracle on small meshes.  The continuous operator runs on the cell
kernel there (the tensor-product kernel stops at degree 8); the golden tests
compare degrees 10-15 with the reference's own outputs."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP32, TOL_FP64, rel_err, sine_metric, to_dev, to_np

pytestmark = pytest.mark.gpu
sfk = pytest.importorskip("sum_factorization_kernels_amd")


@pytest.mark.parametrize("degree", [9, 11, 13, 15])
@pytest.mark.parametrize("basis", [0, 1])
def test_cell_laplacian_high_degree(cuda, degree, basis):
    n3 = (degree + 1) ** 3
    ij = (4.0, 2.0, 4.0 / 3.0)
    src = np.random.default_rng(7000 + degree).standard_normal(3 * n3)
    op = sfk.Operator(sfk.CELL_LAPLACE, degree, (3, 1, 1), basis=basis, inv_jac=ij)
    assert rel_err(to_np(op.apply(to_dev(src))), ob.cell_laplacian(degree, basis, src, ij)) < TOL_FP64


@pytest.mark.parametrize("degree", [9, 12, 15])
@pytest.mark.parametrize("coef", [False, True])
def test_continuous_high_degree(cuda, degree, coef):
    nc = (2, 2, 1)
    ij = (2.0, 2.0, 1.0)
    rng = np.random.default_rng(7100 + degree)
    flags = sfk.FLAG_VARIABLE_COEF if coef else 0
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, flags=flags)
    k = np.exp(0.5 * rng.standard_normal(op.n_cell_q)) if coef else None
    if coef:
        op.set_coefficient(k)
    src = rng.uniform(-1, 1, op.vector_size)
    ref = ob.cont_laplacian(degree, nc, ij, src, coef=k)
    assert rel_err(to_np(op.apply(to_dev(src))), ref) < TOL_FP64
    if coef:
        lam = rng.standard_normal(op.vector_size)
        g = op.grad_coefficient(to_dev(src), to_dev(lam))
        assert rel_err(to_np(g), ob.cont_laplacian_coef_grad(degree, nc, ij, src, lam)) < TOL_FP64
    op32 = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, precision=sfk.FP32, flags=flags)
    if coef:
        op32.set_coefficient(k)
    import torch
    s32 = src.astype(np.float32).astype(np.float64)
    ref32 = ob.cont_laplacian(degree, nc, ij, s32, coef=k)
    assert rel_err(to_np(op32.apply(to_dev(s32, torch.float32))), ref32) < TOL_FP32


@pytest.mark.parametrize("degree", [9, 12])
def test_dg_sipg_high_degree(cuda, degree):
    nc = (2, 2, 3)
    ij = (2.0, 2.0, 3.0)
    rng = np.random.default_rng(7200 + degree)
    N = (degree + 1) ** 3 * 12
    src, rhs, diag, tm = rng.standard_normal(N), rng.standard_normal(N), rng.uniform(0.5, 1.5, N), \
        rng.standard_normal(N)
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, inv_jac=ij)
    assert rel_err(to_np(op.apply(to_dev(src))), ob.dg_sipg(degree, nc, ij, src)) < TOL_FP64
    new_ref, tm_ref = ob.dg_sipg_cheby(degree, nc, ij, rhs, src, diag, tm, 0.6, 0.4)
    d_new, d_tm = to_dev(np.zeros(N)), to_dev(tm)
    op.cheby_step(to_dev(rhs), to_dev(src), to_dev(diag), d_new, d_tm, 0.6, 0.4)
    assert rel_err(to_np(d_new), new_ref) < TOL_FP64
    assert rel_err(to_np(d_tm), tm_ref) < TOL_FP64


@pytest.mark.parametrize("degree", [9, 12])
def test_dg_sipg_per_q_high_degree(cuda, degree):
    nc = (2, 2, 2)
    cm, fm = sine_metric(degree, nc, 0.1)
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    op.set_cell_metric(cm)
    op.set_face_metric(fm)
    src = np.random.default_rng(7300 + degree).standard_normal(op.vector_size)
    assert rel_err(to_np(op.apply(to_dev(src))), ob.dg_sipg_general(degree, nc, cm, fm, src)) < TOL_FP64
    # the device metric generator agrees with the numpy one
    op2 = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    op2.set_sine_deformation(0.1)
    assert rel_err(to_np(op2.apply(to_dev(src))), to_np(op.apply(to_dev(src)))) < 1e-11


@pytest.mark.parametrize("degree", [9, 13])
def test_dg_advect_variable_velocity_high_degree(cuda, degree):
    nc = (2, 3, 2)
    ij = (2.0, 3.0, 2.0)
    n = degree + 1
    ncell = int(np.prod(nc))
    rng = np.random.default_rng(7400 + degree)
    cv = (1.0 + 0.2 * rng.standard_normal((ncell, 3, n ** 3))).ravel()
    fv = (0.5 + 0.2 * rng.standard_normal((3, ncell, n * n))).ravel()
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, inv_jac=ij, flags=sfk.FLAG_VARIABLE_VELOCITY)
    op.set_velocity(cv, fv)
    u = rng.standard_normal(op.vector_size)
    assert rel_err(to_np(op.apply(to_dev(u))), ob.dg_advect_var(degree, nc, ij, cv, fv, u)) < TOL_FP64
    assert rel_err(to_np(op.apply_transpose(to_dev(u))), ob.dg_advect_var_transpose(degree, nc, ij, cv, fv, u)) \
        < TOL_FP64


def test_solver_high_degree(cuda):
    """the mixed-precision solver on the cell-kernel continuous operator"""
    import torch
    degree, nc = 10, (2, 2, 2)
    ij = (2.0, 2.0, 2.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    sm = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, precision=sfk.FP32)
    b = op.rhs_constant(1.0)
    x = torch.zeros_like(b)
    st = sfk.solve_cg_cheby(op, sm, b, x, rel_tol=1e-10, max_iter=1000, cheby_degree=6, lanczos_steps=12)
    assert st.final_residual <= 1e-10 * float(torch.linalg.norm(b))
    r = to_np(b) - ob.cont_laplacian(degree, nc, ij, to_np(x))
    assert np.linalg.norm(r) <= 1.01e-10 * float(torch.linalg.norm(b))
