"""GPU parity of the HIP operators (through the C ABI) against the CPU oracle
on the same seeded inputs.  FP64: max|x-ref|/max|ref| <= 1e-12 (TOL_FP64);
FP32 kernels vs the FP64 oracle: <= 2e-5 (TOL_FP32)."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP32, TOL_FP64, rel_err, to_dev, to_np

pytestmark = pytest.mark.gpu

sfk = pytest.importorskip("sum_factorization_kernels_amd")


def _dtype(prec):
    import torch
    return torch.float64 if prec == sfk.FP64 else torch.float32


def _tol(prec):
    return TOL_FP64 if prec == sfk.FP64 else TOL_FP32


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
@pytest.mark.parametrize("basis", [sfk.BASIS_HERMITE, sfk.BASIS_GLL])
@pytest.mark.parametrize("degree", range(1, 9))
def test_cell_laplacian(cuda, degree, basis, prec):
    ncell = 37  # not a multiple of the cells per workgroup
    ij = (4.0, 2.0, 4.0 / 3.0)  # evaluation_cell_laplacian.h:803-813
    rng = np.random.default_rng(100 + degree)
    src = rng.standard_normal(ncell * (degree + 1) ** 3)
    op = sfk.Operator(sfk.CELL_LAPLACE, degree, (ncell, 1, 1), precision=prec, basis=basis, inv_jac=ij)
    out = to_np(op.apply(to_dev(src, _dtype(prec))))
    ref = ob.cell_laplacian(degree, basis, src, ij)
    assert rel_err(out, ref) < _tol(prec)


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
@pytest.mark.parametrize("refx", [False, True])
@pytest.mark.parametrize("degree", range(1, 9))
def test_dg_sipg_cartesian(cuda, degree, refx, prec):
    nc = (5, 3, 4)
    ij = (4.0, 2.0, 4.0 / 3.0)
    n3 = (degree + 1) ** 3
    N = n3 * int(np.prod(nc))
    rng = np.random.default_rng(200 + degree)
    src = rng.standard_normal(N)
    flags = sfk.FLAG_REF_X_SWEEP if refx else 0
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, precision=prec, inv_jac=ij, flags=flags)
    out = to_np(op.apply(to_dev(src, _dtype(prec))))
    ref = np.zeros(N)
    ob.oracle().sfko_dg_sipg_ex(degree, ob._i3(nc), ob._d3(ij), ob._ptr(src), ob._ptr(ref), int(refx))
    assert rel_err(out, ref) < _tol(prec)
    # fused Chebyshev update (evaluation_dg_laplacian.h:485-502)
    rhs, diag, tm = rng.standard_normal(N), rng.uniform(0.5, 1.5, N), rng.standard_normal(N)
    dt = _dtype(prec)
    d_new, d_tm = to_dev(np.zeros(N), dt), to_dev(tm, dt)
    op.cheby_step(to_dev(rhs, dt), to_dev(src, dt), to_dev(diag, dt), d_new, d_tm, 0.7, 0.3)
    r_new, r_tm = np.zeros(N), tm.copy()
    ob.oracle().sfko_dg_sipg_cheby_ex(degree, ob._i3(nc), ob._d3(ij), ob._ptr(rhs), ob._ptr(src), ob._ptr(diag),
                                      ob._ptr(r_new), ob._ptr(r_tm), ob.ctypes.c_double(0.7),
                                      ob.ctypes.c_double(0.3), int(refx))
    assert rel_err(to_np(d_new), r_new) < _tol(prec)
    assert rel_err(to_np(d_tm), r_tm) < _tol(prec)
    # c_tm == 0 starts a recurrence: tmp is only written (NaN in it must not leak)
    import torch
    d_new, d_tm = to_dev(np.zeros(N), dt), torch.full((N,), float("nan"), dtype=dt, device="cuda")
    op.cheby_step(to_dev(rhs, dt), to_dev(src, dt), to_dev(diag, dt), d_new, d_tm, 0.7, 0.0)
    r_new, r_tm = np.zeros(N), np.zeros(N)
    ob.oracle().sfko_dg_sipg_cheby_ex(degree, ob._i3(nc), ob._d3(ij), ob._ptr(rhs), ob._ptr(src), ob._ptr(diag),
                                      ob._ptr(r_new), ob._ptr(r_tm), ob.ctypes.c_double(0.7),
                                      ob.ctypes.c_double(0.0), int(refx))
    assert rel_err(to_np(d_new), r_new) < _tol(prec)
    assert rel_err(to_np(d_tm), r_tm) < _tol(prec)


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
@pytest.mark.parametrize("degree", range(1, 9))
def test_dg_advect(cuda, degree, prec):
    nc = (6, 3, 4)
    ij = (3.0, 1.5, 2.0)
    vel = (1.0, 0.15, -0.05)  # evaluation_dg_advect.h:46-50
    N = (degree + 1) ** 3 * int(np.prod(nc))
    rng = np.random.default_rng(300 + degree)
    src, upd, tm = rng.standard_normal(N), rng.standard_normal(N), rng.standard_normal(N)
    dt = _dtype(prec)
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, precision=prec, inv_jac=ij, velocity=vel)
    out = to_np(op.apply(to_dev(src, dt)))
    assert rel_err(out, ob.dg_advect(degree, nc, ij, vel, src)) < _tol(prec)
    # fused LSRK stage (evaluation_dg_advect.h:651-667)
    d_np, d_tm = to_dev(np.zeros(N), dt), to_dev(tm, dt)
    op.rk_stage(to_dev(src, dt), to_dev(upd, dt), d_np, d_tm, 0.3, 0.2)
    r_np, r_tm = ob.dg_advect_stage(degree, nc, ij, vel, src, upd, tm, 0.3, 0.2)
    assert rel_err(to_np(d_np), r_np) < _tol(prec)
    assert rel_err(to_np(d_tm), r_tm) < _tol(prec)
    # c_tm == 0: vec_tm untouched
    d_tm2 = to_dev(tm, dt)
    op.rk_stage(to_dev(src, dt), to_dev(upd, dt), d_np, d_tm2, 0.3, 0.0)
    assert np.array_equal(to_np(d_tm2), to_np(to_dev(tm, dt)))


def _var_velocity(degree, nc, rng):
    n = degree + 1
    ncell = int(np.prod(nc))
    cv = rng.uniform(0.2, 1.0, (ncell, 3, n ** 3))
    cv[:, 1, :] *= -1.0
    fv = rng.uniform(-1.0, 1.0, 3 * ncell * n * n)
    return np.ascontiguousarray(cv).ravel(), fv


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
@pytest.mark.parametrize("degree", [1, 2, 3, 4, 6])
def test_dg_advect_variable_velocity_and_adjoint(cuda, degree, prec):
    nc = (4, 3, 5)
    ij = (4.0, 3.0, 5.0)
    N = (degree + 1) ** 3 * int(np.prod(nc))
    rng = np.random.default_rng(400 + degree)
    cv, fv = _var_velocity(degree, nc, rng)
    u, w = rng.standard_normal(N), rng.standard_normal(N)
    dt = _dtype(prec)
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, precision=prec, inv_jac=ij, flags=sfk.FLAG_VARIABLE_VELOCITY)
    op.set_velocity(cv, fv)
    out = to_np(op.apply(to_dev(u, dt)))
    assert rel_err(out, ob.dg_advect_var(degree, nc, ij, cv, fv, u)) < _tol(prec)
    outT = to_np(op.apply_transpose(to_dev(w, dt)))
    assert rel_err(outT, ob.dg_advect_var_transpose(degree, nc, ij, cv, fv, w)) < _tol(prec)
    if prec == sfk.FP64:
        # adjoint identity <A u, w> = <u, A^T w>
        assert abs(out @ w - u @ outT) <= 1e-12 * abs(out) @ abs(w)


@pytest.mark.parametrize("degree", [1, 2, 3, 5])
def test_dg_advect_velocity_gradient(cuda, degree):
    nc = (3, 4, 3)
    ij = (3.0, 4.0, 3.0)
    N = (degree + 1) ** 3 * int(np.prod(nc))
    rng = np.random.default_rng(500 + degree)
    cv, fv = _var_velocity(degree, nc, rng)
    u, f = rng.standard_normal(N), rng.standard_normal(N)
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, inv_jac=ij, flags=sfk.FLAG_VARIABLE_VELOCITY)
    op.set_velocity(cv, fv)
    lam = ob.dg_advect_var(degree, nc, ij, cv, fv, u) - f
    gc, gf = op.grad_velocity(to_dev(u), to_dev(lam))
    rc, rf = ob.dg_advect_velocity_grad(degree, nc, ij, fv, u, lam)
    assert rel_err(to_np(gc), rc) < TOL_FP64
    assert rel_err(to_np(gf), rf) < TOL_FP64


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
@pytest.mark.parametrize("degree", range(1, 9))
def test_cont_laplacian(cuda, degree, prec):
    nc = (3, 2, 4)
    h = 1.0 / np.asarray(nc, dtype=np.float64)
    ij = tuple(1.0 / h)
    g = [c * degree + 1 for c in nc]
    N = int(np.prod(g))
    rng = np.random.default_rng(600 + degree)
    src = rng.uniform(-1, 1, N)
    dt = _dtype(prec)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, precision=prec, inv_jac=ij)
    assert op.vector_size == N
    out = to_np(op.apply(to_dev(src, dt)))
    assert rel_err(out, ob.cont_laplacian(degree, nc, ij, src)) < _tol(prec)


@pytest.mark.parametrize("degree", [2, 4, 8])
def test_cont_laplacian_variable_coefficient_multi_rhs(cuda, degree):
    nc = (2, 3, 2)
    ij = (2.0, 3.0, 2.0)
    n3 = (degree + 1) ** 3
    rng = np.random.default_rng(700 + degree)
    k = np.exp(0.5 * rng.standard_normal(int(np.prod(nc)) * n3))
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, flags=sfk.FLAG_VARIABLE_COEF, n_rhs=3)
    op.set_coefficient(k)
    N = op.vector_size
    src = rng.standard_normal(3 * N)
    out = to_np(op.apply(to_dev(src)))
    for r in range(3):
        ref = ob.cont_laplacian(degree, nc, ij, src[r * N:(r + 1) * N].copy(), coef=k)
        assert rel_err(out[r * N:(r + 1) * N], ref) < TOL_FP64


@pytest.mark.parametrize("kind", [sfk.CELL_LAPLACE, sfk.CONT_LAPLACE])
@pytest.mark.parametrize("degree", [1, 3, 4, 6])
def test_laplacian_diagonal(cuda, kind, degree):
    nc = (2, 2, 3) if kind == sfk.CONT_LAPLACE else (5, 1, 1)
    ij = (2.0, 2.0, 3.0)
    basis = sfk.BASIS_GLL
    op = sfk.Operator(kind, degree, nc, inv_jac=ij, basis=basis)
    N = op.vector_size
    d = to_np(op.diagonal())
    # brute force: e_i^T A e_i with the oracle
    ref = np.zeros(N)
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        if kind == sfk.CONT_LAPLACE:
            ref[i] = ob.cont_laplacian(degree, nc, ij, e)[i]
        else:
            ref[i] = ob.cell_laplacian(degree, basis, e, ij)[i]
    assert rel_err(d, ref) < TOL_FP64


@pytest.mark.parametrize("lanes", [1, 4, 8])
def test_layout_roundtrip(cuda, lanes):
    degree, nc = 3, (8, 3, 2)
    N = (degree + 1) ** 3 * int(np.prod(nc))
    x = np.random.default_rng(800 + lanes).standard_normal(N)
    c = sfk.layout_to_canonical(degree, nc, lanes, to_dev(x))
    assert np.array_equal(to_np(c), ob.to_canonical(degree, nc, lanes, x))
    back = sfk.layout_to_interleaved(degree, nc, lanes, c)
    assert np.array_equal(to_np(back), x)


@pytest.mark.parametrize("nn", range(2, 10))
def test_apply_1d_kernel(cuda, nn):
    rng = np.random.default_rng(900 + nn)
    # a centro-symmetric and a centro-antisymmetric matrix
    M = rng.standard_normal((nn, nn))
    Ms = 0.5 * (M + M[::-1, ::-1])
    Ma = 0.5 * (M - M[::-1, ::-1])
    n_lines, stride = 50, 3
    x = rng.standard_normal(n_lines * nn * stride)
    for mat, sym in ((Ms, 1), (Ma, -1)):
        for fwd in (True, False):
            out = to_np(sfk.apply_1d(mat, sym, fwd, to_dev(x), n_lines, nn * stride, stride))
            xl = x.reshape(n_lines, nn, stride)[:, :, 0]
            ref = xl @ mat if fwd else xl @ mat.T
            got = out.reshape(n_lines, nn, stride)[:, :, 0]
            assert rel_err(got, ref) < 1e-14
