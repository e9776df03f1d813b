"""GPU parity of the operators the reference does not have (per-q metric
SIPG, device metric generator, coefficient gradient, load vector, diagonal
of the deformed SIPG) and the CG-Chebyshev solver, against the oracle and
against numpy/scipy on the same inputs."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP32, TOL_FP64, rel_err, sine_metric, to_dev, to_np

pytestmark = pytest.mark.gpu

sfk = pytest.importorskip("sum_factorization_kernels_amd")


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
@pytest.mark.parametrize("degree", range(1, 9))
def test_sipg_general_matches_oracle(cuda, degree, prec):
    import torch
    nc = (3, 4, 2) if degree <= 5 else (2, 3, 2)
    N = (degree + 1) ** 3 * int(np.prod(nc))
    cm, fm = sine_metric(degree, nc, 0.1)
    rng = np.random.default_rng(1000 + degree)
    src = rng.standard_normal(N)
    dt = torch.float64 if prec == sfk.FP64 else torch.float32
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, precision=prec, geometry=sfk.GEOM_PER_Q)
    op.set_cell_metric(cm)
    op.set_face_metric(fm)
    out = to_np(op.apply(to_dev(src, dt)))
    ref = ob.dg_sipg_general(degree, nc, cm, fm, src)
    assert rel_err(out, ref) < (TOL_FP64 if prec == sfk.FP64 else TOL_FP32)
    if prec == sfk.FP64:
        # fused Chebyshev update
        rhs, diag, tm = rng.standard_normal(N), rng.uniform(0.5, 1.5, N), rng.standard_normal(N)
        d_new, d_tm = to_dev(np.zeros(N)), to_dev(tm)
        op.cheby_step(to_dev(rhs), to_dev(src), to_dev(diag), d_new, d_tm, 0.6, 0.45)
        r_tm = 0.45 * tm + 0.6 * diag * (ref - rhs)
        assert rel_err(to_np(d_tm), r_tm) < TOL_FP64
        assert rel_err(to_np(d_new), src - r_tm) < TOL_FP64
        # c_tm == 0: tmp is not read (NaN in it must not leak)
        d_new, d_tm = to_dev(np.zeros(N)), torch.full((N,), float("nan"), dtype=dt, device="cuda")
        op.cheby_step(to_dev(rhs), to_dev(src), to_dev(diag), d_new, d_tm, 0.6, 0.0)
        r_tm = 0.6 * diag * (ref - rhs)
        assert rel_err(to_np(d_tm), r_tm) < TOL_FP64
        assert rel_err(to_np(d_new), src - r_tm) < TOL_FP64


@pytest.mark.parametrize("degree", [1, 2, 3, 5])
def test_sipg_general_diagonal(cuda, degree):
    nc = (3, 2, 2)
    N = (degree + 1) ** 3 * int(np.prod(nc))
    cm, fm = sine_metric(degree, nc, 0.1)
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    op.set_cell_metric(cm)
    op.set_face_metric(fm)
    d = to_np(op.diagonal())
    ref = np.array([ob.dg_sipg_general(degree, nc, cm, fm, np.eye(1, N, i).ravel())[i] for i in range(N)])
    assert rel_err(d, ref) < TOL_FP64


@pytest.mark.parametrize("degree", [2, 5])
def test_device_sine_metric_matches_numpy(cuda, degree):
    """the device generator of sfk_op_set_sine_deformation == the numpy
    restatement in tests/helpers.py (compared through the operator)"""
    nc = (4, 3, 4)
    N = (degree + 1) ** 3 * int(np.prod(nc))
    src = np.random.default_rng(1100 + degree).standard_normal(N)
    cm, fm = sine_metric(degree, nc, 0.1)
    a = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    a.set_cell_metric(cm)
    a.set_face_metric(fm)
    b = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    b.set_sine_deformation(0.1)
    x = to_dev(src)
    assert rel_err(to_np(b.apply(x)), to_np(a.apply(x))) < TOL_FP64


@pytest.mark.parametrize("degree", [2, 4, 8])
def test_coefficient_gradient(cuda, degree):
    nc = (2, 2, 3)
    ij = (2.0, 2.0, 3.0)
    n3 = (degree + 1) ** 3
    rng = np.random.default_rng(1200 + degree)
    k = np.exp(0.5 * rng.standard_normal(int(np.prod(nc)) * n3))
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, flags=sfk.FLAG_VARIABLE_COEF, n_rhs=2)
    op.set_coefficient(k)
    N = op.vector_size
    u, lam = rng.standard_normal(2 * N), rng.standard_normal(2 * N)
    g = to_np(op.grad_coefficient(to_dev(u), to_dev(lam)))
    ref = ob.cont_laplacian_coef_grad(degree, nc, ij, u, lam, n_rhs=2)
    assert rel_err(g, ref) < TOL_FP64


def test_rhs_constant(cuda):
    degree, nc = 3, (3, 2, 2)
    ij = (3.0, 2.0, 2.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    b = to_np(op.rhs_constant(1.0))
    # direct assembly of b_i = sum_K sum_q phi_i(x_q) JxW_q
    import dense_assembly as da
    n = degree + 1
    xg, wg = ob.gauss_points(n), ob.gauss_weights(n)
    v1, _ = da.basis_1d(degree, 1, xg)
    loc = np.einsum("iq,q->i", v1, wg)  # 1D integrals of the basis functions
    T = da.dof_triples(n)
    g = [c * degree + 1 for c in nc]
    ref = np.zeros(int(np.prod(g)))
    jxw = 1.0 / np.prod(ij)
    for c in range(int(np.prod(nc))):
        cx, cy, cz = c % nc[0], (c // nc[0]) % nc[1], c // (nc[0] * nc[1])
        gi = ((cz * degree + T[2]) * g[1] + cy * degree + T[1]) * g[0] + cx * degree + T[0]
        np.add.at(ref, gi, jxw * loc[T[0]] * loc[T[1]] * loc[T[2]])
    Zi, Yi, Xi = np.meshgrid(np.arange(g[2]), np.arange(g[1]), np.arange(g[0]), indexing="ij")
    bd = ((Xi == 0) | (Yi == 0) | (Zi == 0) | (Xi == g[0] - 1) | (Yi == g[1] - 1) | (Zi == g[2] - 1)).ravel()
    ref[bd] = 0.0
    assert rel_err(b, ref) < TOL_FP64


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("cheby_degree", [1, 2, 3, 6])
def test_cg_chebyshev_solver(cuda, mixed, cheby_degree):
    """FP64 relative residual <= 1e-10 with FP64 or FP32 preconditioner; the
    solution agrees with a dense solve of the assembled system"""
    import torch
    import dense_assembly as da
    degree, nc = 3, (3, 3, 2)
    ij = (3.0, 3.0, 2.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    sm = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, precision=sfk.FP32 if mixed else sfk.FP64)
    b = op.rhs_constant(1.0)
    x = torch.zeros_like(b)
    st = sfk.solve_cg_cheby(op, sm, b, x, rel_tol=1e-10, max_iter=500, cheby_degree=cheby_degree, lanczos_steps=12)
    assert st.final_residual <= 1e-10 * np.linalg.norm(to_np(b))
    A = da.cont_laplacian_dense(degree, nc, ij)
    interior = np.abs(np.diag(A)) > 0
    Ai = A[np.ix_(interior, interior)]
    xr = np.zeros(A.shape[0])
    xr[interior] = np.linalg.solve(Ai, to_np(b)[interior])
    assert rel_err(to_np(x)[interior], xr[interior]) < 1e-8
    # Lanczos bound above the true lambda_max(D^{-1}A)
    D = np.diag(Ai)
    lam = np.linalg.eigvalsh(Ai / np.sqrt(np.outer(D, D)))
    assert st.lambda_max >= lam.max()
    assert st.lambda_max <= 1.25 * lam.max()


@pytest.mark.parametrize("max_iter", [1, 5])
def test_cg_iterate_at_max_iter(cuda, max_iter):
    """max_iter reached: the call reports no convergence and x holds the PCG
    iterate of that many iterations.  Reference: the same flexible PCG in
    numpy on the oracle operator, preconditioned by D^{-1} (Chebyshev degree
    1 is (1/theta) D^{-1}; PCG iterates do not change when the
    preconditioner is scaled)"""
    import torch
    degree, nc = 3, (3, 3, 2)
    ij = (3.0, 3.0, 2.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    b = op.rhs_constant(1.0)
    x = torch.zeros_like(b)
    with pytest.raises(sfk.SfkError) as e:
        sfk.solve_cg_cheby(op, op, b, x, rel_tol=1e-14, max_iter=max_iter, cheby_degree=1, lanczos_steps=12)
    assert e.value.status == 7
    d = to_np(op.diagonal())
    dinv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    A = lambda v: ob.cont_laplacian(degree, nc, ij, v)
    r = to_np(b)
    xr = np.zeros_like(r)
    z = dinv * r
    p, rz = z.copy(), float(r @ z)
    for _ in range(max_iter):
        q = A(p)
        a = rz / float(p @ q)
        xr += a * p
        r = r - a * q
        z = dinv * r
        beta = -a * float(z @ q) / rz
        rz = float(z @ r)
        p = z + beta * p
    assert rel_err(to_np(x), xr) < 1e-11


def test_raw_pointer_wrappers_validate_device_tensors(cuda):
    """wrong dtype or size of a CUDA tensor is rejected before the raw
    pointer reaches the C ABI"""
    import torch
    op = sfk.Operator(sfk.CONT_LAPLACE, 2, (2, 2, 2), inv_jac=(2.0, 2.0, 2.0))
    n = op.vector_size
    b32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        sfk.solve_cg_cheby(op, op, b32, x)
    with pytest.raises(ValueError):
        sfk.solve_cg_cheby(op, op, x[:-1].contiguous(), x)
    with pytest.raises(ValueError):
        sfk.layout_to_canonical(2, (2, 2, 2), 2, b32)
    with pytest.raises(ValueError):
        sfk.vec_dot(x, x[:-1])
    with pytest.raises(ValueError):
        sfk.apply_1d(np.eye(3), 1, True, torch.zeros(8, dtype=torch.float64, device="cuda"), 3, 3, 1)
    # valid calls still go through
    assert sfk.vec_dot(x + 1.0, x + 2.0) == pytest.approx(2.0 * n)
