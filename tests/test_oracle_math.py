"""Pin the oracle where the reference has no answer (and re-pin where it
does) with what the mathematics fixes: closed-form quadrature, basis
identities, brute-force dense assembly with direct basis evaluation,
symmetry / null space / positivity, adjoint identities and central finite
differences.  CPU only, tiny meshes."""
import numpy as np
import pytest

import dense_assembly as da
import oracle_bindings as ob
from helpers import rel_err, sine_metric


@pytest.mark.parametrize("n", range(1, 13))
def test_gauss_closed_form(n):
    x, w = np.polynomial.legendre.leggauss(n)
    assert np.abs(ob.gauss_points(n) - 0.5 * (x + 1)).max() < 1e-15
    assert np.abs(ob.gauss_weights(n) - 0.5 * w).max() < 1e-14


@pytest.mark.parametrize("n", range(2, 13))
def test_gll_closed_form(n):
    # interior GLL points = roots of P'_{n-1}
    c = np.zeros(n)
    c[-1] = 1.0
    r = np.sort(np.polynomial.legendre.Legendre(c).deriv().roots().real)
    ref = np.concatenate([[0.0], 0.5 * (r + 1), [1.0]])
    assert np.abs(ob.gll_points(n) - ref).max() < 1e-14


@pytest.mark.parametrize("p", range(3, 10))
def test_hermite_like_basis_identities(p):
    """lagrange_polynomials.h:86-156: phi_0(0)=1, only phi_0, phi_1 have a
    nonzero derivative at 0 with phi_1'(0) = -phi_0'(0), nodal at the interior
    Jacobi(2,2) roots, mirror symmetry, and (phi_0, phi_1)_L2 = 0."""
    lib = ob.oracle()
    v = lambda i, x: lib.sfko_hermite_value(p, i, x)  # noqa: E731
    d = lambda i, x: lib.sfko_hermite_derivative(p, i, x)  # noqa: E731
    assert abs(v(0, 0.0) - 1) < 1e-13
    for i in range(1, p + 1):
        assert abs(v(i, 0.0)) < 1e-13
    for i in range(2, p + 1):
        assert abs(d(i, 0.0)) < 1e-10
    assert abs(d(0, 0.0) + d(1, 0.0)) < 1e-10 * abs(d(0, 0.0))
    roots = np.zeros(p - 3)
    if p > 3:
        lib.sfko_zeros_of_jacobi_polynomial(p + 1, 2, 2, roots.ctypes.data_as(ob._dp))
        xs = 0.5 * roots + 0.5
        for m, x in enumerate(xs):
            for i in range(p + 1):
                assert abs(v(i, x) - (1.0 if i == m + 2 else 0.0)) < 1e-12
    for x in (0.1, 0.37, 0.8):
        for i in range(p + 1):
            assert abs(v(i, x) - v(p - i, 1 - x)) < 1e-12
    xq, wq = np.polynomial.legendre.leggauss(p + 2)
    xq = 0.5 * (xq + 1)
    assert abs(sum(0.5 * w * v(0, x) * v(1, x) for x, w in zip(xq, wq))) < 1e-14


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("basis", [0, 1])
def test_kernel1d_eo_equals_dense(p, basis):
    """the even-odd 1D kernel restatement against dense S^T x / S x for all
    variants (matrix_vector_kernel.h:30-185); basis 0 with p=2 is the case
    where the reference's backward symmetric sweep drops a term"""
    n = p + 1
    tb = ob.eo_tables(p, basis)
    pts = ob.gauss_points(n)
    S, Dd = da.basis_1d(p, basis, pts)  # S[i][q] = phi_i(x_q); phi' not needed
    lib = ob.oracle()
    pp = pts.ctypes.data_as(ob._dp)
    D = np.array([[lib.sfko_lagrange_derivative(n, pp, i, pts[q]) for q in range(n)] for i in range(n)])
    lv = np.array([[lib.sfko_lagrange_value(n, pp, i, s) for i in range(n)] for s in (0.0, 1.0)])
    ld = np.array([[lib.sfko_lagrange_derivative(n, pp, i, s) for i in range(n)] for s in (0.0, 1.0)])
    rng = np.random.default_rng(p + 10 * basis)
    u = rng.standard_normal(n)
    a = rng.standard_normal(n)
    for typ, M in ((0, S), (1, D)):
        coeffs = tb["values_eo"] if typ == 0 else tb["gradients_eo"]
        out = np.zeros(n)
        ob.apply_1d_eo(n, 1, typ, 1, 0, 0, coeffs, u, out)
        assert rel_err(out, M.T @ u) < 1e-13
        out = np.zeros(n)
        ob.apply_1d_eo(n, 1, typ, 0, 1, 0, coeffs, u, out, a)
        if typ == 0 and basis == 0 and p == 2:
            assert rel_err(out, M @ u + a) > 1e-3  # the reference's dropped middle term
        else:
            assert rel_err(out, M @ u + a) < 1e-13
    out, face = np.zeros(n), np.zeros(4)
    ob.apply_1d_eo(n, 1, 1, 1, 0, 2, tb["gradients_eo"], u, out, None, tb["face_eo"], face)
    assert rel_err(face, [lv[0] @ u, lv[1] @ u, ld[0] @ u, ld[1] @ u]) < 1e-13
    f = rng.standard_normal(4)
    fa = np.array([f[0] + f[1], f[0] - f[1], f[2] + f[3], f[2] - f[3]])
    out = np.zeros(n)
    ob.apply_1d_eo(n, 1, 1, 0, 0, 2, tb["gradients_eo"], u, out, None, tb["face_eo"], fa)
    assert rel_err(out, D @ u + lv[0] * f[0] + lv[1] * f[1] + ld[0] * f[2] + ld[1] * f[3]) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("basis", [0, 1])
def test_cell_laplacian_kronecker(p, basis):
    """element stiffness = sum_d c_d (K1 (x) M1 (x) M1) permuted, with the 1D
    matrices from direct basis evaluation"""
    n = p + 1
    ij = (4.0, 2.0, 4.0 / 3.0)
    xg, wg = ob.gauss_points(n), ob.gauss_weights(n)
    v, d = da.basis_1d(p, basis, xg)
    M1 = (v * wg) @ v.T
    K1 = (d * wg) @ d.T
    jxw = 1.0 / np.prod(ij)
    A = (jxw * ij[0] ** 2 * np.kron(M1, np.kron(M1, K1)) + jxw * ij[1] ** 2 * np.kron(M1, np.kron(K1, M1)) +
         jxw * ij[2] ** 2 * np.kron(K1, np.kron(M1, M1)))
    x = np.random.default_rng(p).standard_normal(3 * n ** 3)
    out = ob.cell_laplacian(p, basis, x, ij)
    for c in range(3):
        assert rel_err(out[c * n ** 3:(c + 1) * n ** 3], A @ x[c * n ** 3:(c + 1) * n ** 3]) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 3])
def test_cell_laplacian_linear_function_kat(p):
    """the GLL-basis variant of the reference self-check
    (evaluation_cell_laplacian_scalar_coef.h:90-184): for u linear in x_d,
    A_K u equals the boundary integral of du/dn v"""
    n = p + 1
    ij = (4.0, 2.0, 4.0 / 3.0)
    gl = ob.gll_points(n)
    T = da.dof_triples(n)
    xg, wg = ob.gauss_points(n), ob.gauss_weights(n)
    v, _ = da.basis_1d(p, 1, xg)
    b1 = v @ wg  # 1D integrals of the basis
    vb = da.basis_1d(p, 1, [0.0, 1.0])[0]
    for test in range(3):
        u = gl[T[test]] / ij[test]  # physical coordinate x_test
        out = ob.cell_laplacian(p, 1, u, ij)
        # boundary term: +-1 on the faces normal to `test`, face JxW = prod h_e (e != test)
        other = [e for e in range(3) if e != test]
        fac = 1.0 / (ij[other[0]] * ij[other[1]])
        ref = fac * (vb[T[test], 1] - vb[T[test], 0]) * b1[T[other[0]]] * b1[T[other[1]]]
        assert np.abs(out - ref).max() < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3])
def test_sipg_cartesian_dense_assembly(p):
    nc = (3, 2, 2)
    ij = (3.0, 2.0, 2.0)
    cm, fm = ob.cartesian_metric(p, nc, ij)
    A = da.sipg_general_dense(p, nc, cm, fm)
    x = np.random.default_rng(p).standard_normal(A.shape[0])
    assert rel_err(ob.dg_sipg(p, nc, ij, x), A @ x) < 1e-13
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > -1e-10 * np.abs(A).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_sipg_deformed_dense_assembly(p):
    nc = (3, 2, 2)
    cm, fm = sine_metric(p, nc, 0.1)
    A = da.sipg_general_dense(p, nc, cm, fm)
    x = np.random.default_rng(10 + p).standard_normal(A.shape[0])
    assert rel_err(ob.dg_sipg_general(p, nc, cm, fm, x), A @ x) < 1e-13
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-12 * np.abs(A).max()


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_sipg_general_reduces_to_cartesian(p):
    nc = (4, 3, 2)
    ij = (4.0, 3.0, 2.0)
    cm, fm = ob.cartesian_metric(p, nc, ij)
    cm2, fm2 = sine_metric(p, nc, 0.0)  # identity mapping through the numpy generator
    assert np.abs(cm - cm2).max() < 1e-15 and np.abs(fm - fm2).max() < 1e-14
    x = np.random.default_rng(p).standard_normal((p + 1) ** 3 * 24)
    assert rel_err(ob.dg_sipg_general(p, nc, cm, fm, x), ob.dg_sipg(p, nc, ij, x)) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 3])
def test_continuous_laplacian_dense_assembly(p):
    nc = (2, 3, 2)
    ij = (2.0, 3.0, 2.0)
    k = np.exp(0.5 * np.random.default_rng(p).standard_normal(12 * (p + 1) ** 3))
    A = da.cont_laplacian_dense(p, nc, ij, k)
    x = np.random.default_rng(20 + p).standard_normal(A.shape[0])
    assert rel_err(ob.cont_laplacian(p, nc, ij, x, coef=k), A @ x) < 1e-13
    assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_advection_adjoint_dense(p):
    nc = (3, 2, 3)
    ij = (3.0, 2.0, 3.0)
    n = p + 1
    rng = np.random.default_rng(30 + p)
    cv = rng.uniform(-1, 1, 18 * 3 * n ** 3)
    fv = rng.uniform(-1, 1, 3 * 18 * n * n)
    N = 18 * n ** 3
    A = da.dense_from_apply(lambda e: ob.dg_advect_var(p, nc, ij, cv, fv, e), N)
    AT = da.dense_from_apply(lambda e: ob.dg_advect_var_transpose(p, nc, ij, cv, fv, e), N)
    assert np.abs(AT - A.T).max() < 1e-12 * np.abs(A).max()
    # constant velocity: the reference-derived operator
    vel = (1.0, 0.15, -0.05)
    cvc = np.repeat(np.array(vel)[None, :, None], 18, 0).repeat(n ** 3, 2).ravel()
    fvc = np.repeat(np.array(vel)[:, None], 18 * n * n, 1).ravel()
    x = rng.standard_normal(N)
    assert rel_err(ob.dg_advect_var(p, nc, ij, cvc, fvc, x), ob.dg_advect(p, nc, ij, vel, x)) < 1e-14


@pytest.mark.parametrize("p", [1, 3])
def test_advection_velocity_gradient_finite_differences(p):
    """dJ/da for J = 0.5 |A(a) u - f|^2 against central differences with
    eps = 1e-6 in 16 random directions (FP64)"""
    nc = (2, 3, 2)
    ij = (2.0, 3.0, 2.0)
    n = p + 1
    ncell = 12
    rng = np.random.default_rng(40 + p)
    cv = rng.uniform(0.2, 1.0, ncell * 3 * n ** 3)
    fv = rng.uniform(0.2, 1.0, 3 * ncell * n * n) * rng.choice([-1.0, 1.0], 3 * ncell * n * n)
    N = ncell * n ** 3
    u, f = rng.standard_normal(N), rng.standard_normal(N)

    def J(c, fa):
        r = ob.dg_advect_var(p, nc, ij, c, fa, u) - f
        return 0.5 * r @ r

    lam = ob.dg_advect_var(p, nc, ij, cv, fv, u) - f
    gc, gf = ob.dg_advect_velocity_grad(p, nc, ij, fv, u, lam)
    eps = 1e-6
    for _ in range(16):
        dc, df = rng.standard_normal(cv.size), rng.standard_normal(fv.size)
        fd = (J(cv + eps * dc, fv + eps * df) - J(cv - eps * dc, fv - eps * df)) / (2 * eps)
        an = gc @ dc + gf @ df
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an))


@pytest.mark.parametrize("p", [2, 4])
def test_coefficient_gradient_finite_differences(p):
    nc = (2, 2, 2)
    ij = (2.0, 2.0, 2.0)
    ncq = 8 * (p + 1) ** 3
    rng = np.random.default_rng(50 + p)
    k = np.exp(0.5 * rng.standard_normal(ncq))
    g = [c * p + 1 for c in nc]
    Nn = int(np.prod(g))
    u, lam = rng.standard_normal(Nn), rng.standard_normal(Nn)
    grad = ob.cont_laplacian_coef_grad(p, nc, ij, u, lam)
    eps = 1e-6
    for _ in range(8):
        dk = rng.standard_normal(ncq)
        fp = lam @ ob.cont_laplacian(p, nc, ij, u, coef=k + eps * dk)
        fm = lam @ ob.cont_laplacian(p, nc, ij, u, coef=k - eps * dk)
        assert abs((fp - fm) / (2 * eps) - grad @ dk) < 1e-7 * max(1.0, abs(grad @ dk))


def test_interleaved_layout_roundtrip():
    p, nc = 2, (8, 2, 3)
    x = np.random.default_rng(0).standard_normal(27 * 48)
    for L in (1, 2, 4, 8):
        y = ob.to_interleaved(p, nc, L, x)
        assert np.array_equal(ob.to_canonical(p, nc, L, y), x)
        # lane 0 of the first batch holds cell 0's dof 0, lane 1 holds cell 1's
        assert y[0] == x[0] and (L == 1 or y[1] == x[27])
