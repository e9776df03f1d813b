"""Brute-force dense assembly of the operators on tiny meshes (numpy).

An independent second oracle: basis values and derivatives are evaluated
directly at the quadrature and face points (no sum factorization, no
collocation derivative, no extrapolation from Gauss points), element and face
matrices are formed explicitly and assembled into a dense global matrix.  Only
the 1D basis functions themselves (pinned against the reference by
test_oracle_golden.py) and the quadrature rule come from the oracle.
"""
import itertools

import numpy as np

import oracle_bindings as ob


def basis_1d(degree, basis, x):
    """phi_i(x), phi_i'(x) for i = 0..p at points x (arrays (p+1, len(x)))."""
    n = degree + 1
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    lib = ob.oracle()
    if basis == 0:
        v = np.array([[lib.sfko_hermite_value(degree, i, t) for t in x] for i in range(n)])
        d = np.array([[lib.sfko_hermite_derivative(degree, i, t) for t in x] for i in range(n)])
    else:
        pts = ob.gll_points(n)
        pp = pts.ctypes.data_as(ob._dp)
        v = np.array([[lib.sfko_lagrange_value(n, pp, i, t) for t in x] for i in range(n)])
        d = np.array([[lib.sfko_lagrange_derivative(n, pp, i, t) for t in x] for i in range(n)])
    return v, d


def cell_id(nc, ix, iy, iz):
    return (iz * nc[1] + iy) * nc[0] + ix


def _tensor_grad(v1, d1, ii, qq):
    """reference gradient of phi_(i0,i1,i2) at points (q0,q1,q2): (3,) arrays
    over the flattened (dof, point) grid; ii, qq: (3,) index arrays"""
    return np.stack([d1[ii[0], qq[0]] * v1[ii[1], qq[1]] * v1[ii[2], qq[2]],
                     v1[ii[0], qq[0]] * d1[ii[1], qq[1]] * v1[ii[2], qq[2]],
                     v1[ii[0], qq[0]] * v1[ii[1], qq[1]] * d1[ii[2], qq[2]]])


def dof_triples(n):
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()])  # dof = (k n + j) n + i


def sipg_general_dense(degree, nc, cell_metric, face_metric):
    """Dense DG-SIPG matrix for the per-q metric layouts of include/sfk.h,
    periodic mesh, Hermite-like basis."""
    n = degree + 1
    n2, n3 = n * n, n ** 3
    ncell = int(np.prod(nc))
    xg = ob.gauss_points(n)
    v1, d1 = basis_1d(degree, 0, xg)
    f0v, f0d = basis_1d(degree, 0, [0.0])
    f1v, f1d = basis_1d(degree, 0, [1.0])
    T = dof_triples(n)
    A = np.zeros((ncell * n3, ncell * n3))
    # cell terms: grad phi_i at all q: (3, n3 dofs, n3 q)
    gi = np.zeros((3, n3, n3))
    for q in range(n3):
        qq = T[:, q][:, None].repeat(n3, axis=1)
        gi[:, :, q] = _tensor_grad(v1, d1, T, qq)
    cm = cell_metric.reshape(ncell, 6, n3)
    for c in range(ncell):
        G = np.zeros((n3, 3, 3))
        G[:, 0, 0], G[:, 1, 1], G[:, 2, 2] = cm[c, 0], cm[c, 1], cm[c, 2]
        G[:, 0, 1] = G[:, 1, 0] = cm[c, 3]
        G[:, 0, 2] = G[:, 2, 0] = cm[c, 4]
        G[:, 1, 2] = G[:, 2, 1] = cm[c, 5]
        K = np.einsum("aiq,qab,bjq->ij", gi, G, gi)
        A[c * n3:(c + 1) * n3, c * n3:(c + 1) * n3] += K
    # face terms
    fm = face_metric.reshape(3, ncell, 7, n2)
    for d in range(3):
        lo, hi = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
        for c in range(ncell):
            x = [c % nc[0], (c // nc[0]) % nc[1], c // (nc[0] * nc[1])]
            xu = list(x)
            xu[d] = (xu[d] + 1) % nc[d]
            u = cell_id(nc, *xu)
            aL, aU = fm[d, c, 0:3], fm[d, c, 3:6]
            sigma = fm[d, c, 6]
            for fa in range(n2):
                a_lo, a_hi = fa % n, fa // n
                # traces and reference gradients of the L (at xi_d=1) and U (xi_d=0) bases
                def evaluate(side):
                    vv = np.ones(n3)
                    gg = np.ones((3, n3))
                    for e in range(3):
                        if e == d:
                            val = (f1v if side == "L" else f0v)[T[e], 0]
                            der = (f1d if side == "L" else f0d)[T[e], 0]
                        else:
                            qi = a_lo if e == lo else a_hi
                            val = v1[T[e], qi]
                            der = d1[T[e], qi]
                        vv = vv * val
                        for comp in range(3):
                            gg[comp] = gg[comp] * (der if comp == e else val)
                    return vv, gg
                vL, gL = evaluate("L")
                vU, gU = evaluate("U")
                dnL = aL[:, fa] @ gL  # JxW_f dphi/dn for L dofs (n from L to U)
                dnU = aU[:, fa] @ gU
                jump = np.concatenate([vL, -vU])
                avg = 0.5 * np.concatenate([dnL, dnU])
                blk = -np.outer(jump, avg) - np.outer(avg, jump) + sigma[fa] * np.outer(jump, jump)
                idx = np.concatenate([np.arange(c * n3, (c + 1) * n3), np.arange(u * n3, (u + 1) * n3)])
                A[np.ix_(idx, idx)] += blk
    return A


def cont_laplacian_dense(degree, nc, inv_jac, coef=None):
    """Dense continuous Q_p Laplacian (GLL basis, Gauss quadrature), Dirichlet
    rows and columns zero."""
    n = degree + 1
    n3 = n ** 3
    g = [c * degree + 1 for c in nc]
    N = int(np.prod(g))
    xg = ob.gauss_points(n)
    wg = ob.gauss_weights(n)
    v1, d1 = basis_1d(degree, 1, xg)
    T = dof_triples(n)
    gi = np.zeros((3, n3, n3))
    wq = np.zeros(n3)
    for q in range(n3):
        qq = T[:, q][:, None].repeat(n3, axis=1)
        gi[:, :, q] = _tensor_grad(v1, d1, T, qq)
        wq[q] = wg[T[0, q]] * wg[T[1, q]] * wg[T[2, q]]
    jxw = 1.0 / np.prod(inv_jac)
    A = np.zeros((N, N))
    ncell = int(np.prod(nc))
    for c in range(ncell):
        cx, cy, cz = c % nc[0], (c // nc[0]) % nc[1], c // (nc[0] * nc[1])
        k = coef[c * n3:(c + 1) * n3] if coef is not None else np.ones(n3)
        K = np.zeros((n3, n3))
        for d in range(3):
            K += np.einsum("iq,q,jq->ij", gi[d], jxw * inv_jac[d] ** 2 * wq * k, gi[d])
        gidx = ((cz * degree + T[2]) * g[1] + cy * degree + T[1]) * g[0] + cx * degree + T[0]
        A[np.ix_(gidx, gidx)] += K
    X, Y, Z = np.meshgrid(np.arange(g[0]), np.arange(g[1]), np.arange(g[2]), indexing="xy")
    Zi, Yi, Xi = np.meshgrid(np.arange(g[2]), np.arange(g[1]), np.arange(g[0]), indexing="ij")
    bd = ((Xi == 0) | (Yi == 0) | (Zi == 0) | (Xi == g[0] - 1) | (Yi == g[1] - 1) | (Zi == g[2] - 1)).ravel()
    A[bd, :] = 0.0
    A[:, bd] = 0.0
    return A


def dense_from_apply(apply, N):
    """Dense matrix of a linear map given as a function on vectors."""
    A = np.zeros((N, N))
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        A[:, i] = apply(e)
    return A


def dg_precomputed_dense(degree, nc, inv_jac):
    """Dense periodic DG-SIPG matrix of the Hermite-like basis on Cartesian
    cells of size h_d = 1/inv_jac_d: volume term grad u . grad v and, on every
    face with unit normal e_d from the lower cell (-) to the upper cell (+),
        -{du/dn}[v] - [u]{dv/dn} + sigma [u][v],
    [w] = w- - w+, {dw/dn} = (dw-/dx_d + dw+/dx_d) / 2, sigma = p(p+1)/h_d
    (the penalty of precomputed_dg_laplacian.h:310 on cells of size h_d).
    Quadrature: p+3 Gauss points per direction (more than the operator's)."""
    n = degree + 1
    n3 = n ** 3
    h = 1.0 / np.asarray(inv_jac, dtype=np.float64)
    xq, wq = np.polynomial.legendre.leggauss(n + 2)
    xq, wq = 0.5 * (xq + 1.0), 0.5 * wq
    v, d = basis_1d(degree, 0, xq)          # (n, nq)
    v01, d01 = basis_1d(degree, 0, [0.0, 1.0])  # (n, 2)
    ii = dof_triples(n)                      # (3, n3)
    ncell = int(np.prod(nc))
    A = np.zeros((ncell * n3, ncell * n3))
    det = np.prod(h)
    # element stiffness (identical for every cell)
    K = np.zeros((n3, n3))
    for dd in range(3):
        # 1D factors: derivative in dd, values in the others
        f = []
        for e in range(3):
            a = d if e == dd else v
            f.append((a[:, None, :] * a[None, :, :] * wq).sum(axis=2))  # (n, n)
        K += det / h[dd] ** 2 * f[0][ii[0][:, None], ii[0][None, :]] * f[1][ii[1][:, None], ii[1][None, :]] \
            * f[2][ii[2][:, None], ii[2][None, :]]
    for c in range(ncell):
        A[c * n3:(c + 1) * n3, c * n3:(c + 1) * n3] += K
    # faces
    sigma = degree * (degree + 1)
    for c in range(ncell):
        ix, iy, iz = c % nc[0], (c // nc[0]) % nc[1], c // (nc[0] * nc[1])
        for dd in range(3):
            up = [ix, iy, iz]
            up[dd] = (up[dd] + 1) % nc[dd]
            cp = cell_id(nc, *up)
            t = [e for e in range(3) if e != dd]
            # tangential mass factors (face quadrature over the two other directions)
            tm = [(v[:, None, :] * v[None, :, :] * wq).sum(axis=2) * h[e] for e in t]  # (n, n) each
            T = tm[0][ii[t[0]][:, None], ii[t[0]][None, :]] * tm[1][ii[t[1]][:, None], ii[t[1]][None, :]]
            # normal-direction traces: side -: x_dd = 1, side +: x_dd = 0
            val = {-1: v01[:, 1], 1: v01[:, 0]}
            der = {-1: d01[:, 1] / h[dd], 1: d01[:, 0] / h[dd]}
            cells = {-1: c, 1: cp}
            for s in (-1, 1):      # test side
                for r in (-1, 1):  # trial side
                    js, jr = (1.0 if s == -1 else -1.0), (1.0 if r == -1 else -1.0)
                    vs, ds = val[s][ii[dd]], der[s][ii[dd]]
                    vr, dr = val[r][ii[dd]], der[r][ii[dd]]
                    loc = (-0.5 * dr[None, :] * js * vs[:, None] - jr * vr[None, :] * 0.5 * ds[:, None]
                           + sigma / h[dd] * jr * js * vr[None, :] * vs[:, None]) * T
                    A[cells[s] * n3:(cells[s] + 1) * n3, cells[r] * n3:(cells[r] + 1) * n3] += loc
    return A
