"""Shared helpers of the tests: tolerances, seeded inputs, torch <-> numpy."""
import numpy as np

# max |x - ref| / max |ref| per vector
TOL_FP64 = 1e-12
TOL_FP32 = 2e-5


def rel_err(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.abs(ref).max()
    return float(np.abs(x - ref).max() / (den if den > 0 else 1.0))


def to_dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def to_np(t):
    return t.double().cpu().numpy()


def sine_metric(degree, nc, amp):
    """Cell and face metric of x = xhat + amp sin(pi xhat) sin(pi yhat)
    sin(pi zhat) (1,1,1) on the uniform mesh of [0,1]^3 (numpy restatement
    used to pin the device generator and to feed the oracle).  Layouts of
    sfk_op_set_cell_metric / sfk_op_set_face_metric."""
    import oracle_bindings as ob
    n = degree + 1
    xg = ob.gauss_points(n)
    wg = ob.gauss_weights(n)
    h = 1.0 / np.asarray(nc, dtype=np.float64)

    def jac(xh):
        # xh: (..., 3) undeformed coordinates; returns J = dx/dxhat (..., 3, 3)
        s = np.sin(np.pi * xh)
        c = np.cos(np.pi * xh)
        grad = amp * np.pi * np.stack([c[..., 0] * s[..., 1] * s[..., 2], s[..., 0] * c[..., 1] * s[..., 2],
                                       s[..., 0] * s[..., 1] * c[..., 2]], axis=-1)
        J = np.broadcast_to(np.eye(3), xh.shape[:-1] + (3, 3)).copy()
        J += grad[..., None, :]  # every row gets grad s (x_i = xhat_i + s)
        return J

    def det_inv(J):
        # closed-form 3x3 determinant and inverse (adjugate / det)
        a = [[J[..., i, j] for j in range(3)] for i in range(3)]
        adj = np.empty_like(J)  # adj[j, i] = cofactor (i, j)
        for i in range(3):
            for j in range(3):
                i1, i2, j1, j2 = (i + 1) % 3, (i + 2) % 3, (j + 1) % 3, (j + 2) % 3
                adj[..., j, i] = a[i1][j1] * a[i2][j2] - a[i1][j2] * a[i2][j1]
        det = a[0][0] * adj[..., 0, 0] + a[0][1] * adj[..., 1, 0] + a[0][2] * adj[..., 2, 0]
        return det, adj / det[..., None, None]

    ncell = int(np.prod(nc))
    iz, iy, ix = np.meshgrid(np.arange(nc[2]), np.arange(nc[1]), np.arange(nc[0]), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=-1).astype(np.float64)  # (ncell, 3)
    qz, qy, qx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    qi = np.stack([qx.ravel(), qy.ravel(), qz.ravel()], axis=-1)  # (n3, 3) point indices
    xi = xg[qi]  # (n3, 3)
    wq = np.prod(wg[qi], axis=-1)
    xh = (cells[:, None, :] + xi[None, :, :]) * h  # (ncell, n3, 3)
    J = jac(xh) * h  # dx/dxi
    det, Ji = det_inv(J)
    G = (det * wq)[..., None, None] * np.einsum("cqij,cqkj->cqik", Ji, Ji)
    cm = np.stack([G[..., 0, 0], G[..., 1, 1], G[..., 2, 2], G[..., 0, 1], G[..., 0, 2], G[..., 1, 2]], axis=1)
    cm = np.ascontiguousarray(cm.reshape(ncell, 6, n ** 3)).ravel()
    # faces
    fm = np.zeros((3, ncell, 7, n * n))
    vol = np.sum((det * wq), axis=1)  # cell volumes
    for d in range(3):
        lo, hi = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
        fa_lo, fa_hi = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
        fa_lo, fa_hi = fa_lo.ravel(), fa_hi.ravel()  # fa = hi*n + lo
        xiL = np.zeros((n * n, 3))
        xiL[:, lo] = xg[fa_lo]
        xiL[:, hi] = xg[fa_hi]
        xiU = xiL.copy()
        xiL[:, d] = 1.0
        xiU[:, d] = 0.0
        wf = wg[fa_lo] * wg[fa_hi]
        e = np.zeros(3)
        e[d] = 1.0
        up = cells.copy()
        up[:, d] = (up[:, d] + 1) % nc[d]
        JL = jac((cells[:, None, :] + xiL[None]) * h) * h
        JU = jac((up[:, None, :] + xiU[None]) * h) * h
        detL, JLi = det_inv(JL)
        JUi = det_inv(JU)[1]
        nvec = np.einsum("cqji,j->cqi", JLi, e)  # J^{-T} e_d
        nrm = np.linalg.norm(nvec, axis=-1)
        area = detL * nrm
        nunit = nvec / nrm[..., None]
        jxwf = area * wf
        aL = jxwf[..., None] * np.einsum("cqij,cqj->cqi", JLi, nunit)
        aU = jxwf[..., None] * np.einsum("cqij,cqj->cqi", JUi, nunit)
        farea = np.sum(area * wf, axis=1)
        volU = vol.reshape(nc[2], nc[1], nc[0])
        volU = np.roll(volU, -1, axis=2 - d).ravel()
        tau = (degree + 1) ** 2 * 0.5 * (farea / vol + farea / volU)
        fm[d, :, 0:3, :] = np.transpose(aL, (0, 2, 1))
        fm[d, :, 3:6, :] = np.transpose(aU, (0, 2, 1))
        fm[d, :, 6, :] = tau[:, None] * jxwf
    return cm, np.ascontiguousarray(fm).ravel()


def advect_velocity_field(degree, nc):
    """a(x) = (1, 0.5 + 0.25 sin(2 pi x), 0.25) on the uniform periodic mesh
    of [0,1]^3 at the cell Gauss points ([cell][3][q]) and as a.n on the
    upper face of every cell in each direction ([d][cell][fa], fa = hi*n +
    lo over the two in-face directions, lo the faster one)."""
    import oracle_bindings as ob
    n = degree + 1
    xg = ob.gauss_points(n)
    ncell = int(np.prod(nc))
    cx = np.arange(ncell) % nc[0]
    # x coordinate of the Gauss points along x of every cell: (ncell, n)
    x = (cx[:, None] + xg[None, :]) / nc[0]
    ay = 0.5 + 0.25 * np.sin(2 * np.pi * x)  # (ncell, n): depends on the x index only
    cell = np.empty((ncell, 3, n, n, n))  # [cell][comp][k][j][i]
    cell[:, 0] = 1.0
    cell[:, 1] = ay[:, None, None, :]
    cell[:, 2] = 0.25
    face = np.empty((3, ncell, n, n))  # [d][cell][hi][lo]
    face[0] = 1.0
    face[1] = ay[:, None, :]  # d = 1: lo = x, hi = z
    face[2] = 0.25
    return cell.ravel(), face.ravel()
