"""Pins of the sine-deformed metric (tests/helpers.py sine_metric, the numpy
twin that the device generator sfk_op_set_sine_deformation is compared with)
by facts of the mapping itself, independent of how the twin computes them:

* the cell metric G = det(J) w J^{-1} J^{-T}, J = dx/dxi, against central
  differences of the map x(xi) with numpy's inverse and determinant;
* the volume: the map keeps the unit cube (the sine vanishes on its faces),
  so the sum of det(J) w, recovered as det(G) / w^3, is 1;
* the face metric: with J from central differences, J (JxW_f J^{-1} n) is
  the area vector n dA of a quadrature point; over the six faces of every
  cell these sum to zero (closed surface), and on the domain face
  xhat_d = 1 the map is the identity, so n dA = w h^2 e_d there;
* the penalty row: tau JxW_f / |n dA| is constant per face and equals
  (p+1)^2 / 2 (A_f / V_lower + A_f / V_upper) with A_f and V from the sums
  above: the reference's Cartesian penalty (p+1)^2 / h_d
  (evaluation_dg_laplacian.h:185,223,346-348) with 1/h_d taken as face
  area over cell volume, which it equals on the Cartesian mesh.

Each check is shown to reject a plausible convention mistake (J^T for J,
J^{-T} for J^{-1}) built here on the CPU."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import sine_metric

AMP = 0.1
FD = 1e-5


def xmap(xh, amp=AMP):
    """x = xhat + amp sin(pi xhat) sin(pi yhat) sin(pi zhat) (1, 1, 1)"""
    s = amp * np.prod(np.sin(np.pi * xh), axis=-1)
    return xh + s[..., None]


def fd_jacobian(cells, xi, h):
    """J[..., i, j] = d x_i / d xi_j by central differences of xmap, with
    xhat = (cell + xi) h; cells (c, 3), xi (q, 3) -> (c, q, 3, 3)"""
    J = np.empty(cells.shape[:1] + xi.shape[:1] + (3, 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = FD
        xp = xmap((cells[:, None, :] + xi[None] + e) * h)
        xm = xmap((cells[:, None, :] + xi[None] - e) * h)
        J[..., :, j] = (xp - xm) / (2 * FD)
    return J


def mesh(degree, nc):
    n = degree + 1
    xg, wg = ob.gauss_points(n), ob.gauss_weights(n)
    iz, iy, ix = np.meshgrid(np.arange(nc[2]), np.arange(nc[1]), np.arange(nc[0]), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=-1).astype(np.float64)
    h = 1.0 / np.asarray(nc, dtype=np.float64)
    return n, xg, wg, cells, h


def cell_points(n, xg, wg):
    qz, qy, qx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    qi = np.stack([qx.ravel(), qy.ravel(), qz.ravel()], axis=-1)
    return xg[qi], np.prod(wg[qi], axis=-1)


def face_points(n, xg, wg, d, side):
    """points of the face xi_d = side, index fa = hi*n + lo (lo the faster of
    the two in-face directions)"""
    lo, hi = [k for k in range(3) if k != d]
    fh, fl = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    fl, fh = fl.ravel(), fh.ravel()
    xi = np.zeros((n * n, 3))
    xi[:, lo] = xg[fl]
    xi[:, hi] = xg[fh]
    xi[:, d] = side
    return xi, wg[fl] * wg[fh]


def unpack_cell(cm, ncell, n):
    c = cm.reshape(ncell, 6, n ** 3)
    G = np.empty((ncell, n ** 3, 3, 3))
    for k, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        G[..., i, j] = c[:, k]
        G[..., j, i] = c[:, k]
    return G


def g_from_jacobian(J, w, transpose_mistake=False):
    Ji = np.linalg.inv(J)
    det = np.linalg.det(J)
    M = np.einsum("cqji,cqjk->cqik", Ji, Ji) if transpose_mistake else np.einsum("cqij,cqkj->cqik", Ji, Ji)
    return (det * w)[..., None, None] * M


@pytest.mark.parametrize("degree,nc", [(2, (3, 2, 2)), (4, (4, 3, 2)), (5, (2, 2, 3))])
def test_cell_metric_matches_finite_differences_of_the_map(degree, nc):
    n, xg, wg, cells, h = mesh(degree, nc)
    xi, w = cell_points(n, xg, wg)
    cm, _ = sine_metric(degree, nc, AMP)
    G = unpack_cell(cm, len(cells), n)
    J = fd_jacobian(cells, xi, h)
    ref = g_from_jacobian(J, w)
    scale = np.abs(ref).max()
    assert np.abs(G - ref).max() < 1e-8 * scale
    # the check rejects the transposed convention det J^{-T} J^{-1}
    wrong = g_from_jacobian(J, w, transpose_mistake=True)
    assert np.abs(wrong - ref).max() > 1e-3 * scale


@pytest.mark.parametrize("degree,nc", [(3, (4, 4, 4)), (5, (3, 4, 2))])
def test_cell_metric_preserves_volume(degree, nc):
    n, xg, wg, cells, h = mesh(degree, nc)
    _, w = cell_points(n, xg, wg)
    cm, _ = sine_metric(degree, nc, AMP)
    G = unpack_cell(cm, len(cells), n)
    detj_w = np.linalg.det(G) / w ** 3 * w  # det(G) = det(J) w^3
    assert detj_w.min() > 0
    assert abs(detj_w.sum() - 1.0) < 1e-9
    # and the deformation is not trivial: cell volumes differ from h^3
    assert np.abs(detj_w.reshape(len(cells), -1).sum(axis=1) - np.prod(h)).max() > 1e-3 * np.prod(h)


def area_vectors(fm, degree, nc, cells, h, n, xg, wg, mistake=False):
    """n dA per upper face q-point of every cell from both stored sides:
    lower side J_cell(xi_d = 1) aL, upper side J_upper(xi_d = 0) aU"""
    ncell = len(cells)
    f = fm.reshape(3, ncell, 7, n * n)
    out = []
    for d in range(3):
        xiL, wf = face_points(n, xg, wg, d, 1.0)
        xiU, _ = face_points(n, xg, wg, d, 0.0)
        up = cells.copy()
        up[:, d] = (up[:, d] + 1) % nc[d]
        JL = fd_jacobian(cells, xiL, h)
        JU = fd_jacobian(up, xiU, h)
        if mistake:
            JL, JU = np.swapaxes(JL, -1, -2), np.swapaxes(JU, -1, -2)
        aL = np.transpose(f[d, :, 0:3, :], (0, 2, 1))
        aU = np.transpose(f[d, :, 3:6, :], (0, 2, 1))
        AL = np.einsum("cqij,cqj->cqi", JL, aL)
        AU = np.einsum("cqij,cqj->cqi", JU, aU)
        out.append((AL, AU, f[d, :, 6, :], wf))
    return out


@pytest.mark.parametrize("degree,nc", [(3, (4, 4, 4)), (4, (3, 2, 4))])
def test_face_metric_closed_surface_and_boundary(degree, nc):
    n, xg, wg, cells, h = mesh(degree, nc)
    cm, fm = sine_metric(degree, nc, AMP)
    ncell = len(cells)
    faces = area_vectors(fm, degree, nc, cells, h, n, xg, wg)
    face_area = np.prod(h) / h.min()
    net = np.zeros((ncell, 3))
    for d, (AL, AU, _, wf) in enumerate(faces):
        # both sides of a face describe the same area vector
        assert np.abs(AL - AU).max() < 1e-9 * face_area
        # outward on the upper face, inward on the lower face of the upper neighbour
        net += AL.sum(axis=1)
        lower = np.arange(ncell)
        stride = [1, nc[0], nc[0] * nc[1]][d]
        idx = cells[:, d].astype(int)
        up = lower + np.where(idx == nc[d] - 1, -(nc[d] - 1) * stride, stride)
        np.subtract.at(net, up, AU.sum(axis=1))
        # domain face xhat_d = 1 (upper faces of the last layer): identity map
        last = idx == nc[d] - 1
        hh = np.prod(h) / h[d]
        e = np.zeros(3)
        e[d] = 1.0
        exact = (wf * hh)[None, :, None] * e[None, None, :]
        assert np.abs(AL[last] - exact).max() < 1e-9 * face_area
    assert np.abs(net).max() < 1e-8 * face_area
    # the check rejects a transposed Jacobian (J^{-T} in place of J^{-1} in
    # the stored vectors has the same effect)
    bad = area_vectors(fm, degree, nc, cells, h, n, xg, wg, mistake=True)
    worst = max(np.abs(AL - AU).max() for AL, AU, _, _ in bad)
    err_bnd = 0.0
    for d, (AL, _, _, wf) in enumerate(bad):
        last = cells[:, d].astype(int) == nc[d] - 1
        e = np.zeros(3)
        e[d] = 1.0
        err_bnd = max(err_bnd, np.abs(AL[last] - (wf * np.prod(h) / h[d])[None, :, None] * e).max())
    assert max(worst, err_bnd) > 1e-3 * face_area


@pytest.mark.parametrize("degree,nc", [(3, (4, 3, 2))])
def test_face_penalty_row(degree, nc):
    n, xg, wg, cells, h = mesh(degree, nc)
    cm, fm = sine_metric(degree, nc, AMP)
    ncell = len(cells)
    _, w = cell_points(n, xg, wg)
    vol = (np.linalg.det(unpack_cell(cm, ncell, n)) / w ** 2).sum(axis=1)
    for d, (AL, _, tj, _) in enumerate(area_vectors(fm, degree, nc, cells, h, n, xg, wg)):
        dA = np.linalg.norm(AL, axis=-1)
        tau = tj / dA
        assert np.abs(tau - tau[:, :1]).max() < 1e-8 * np.abs(tau).max()
        area = dA.sum(axis=1)
        vol_up = np.roll(vol.reshape(nc[2], nc[1], nc[0]), -1, axis=2 - d).ravel()
        expect = (degree + 1) ** 2 * 0.5 * (area / vol + area / vol_up)
        assert np.abs(tau[:, 0] - expect).max() < 1e-8 * expect.max()
