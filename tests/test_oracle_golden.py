"""Pin the C oracle against the reference's own outputs (golden fixtures
written from the compiled reference by tests/golden/make_golden.py) and its
known-answer tests.  CPU only."""
import os

import numpy as np
import pytest

import oracle_bindings as ob
from helpers import rel_err

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_quadrature_matches_reference():
    q = load("quadrature_bases.npz")
    for n in range(1, 13):
        # the reference build contracts to FMA (-ffp-contract=fast), which
        # moves the last Newton iterate by a few ulp
        assert np.abs(ob.gauss_points(n) - q[f"gauss_points_{n}"]).max() < 4e-16
        assert rel_err(ob.gauss_weights(n), q[f"gauss_weights_{n}"]) < 1e-13
        if n >= 2:
            assert np.abs(ob.gll_points(n) - q[f"gll_points_{n}"]).max() < 1e-15


def test_hermite_basis_matches_reference():
    q = load("quadrature_bases.npz")
    xs = q["hermite_x"]
    lib = ob.oracle()
    for p in range(1, 10):
        v = np.array([[lib.sfko_hermite_value(p, i, x) for x in xs] for i in range(p + 1)])
        d = np.array([[lib.sfko_hermite_derivative(p, i, x) for x in xs] for i in range(p + 1)])
        assert rel_err(v, q[f"hermite_value_{p}"]) < 1e-13
        assert rel_err(d, q[f"hermite_derivative_{p}"]) < 1e-13


@pytest.mark.parametrize("degree", range(1, 9))
def test_eo_tables_match_reference(degree):
    t = load("tables.npz")
    for kind, basis in ((0, 0), (1, 0), (2, 1)):
        mine = ob.eo_tables(degree, basis)
        keys = ["values_eo", "gradients_eo"] + (["face_eo"] if kind > 0 else []) + (
            ["inv_values_eo"] if kind == 2 else [])
        for k in keys:
            assert rel_err(mine[k], t[f"{kind}_{degree}_{k}"]) < 1e-13, (kind, k)
        if kind < 2:
            assert abs(mine["hermite_d"] - t[f"{kind}_{degree}_hermite_d"][0]) < 1e-12 * abs(mine["hermite_d"])


@pytest.mark.parametrize("nn", range(2, 10))
def test_kernel1d_matches_reference(nn):
    g = load("kernel1d.npz")
    tb = ob.eo_tables(nn - 1, 0)
    for vi, (typ, cor, add, dg) in enumerate(g["variants"]):
        for stride in (1, nn):
            seed = 1000 + 100 * nn + 10 * vi + stride
            rng = np.random.default_rng(seed)
            ln = stride * (nn - 1) + 1
            inp, out, addv, face = (rng.standard_normal(ln), rng.standard_normal(ln), rng.standard_normal(ln),
                                    rng.standard_normal(4))
            coeffs = tb["values_eo"] if typ == 0 else tb["gradients_eo"]
            # the reference tables are the Hermite ones (kind 1); use them
            t = load("tables.npz")
            coeffs = t[f"1_{nn - 1}_values_eo"] if typ == 0 else t[f"1_{nn - 1}_gradients_eo"]
            ob.apply_1d_eo(nn, stride, int(typ), int(cor), int(add), int(dg), coeffs, inp, out, addv,
                           t[f"1_{nn - 1}_face_eo"], face)
            assert rel_err(out, g[f"{nn}_{vi}_{stride}_out"]) < 1e-14
            if dg:
                assert rel_err(face, g[f"{nn}_{vi}_{stride}_face"]) < 1e-14


@pytest.mark.parametrize("degree", [1, 3, 4, 5, 6, 7, 8, 10, 12, 15])
def test_cell_laplacian_matches_reference(degree):
    g = load(f"cell_laplacian_p{degree}.npz")
    n3 = (degree + 1) ** 3
    nc = g["n_cells"]
    src = np.random.default_rng(int(g["seed"])).standard_normal(int(nc[0]) * n3)
    out = ob.cell_laplacian(degree, 0, src, g["inv_jac"])
    assert rel_err(out, g["out"]) < 1e-13


def _dg_inputs(g, degree):
    nc = [int(x) for x in g["n_cells"]]
    N = (degree + 1) ** 3 * int(np.prod(nc))
    rng = np.random.default_rng(int(g["seed"]))
    return nc, rng.standard_normal(N), rng.standard_normal(N), rng.uniform(0.5, 1.5, N), rng.standard_normal(N)


@pytest.mark.parametrize("degree", [1, 3, 4, 5, 6, 7, 8, 10, 12])
def test_dg_sipg_reference_mode_matches_reference(degree):
    """With the reference's last x sweep (ref_x_sweep=1) the oracle equals the
    compiled reference on every cell whose +x neighbour it reads correctly."""
    g = load(f"dg_laplacian_p{degree}.npz")
    nc, src, rhs, diag, tm = _dg_inputs(g, degree)
    n3 = (degree + 1) ** 3
    mask = np.repeat(g["mask"].astype(bool), n3)
    out = np.zeros_like(src)
    ob.oracle().sfko_dg_sipg_ex(degree, ob._i3(nc), ob._d3(g["inv_jac"]), ob._ptr(src), ob._ptr(out), 1)
    assert rel_err(out[mask], g["matvec"][mask]) < 1e-13
    new, tmo = np.zeros_like(src), tm.copy()
    ob.oracle().sfko_dg_sipg_cheby_ex(degree, ob._i3(nc), ob._d3(g["inv_jac"]), ob._ptr(rhs), ob._ptr(src),
                                      ob._ptr(diag), ob._ptr(new), ob._ptr(tmo), ob.ctypes.c_double(0.7),
                                      ob.ctypes.c_double(0.3), 1)
    assert rel_err(new[mask], g["cheby_new"][mask]) < 1e-13
    assert rel_err(tmo[mask], g["cheby_tm"][mask]) < 1e-13
    # the masked cells (global x index nx-1) are exactly where the
    # reference's periodic x wrap reads a wrong address
    assert rel_err(out[~mask], g["matvec"][~mask]) > 1e-3


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5, 6, 7, 8, 11, 14])
def test_dg_advect_matches_reference(degree):
    g = load(f"dg_advect_p{degree}.npz")
    nc = [int(x) for x in g["n_cells"]]
    N = (degree + 1) ** 3 * int(np.prod(nc))
    rng = np.random.default_rng(int(g["seed"]))
    src, upd, tm = rng.standard_normal(N), rng.standard_normal(N), rng.standard_normal(N)
    out = ob.dg_advect(degree, nc, g["inv_jac"], g["velocity"], src)
    # upwinding with a_x > 0 masks the reference's x-wrap defect: all cells
    assert rel_err(out, g["matvec"]) < 1e-13
    vnp, vtm = ob.dg_advect_stage(degree, nc, g["inv_jac"], g["velocity"], src, upd, tm, float(g["c_np"]),
                                  float(g["c_tm"]))
    assert rel_err(vnp, g["stage_np"]) < 1e-13
    assert rel_err(vtm, g["stage_tm"]) < 1e-13


def exact_solution(x, y, z, t, deriv=False):
    """ExactSolution of evaluation_dg_advect.h:37-92 (wave number 4,
    velocity (1, 0.15, -0.05))."""
    a = (1.0, 0.15, -0.05)
    k = 4.0 * np.pi
    px, py, pz = x - t * a[0], y - t * a[1], z - t * a[2]
    if not deriv:
        return np.sin(k * px) * np.cos(k * py) * np.cos(k * pz)
    r = -a[0] * k * np.cos(k * px) * np.cos(k * py) * np.cos(k * pz)
    add = k * np.sin(k * px) * (a[1] * np.sin(k * py) * np.cos(k * pz) + a[2] * np.cos(k * py) * np.sin(k * pz))
    return r + add


def gll_nodal(degree, nc, ij, fn):
    n = degree + 1
    gl = ob.gll_points(n)
    iz, iy, ix = np.meshgrid(np.arange(nc[2]), np.arange(nc[1]), np.arange(nc[0]), indexing="ij")
    lz, ly, lx = np.meshgrid(gl, gl, gl, indexing="ij")
    X = (ix.ravel()[:, None] + lx.ravel()[None, :]) / ij[0]
    Y = (iy.ravel()[:, None] + ly.ravel()[None, :]) / ij[1]
    Z = (iz.ravel()[:, None] + lz.ravel()[None, :]) / ij[2]
    return fn(X, Y, Z).ravel()


def _kat(degree):
    return load("advect_kat.npz" if degree < 8 else "advect_kat_high.npz")


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5, 6, 7, 10, 14])
def test_advect_derivative_kat(degree):
    """max |du/dt_exact - M^{-1}L(u_exact)| at the GLL nodes equals the
    reference's verify_derivative value (evaluation_dg_advect.h:221-251)."""
    k = _kat(degree)
    nc = [int(x) for x in k["n_cells"]]
    ij = k[f"inv_jac_p{degree}"]
    u = gll_nodal(degree, nc, ij, lambda x, y, z: exact_solution(x, y, z, 0.0))
    dudt = gll_nodal(degree, nc, ij, lambda x, y, z: exact_solution(x, y, z, 0.0, deriv=True))
    r = ob.dg_advect(degree, nc, ij, (1.0, 0.15, -0.05), u)
    err = np.abs(dudt - r).max()
    assert abs(err - k[f"p{degree}"][0]) <= 1e-10 * max(1.0, err)
