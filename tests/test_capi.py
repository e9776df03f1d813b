"""The C ABI library loads without a GPU, exports every entry point that
include/sfk.h declares, validates descriptors on the host, and its host-side
tables equal the oracle's (pinned against the reference).  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_bindings as ob
from helpers import rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sfk.h")

sfk = pytest.importorskip("sum_factorization_kernels_amd")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sfk.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string():
    lib = sfk.lib()
    assert b"gfx950" in lib.sfk_version()
    d = sfk.OpDesc()
    d.kind, d.degree, d.n_rhs = sfk.DG_SIPG, 99, 1
    d.n_cells[:] = [2, 2, 2]
    d.inv_jac[:] = [1, 1, 1]
    h = ctypes.c_void_p()
    assert lib.sfk_op_create(ctypes.byref(d), ctypes.byref(h)) == 1  # SFK_ERR_INVALID_ARG
    assert b"degree" in lib.sfk_last_error()


@pytest.mark.parametrize("bad", [
    dict(degree=0), dict(degree=16), dict(n_cells=(0, 1, 1)), dict(inv_jac=(1.0, -1.0, 1.0)),
    dict(inv_jac=(1.0, float("nan"), 1.0)), dict(n_rhs=0), dict(precision=7), dict(kind=9),
])
def test_descriptor_validation(bad):
    kw = dict(kind=sfk.DG_SIPG, degree=3, n_cells=(2, 2, 2), inv_jac=(2.0, 2.0, 2.0))
    kw.update(bad)
    with pytest.raises(sfk.SfkError) as e:
        sfk.Operator(**kw)
    assert e.value.status == 1


@pytest.mark.parametrize("kw,status", [
    (dict(kind=sfk.DG_SIPG, basis=sfk.BASIS_GLL), 2),
    (dict(kind=sfk.DG_ADVECT, basis=sfk.BASIS_HERMITE), 2),
    (dict(kind=sfk.DG_SIPG, n_rhs=2), 2),
    (dict(kind=sfk.DG_SIPG, geometry=sfk.GEOM_PER_Q, flags=sfk.FLAG_REF_X_SWEEP), 2),
    (dict(kind=sfk.DG_ADVECT, flags=sfk.FLAG_VARIABLE_COEF), 2),
    (dict(kind=sfk.DG_PRECOMP, basis=sfk.BASIS_GLL), 2),
    (dict(kind=sfk.DG_PRECOMP, geometry=sfk.GEOM_PER_Q), 2),
    (dict(kind=sfk.DG_SIPG, flags=sfk.FLAG_REF_PRECOMP_COUPLING), 2),
])
def test_unsupported_combinations(kw, status):
    base = dict(degree=3, n_cells=(2, 2, 2), inv_jac=(2.0, 2.0, 2.0))
    base.update(kw)
    with pytest.raises(sfk.SfkError) as e:
        sfk.Operator(**base)
    assert e.value.status == status


def test_vector_sizes():
    op = sfk.Operator(sfk.CONT_LAPLACE, 4, (16, 16, 16), inv_jac=(16, 16, 16))
    assert op.vector_size == 65 ** 3  # 274,625 nodes of the Q4 16^3 grid
    op = sfk.Operator(sfk.DG_SIPG, 5, (32, 32, 32), geometry=sfk.GEOM_PER_Q)
    assert op.vector_size == 7077888 and op.n_cell_q == 7077888 and op.n_face_q == 32 ** 3 * 36
    op = sfk.Operator(sfk.DG_PRECOMP, 3, (56, 50, 55))
    assert op.vector_size == 56 * 50 * 55 * 64  # the reference driver's default grid at degree 3


def test_setters_validate_sizes():
    op = sfk.Operator(sfk.DG_SIPG, 2, (2, 2, 2), geometry=sfk.GEOM_PER_Q)
    with pytest.raises(sfk.SfkError):
        op.set_cell_metric(np.zeros(10))
    op2 = sfk.Operator(sfk.CONT_LAPLACE, 2, (2, 2, 2), inv_jac=(2, 2, 2))
    with pytest.raises(sfk.SfkError):
        op2.set_coefficient(np.zeros(8 * 27))  # operator lacks SFK_FLAG_VARIABLE_COEF


@pytest.mark.parametrize("degree,kind,geom,default_cb", [
    (5, sfk.DG_SIPG, sfk.GEOM_PER_Q, 3), (5, sfk.DG_SIPG, sfk.GEOM_CARTESIAN, 7),
    (4, sfk.CONT_LAPLACE, sfk.GEOM_CARTESIAN, 10), (8, sfk.CELL_LAPLACE, sfk.GEOM_CARTESIAN, 3),
    (3, sfk.DG_ADVECT, sfk.GEOM_CARTESIAN, 16), (3, sfk.DG_PRECOMP, sfk.GEOM_CARTESIAN, 16),
    (8, sfk.DG_PRECOMP, sfk.GEOM_CARTESIAN, 3)])
def test_schedule_defaults_and_validation(degree, kind, geom, default_cb):
    kw = {"basis": sfk.BASIS_GLL} if kind in (sfk.CONT_LAPLACE, sfk.DG_ADVECT) else {}
    if kind == sfk.DG_ADVECT:
        kw["velocity"] = (1.0, 0.15, -0.05)
    op = sfk.Operator(kind, degree, (4, 4, 4), geometry=geom, inv_jac=(4, 4, 4), **kw)
    # the per-q SIPG kernel walks slabs by default (block order 2)
    assert op.schedule == (default_cb, 2 if geom == sfk.GEOM_PER_Q else 1, -1, 0, 0)
    m = op.max_cells_per_block
    assert default_cb <= m <= 256 // (degree + 1) ** 2
    op.set_schedule(1, 0, 1)
    assert op.schedule == (1, 0, 1, 0, 0)
    # scatter 2 (tensor-product kernel) exists for the continuous Laplacian only
    bad_scatter = 3 if kind == sfk.CONT_LAPLACE else 2
    for bad in [(m + 1, 1, -1), (-1, 1, -1), (0, 3, -1), (0, 1, bad_scatter), (0, 1, -1, -1, 0),
                (0, 1, -1, 99, 0), (0, 1, -1, 0, -1)]:
        with pytest.raises(sfk.SfkError):
            op.set_schedule(*bad)
    with pytest.raises(sfk.SfkError):
        op.tune(repeats=0, stream=0)  # rejected before any device work


@pytest.mark.parametrize("degree", range(1, 9))
@pytest.mark.parametrize("basis", [0, 1])
def test_library_tables_equal_oracle(degree, basis):
    t = sfk.tables(degree, basis)
    n = degree + 1
    mine = ob.eo_tables(degree, basis)
    # rebuild the even-odd values form from the library's dense S
    S = t["S"]
    h = (n + 1) // 2
    eo = np.zeros(n * h)
    for i in range((degree + 1) // 2):
        for q in range(h):
            eo[i * h + q] = 0.5 * (S[i, q] + S[i, n - 1 - q])
            eo[(degree - i) * h + q] = 0.5 * (S[i, q] - S[i, n - 1 - q])
    if degree % 2 == 0:
        eo[degree // 2 * h:degree // 2 * h + h] = S[degree // 2, :h]
    assert rel_err(eo, mine["values_eo"]) < 1e-13
    p, w, g = sfk.quadrature(n)
    assert np.abs(p - ob.gauss_points(n)).max() == 0.0
    assert np.abs(w - ob.gauss_weights(n)).max() == 0.0
    assert np.abs(g - ob.gll_points(n)).max() < 1e-16
    if basis == 0:
        assert abs(t["hermite_d"] - mine["hermite_d"]) <= 1e-13 * abs(mine["hermite_d"])


def test_no_oracle_in_product():
    """the shipped library never references the oracle or the reference"""
    data = open(sfk.LIB_PATH, "rb").read()
    assert b"sfko_" not in data and b"libsfkref" not in data and b"liboracle" not in data


@pytest.mark.parametrize("kind,kw", [(sfk.DG_SIPG, dict(geometry=sfk.GEOM_PER_Q)), (sfk.DG_PRECOMP, {})])
def test_dg_diagonal_rejects_self_neighbour(kind, kw):
    """one cell in a direction: the cell is its own periodic neighbour, which
    the diagonal kernels do not model; rejected before any device access (the
    pointer below is host memory and is never read)"""
    import ctypes
    op = sfk.Operator(kind, 1, (3, 1, 2), **kw)
    host = (ctypes.c_double * op.vector_size)()
    st = sfk.lib().sfk_compute_diagonal(op.handle, ctypes.cast(host, ctypes.c_void_p), None)
    assert st == 2  # SFK_ERR_UNSUPPORTED
    assert "own periodic neighbour" in sfk.lib().sfk_last_error().decode()


def test_raw_pointer_wrappers_validate_tensors():
    """the wrappers over raw-pointer entry points reject tensors the C ABI
    would read out of bounds (checked before any call into the library)"""
    import torch
    op = sfk.Operator(sfk.CONT_LAPLACE, 2, (2, 2, 2), inv_jac=(2.0, 2.0, 2.0))
    cpu = torch.zeros(op.vector_size, dtype=torch.float64)
    with pytest.raises(ValueError):
        sfk.solve_cg_cheby(op, op, cpu, cpu)
    with pytest.raises(ValueError):
        sfk.layout_to_canonical(2, (2, 2, 2), 2, torch.zeros(8 * 27, dtype=torch.float64))
    with pytest.raises(ValueError):
        sfk.vec_dot(torch.zeros(4), torch.zeros(4))
    with pytest.raises(ValueError):
        sfk.apply_1d(np.eye(3), 1, True, torch.zeros(9, dtype=torch.float64), 3, 3, 1)


def test_degree_limits():
    """degrees 1..15 ((p+1)^2 <= 256 threads per cell); the precomputed DG
    Laplacian and the tensor-product continuous kernel stop at 8"""
    sfk.Operator(sfk.CELL_LAPLACE, 15, (1, 1, 1))
    with pytest.raises(sfk.SfkError):
        sfk.Operator(sfk.CELL_LAPLACE, 16, (1, 1, 1))
    with pytest.raises(sfk.SfkError):
        sfk.Operator(sfk.DG_PRECOMP, 9, (2, 2, 2))
    op = sfk.Operator(sfk.CONT_LAPLACE, 9, (2, 2, 2), inv_jac=(2.0, 2.0, 2.0))
    with pytest.raises(sfk.SfkError):
        op.set_schedule(0, 1, 2)
