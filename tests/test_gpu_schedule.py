"""Workgroup schedules and the autotuner: every schedule of an operator
kernel (cells per workgroup, block order, continuous scatter) gives the
oracle's result, and DG/element-local results do not depend on the schedule
at all (same arithmetic per cell, bit-identical)."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP32, TOL_FP64, rel_err, sine_metric, to_dev, to_np

pytestmark = pytest.mark.gpu

sfk = pytest.importorskip("sum_factorization_kernels_amd")


def _schedules(op, scatters=(-1,), orders=(0, 1)):
    m = op.max_cells_per_block
    cbs = sorted({1, 2, max(1, m // 2), m})
    return [(cb, order, sc) for sc in scatters for order in orders for cb in cbs]


@pytest.mark.parametrize("degree", [1, 2, 4, 6, 8])
def test_cont_schedules_match_oracle(cuda, degree):
    # odd cell counts: partial segments and single-row colours
    nc = (7, 3, 5) if degree <= 4 else (5, 3, 3)
    ij = tuple(float(c) for c in nc)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    rng = np.random.default_rng(4000 + degree)
    src = rng.uniform(-1, 1, op.vector_size)
    ref = ob.cont_laplacian(degree, nc, ij, src)
    d_src = to_dev(src)
    colored = {}
    for sched in _schedules(op, scatters=(0, 1)):
        op.set_schedule(*sched)
        assert op.schedule[:3] == sched
        out = to_np(op.apply(d_src))
        assert rel_err(out, ref) < TOL_FP64, sched
        if sched[2] == 1:
            # colored stores: the summation order of a node depends on the
            # segment length only, not on the block order
            cb = sched[0]
            if cb in colored:
                assert np.array_equal(out, colored[cb]), sched
            else:
                colored[cb] = out
            # repeated applies are deterministic
            assert np.array_equal(to_np(op.apply(d_src)), out)


def _nan_vec(op, n=None):
    import torch
    return torch.full((op.vector_size * (n or 1),), float("nan"), dtype=op.dtype, device="cuda")


def _set_tp(op, order, shape, zc):
    """set a tensor-product schedule; False when the tile shape needs more
    LDS than a workgroup has in this precision (a shape that only FP32 fits)"""
    try:
        op.set_schedule(0, order, 2, shape, zc)
    except sfk.SfkError as e:
        if e.status == 2 and "LDS" in str(e):
            return False
        raise
    return True


@pytest.mark.parametrize("degree", range(1, 9))
def test_cont_tp_schedules_match_oracle(cuda, degree):
    """tensor-product kernel: every tile shape and z range (ragged tiles,
    z ranges that start above the bottom, tiles larger than the mesh) against
    the oracle's quadrature-form operator; every node is written (the output
    starts as NaN), Dirichlet nodes exactly 0"""
    nc = (7, 5, 6) if degree <= 4 else (5, 3, 4)
    ij = (7.0, 2.5, 4.5)  # anisotropic cells: distinct direction constants
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    rng = np.random.default_rng(4050 + degree)
    src = rng.uniform(-1, 1, op.vector_size)
    ref = ob.cont_laplacian(degree, nc, ij, src)
    d_src = to_dev(src)
    g = [c * degree + 1 for c in nc]
    count = op.tile_shape(0)[0]
    too_big = set()
    for shape in range(count):
        for zc in (0, 1, 2, 3, nc[2]):
            for order in (0, 1):
                if not _set_tp(op, order, shape, zc):
                    too_big.add(shape)
                    continue
                out = to_np(op.apply(d_src, dst=_nan_vec(op)))
                assert np.isfinite(out).all(), (shape, zc, order)
                assert rel_err(out, ref) < TOL_FP64, (shape, zc, order)
                o3 = out.reshape(g[2], g[1], g[0])
                for face in (o3[0], o3[-1], o3[:, 0], o3[:, -1], o3[:, :, 0], o3[:, :, -1]):
                    assert (face == 0).all(), (shape, zc, order)
    # only Q6's 24-cell-wide FP32 tile exceeds the LDS of an FP64 workgroup
    assert too_big == ({5} if degree == 6 else set()), too_big


def test_cont_tp_fp32_every_shape(cuda):
    """FP32 Q6 on every tile shape (the 24-cell-wide one only fits FP32),
    x cells not a multiple of any tile width: the matvec against the oracle
    and the solver's three-term Chebyshev preconditioner against the
    tm-form sweep (two kernel paths on the same tile)"""
    import torch
    degree, nc, ij = 6, (25, 3, 4), (25.0, 3.0, 4.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, precision=sfk.FP32)
    rng = np.random.default_rng(4495)
    src = rng.uniform(-1, 1, op.vector_size).astype(np.float32).astype(np.float64)
    ref = ob.cont_laplacian(degree, nc, ij, src)
    d = op.diagonal()
    dinv = torch.where(d != 0, 1.0 / d, torch.zeros_like(d))
    rhs = to_dev(rng.standard_normal(op.vector_size), torch.float32) * (d != 0)
    for shape in range(op.tile_shape(0)[0]):
        for zc in (0, 3):
            assert _set_tp(op, 1, shape, zc), shape
            out = to_np(op.apply(to_dev(src, torch.float32), dst=_nan_vec(op)))
            assert rel_err(out, ref) < TOL_FP32, (shape, zc)
            z = op.cheby_precondition(rhs, 6, 2.5 / 30.0, 2.5)
            x, spare, tm = torch.zeros_like(rhs), _nan_vec(op), _nan_vec(op)
            op.cheby_sweep(rhs, x, spare, dinv, tm, 6, 2.5 / 30.0, 2.5)
            assert rel_err(to_np(z), to_np(x)) < 5e-5, (shape, zc)


@pytest.mark.parametrize("degree,nc", [(4, (1, 1, 1)), (4, (2, 1, 3)), (6, (1, 2, 1)), (2, (40, 3, 2)),
                                       (3, (3, 30, 2))])
def test_cont_tp_small_and_long_meshes(cuda, degree, nc):
    ij = tuple(float(c) for c in nc)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    src = np.random.default_rng(4060 + degree).uniform(-1, 1, op.vector_size)
    ref = ob.cont_laplacian(degree, nc, ij, src)
    for shape in range(op.tile_shape(0)[0]):
        if not _set_tp(op, 1, shape, 0):
            continue
        out = to_np(op.apply(to_dev(src), dst=_nan_vec(op)))
        assert rel_err(out, ref) < TOL_FP64 or np.abs(ref).max() == 0 and (out == 0).all(), shape


@pytest.mark.parametrize("degree", [2, 4, 6, 8])
def test_cont_tp_fp32_multi_rhs(cuda, degree):
    import torch
    nc = (6, 5, 4) if degree <= 4 else (4, 3, 3)
    ij = (6.0, 5.0, 4.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, precision=sfk.FP32, n_rhs=3)
    op.set_schedule(0, 1, 2, 0, 2)
    N = op.vector_size
    rng = np.random.default_rng(4070 + degree)
    src = rng.standard_normal(3 * N)
    out = to_np(op.apply(to_dev(src, torch.float32), dst=_nan_vec(op, 3)))
    for r in range(3):
        ref = ob.cont_laplacian(degree, nc, ij, src[r * N:(r + 1) * N].copy())
        assert rel_err(out[r * N:(r + 1) * N], ref) < TOL_FP32, r


@pytest.mark.parametrize("degree", [1, 4, 6])
@pytest.mark.parametrize("c_tm", [0.0, 0.37])
@pytest.mark.parametrize("zc", [0, 1])
def test_cont_tp_chebyshev_step(cuda, degree, c_tm, zc):
    """Chebyshev update fused into the tensor-product kernel against
    tm' = c_tm tm + c_np diag (A u - rhs), u' = u - tm' with A u from the
    oracle, on every node (Dirichlet rows too); tm not read when c_tm == 0"""
    nc = (5, 3, 4)
    ij = (5.0, 3.0, 4.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    op.set_schedule(0, 1, 2, 0, zc)
    N = op.vector_size
    rng = np.random.default_rng(4450 + degree)
    u, rhs = rng.uniform(-1, 1, N), rng.standard_normal(N)
    diag, tm = rng.uniform(0.5, 1.5, N), rng.standard_normal(N)
    if c_tm == 0.0:
        tm[:] = np.nan
    d_new, d_tm = _nan_vec(op), to_dev(tm)
    op.cheby_step(to_dev(rhs), to_dev(u), to_dev(diag), d_new, d_tm, 0.8, c_tm)
    au = ob.cont_laplacian(degree, nc, ij, u)
    r_tm = (c_tm * tm if c_tm != 0.0 else 0.0) + 0.8 * diag * (au - rhs)
    assert rel_err(to_np(d_tm), r_tm) < TOL_FP64
    assert rel_err(to_np(d_new), u - r_tm) < TOL_FP64


@pytest.mark.parametrize("degree", [2, 5])
def test_cont_colored_multi_rhs_coefficient(cuda, degree):
    nc = (5, 4, 3)
    ij = (5.0, 4.0, 3.0)
    n3 = (degree + 1) ** 3
    rng = np.random.default_rng(4100 + degree)
    k = np.exp(0.5 * rng.standard_normal(int(np.prod(nc)) * n3))
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, flags=sfk.FLAG_VARIABLE_COEF, n_rhs=2)
    op.set_coefficient(k)
    op.set_schedule(2, 1, 1)
    N = op.vector_size
    src = rng.standard_normal(2 * N)
    out = to_np(op.apply(to_dev(src)))
    for r in range(2):
        ref = ob.cont_laplacian(degree, nc, ij, src[r * N:(r + 1) * N].copy(), coef=k)
        assert rel_err(out[r * N:(r + 1) * N], ref) < TOL_FP64


@pytest.mark.parametrize("degree", [1, 3, 5, 8])
def test_dg_schedules_bit_identical(cuda, degree):
    nc = (5, 3, 4)
    ij = (4.0, 2.0, 4.0 / 3.0)
    n3 = (degree + 1) ** 3
    rng = np.random.default_rng(4200 + degree)
    src = rng.standard_normal(n3 * int(np.prod(nc)))
    d_src = to_dev(src)
    cases = [
        (sfk.Operator(sfk.DG_SIPG, degree, nc, inv_jac=ij), None),
        (sfk.Operator(sfk.CELL_LAPLACE, degree, (37, 1, 1), inv_jac=ij), None),
        (sfk.Operator(sfk.DG_ADVECT, degree, nc, inv_jac=ij, velocity=(1.0, 0.15, -0.05)), None),
    ]
    gen = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    cm, fm = sine_metric(degree, nc, 0.1)
    gen.set_cell_metric(cm)
    gen.set_face_metric(fm)
    cases.append((gen, ob.dg_sipg_general(degree, nc, cm, fm, src)))
    for op, ref in cases:
        x = d_src if op.vector_size == len(src) else to_dev(rng.standard_normal(op.vector_size))
        base = to_np(op.apply(x))
        if ref is not None:
            assert rel_err(base, ref) < TOL_FP64
        # order 2: DG kernels walk slabs of z-layers (a permutation of the cells)
        for sched in _schedules(op, orders=(0, 1, 2)):
            op.set_schedule(*sched)
            assert np.array_equal(to_np(op.apply(x)), base), (op.kind, sched)


def test_schedule_validation(cuda):
    op = sfk.Operator(sfk.DG_SIPG, 5, (4, 4, 4), geometry=sfk.GEOM_PER_Q)
    m = op.max_cells_per_block
    assert 1 <= m <= 256 // 36
    for bad in [(m + 1, 1, -1), (-1, 1, -1), (0, 3, -1), (0, 1, 2)]:
        with pytest.raises(sfk.SfkError):
            op.set_schedule(*bad)
    op.set_schedule(0, 0, -1)
    assert op.schedule[0] == 3  # the degree's default


def test_tune_picks_valid_schedule(cuda):
    degree, nc = 4, (6, 5, 4)
    ij = (6.0, 5.0, 4.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    sched, t = op.tune(repeats=3)
    assert t > 0
    assert 1 <= sched[0] <= op.max_cells_per_block and sched[1] in (0, 1) and sched[2] in (0, 1, 2)
    assert op.schedule == sched
    src = np.random.default_rng(4300).uniform(-1, 1, op.vector_size)
    assert rel_err(to_np(op.apply(to_dev(src))), ob.cont_laplacian(degree, nc, ij, src)) < TOL_FP64


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tune_sets_smoother_schedule(cuda, prec):
    """sfk_tune also picks the fused Chebyshev launches' tile, z_cells and
    order; set_schedule makes them follow the matvec schedule again.  The
    Chebyshev step with the tuned schedule matches the oracle, and the
    solver (whose smoother uses that schedule) reaches 1e-10."""
    import torch
    degree, nc = 4, (6, 5, 8)
    ij = (6.0, 5.0, 8.0)
    p = sfk.FP32 if prec == "fp32" else sfk.FP64
    tol = TOL_FP32 if prec == "fp32" else TOL_FP64
    dt = torch.float32 if prec == "fp32" else torch.float64
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, precision=p)
    sched, _ = op.tune(repeats=2)
    sm = op.smoother_schedule
    count = op.tile_shape(0)[0]
    assert sm[2] == sched[2] == 2
    assert sm[1] in (0, 1) and 0 <= sm[3] < count and 1 <= sm[4] <= nc[2]
    N = op.vector_size
    rng = np.random.default_rng(4470)
    u, rhs = rng.uniform(-1, 1, N), rng.standard_normal(N)
    diag, tm = rng.uniform(0.5, 1.5, N), rng.standard_normal(N)
    if prec == "fp32":
        u, rhs, diag, tm = (v.astype(np.float32).astype(np.float64) for v in (u, rhs, diag, tm))
    d_new, d_tm = _nan_vec(op), to_dev(tm, dt)
    op.cheby_step(to_dev(rhs, dt), to_dev(u, dt), to_dev(diag, dt), d_new, d_tm, 0.8, 0.37)
    r_tm = 0.37 * tm + 0.8 * diag * (ob.cont_laplacian(degree, nc, ij, u) - rhs)
    assert rel_err(to_np(d_tm), r_tm) < tol
    assert rel_err(to_np(d_new), u - r_tm) < tol
    # the solver with this operator as the smoother
    outer = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    b = outer.rhs_constant(1.0)
    x = torch.zeros_like(b)
    st = sfk.solve_cg_cheby(outer, op, b, x, rel_tol=1e-10, max_iter=500, cheby_degree=5, lanczos_steps=12)
    assert st.final_residual <= 1e-10 * float(torch.linalg.norm(b))
    # a schedule set by hand applies to the Chebyshev launches as well
    op.set_schedule(0, 0, 2, 1, 2)
    assert op.smoother_schedule == op.schedule


@pytest.mark.parametrize("degree", [1, 3, 6])
@pytest.mark.parametrize("scatter", [0, 1])
@pytest.mark.parametrize("c_tm", [0.0, 0.37])
def test_cont_chebyshev_step(cuda, degree, scatter, c_tm):
    """continuous Chebyshev step under both scatter schedules against
    tm' = c_tm tm + c_np diag (A u - rhs), u' = u - tm' with A u from the oracle"""
    nc = (5, 3, 4)
    ij = (5.0, 3.0, 4.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij)
    op.set_schedule(2, 1, scatter)
    N = op.vector_size
    rng = np.random.default_rng(4400 + degree)
    u, rhs = rng.uniform(-1, 1, N), rng.standard_normal(N)
    diag, tm = rng.uniform(0.5, 1.5, N), rng.standard_normal(N)
    if c_tm == 0.0:
        tm[:] = np.nan  # not read when c_tm == 0
    d_new, d_tm = to_dev(np.zeros(N)), to_dev(tm)
    op.cheby_step(to_dev(rhs), to_dev(u), to_dev(diag), d_new, d_tm, 0.8, c_tm)
    au = ob.cont_laplacian(degree, nc, ij, u)
    r_tm = (c_tm * tm if c_tm != 0.0 else 0.0) + 0.8 * diag * (au - rhs)
    assert rel_err(to_np(d_tm), r_tm) < TOL_FP64
    assert rel_err(to_np(d_new), u - r_tm) < TOL_FP64


def test_tune_cache(cuda, tmp_path, monkeypatch):
    """SFK_TUNE_CACHE: the first tune appends its result, a second operator
    with the same key reads it back without timing (shown by editing the
    stored line), a different mesh appends a new line"""
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("SFK_TUNE_CACHE", str(path))
    nc, ij = (6, 5, 8), (6.0, 5.0, 8.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, 4, nc, inv_jac=ij)
    sched, t = op.tune(repeats=2)
    lines = path.read_text().splitlines()
    assert len(lines) == 1
    f = lines[0].split()
    assert [int(v) for v in f[1:6]] == list(sched) and float(f[9]) == pytest.approx(t, rel=1e-6)
    # a hand-edited line (another valid schedule, another time) is what the next tune returns
    other = (0, 1 - sched[1], 2, (sched[3] + 1) % op.tile_shape(0)[0], 2)
    f[1:6] = [str(v) for v in other]
    f[9] = "0.125"
    path.write_text(" ".join(f) + "\n")
    op2 = sfk.Operator(sfk.CONT_LAPLACE, 4, nc, inv_jac=ij)
    sched2, t2 = op2.tune(repeats=2)
    assert tuple(sched2) == other and t2 == 0.125
    assert op2.schedule[1:] == other[1:]
    src = np.random.default_rng(4480).uniform(-1, 1, op2.vector_size)
    assert rel_err(to_np(op2.apply(to_dev(src))), ob.cont_laplacian(4, nc, ij, src)) < TOL_FP64
    # another mesh: a new line
    op3 = sfk.Operator(sfk.CONT_LAPLACE, 4, (6, 5, 9), inv_jac=(6.0, 5.0, 9.0))
    op3.tune(repeats=2)
    assert len(path.read_text().splitlines()) == 2


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("degree", [2, 3, 6])
@pytest.mark.parametrize("p", [2, 4, 6])
def test_cheby_precondition_matches_sweep(cuda, prec, degree, p):
    """the solver's preconditioner (three-term Chebyshev on the tensor-product
    kernel, first iterate formed from rhs in the kernel) equals the tm-form
    sweep from x = 0 (a separate kernel path) with the same bounds, on every
    node, for NaN-prefilled work vectors"""
    import torch
    nc, ij = (5, 4, 6), (5.0, 4.0, 6.0)
    pr = sfk.FP32 if prec == "fp32" else sfk.FP64
    dt = torch.float32 if prec == "fp32" else torch.float64
    op = sfk.Operator(sfk.CONT_LAPLACE, p, nc, inv_jac=ij, precision=pr)
    op.set_schedule(0, 1, 2, 1, 2)
    N = op.vector_size
    rng = np.random.default_rng(4490 + p + degree)
    rhs = to_dev(rng.standard_normal(N), dt) * (op.diagonal() != 0)  # Dirichlet rows of a residual are 0
    d = op.diagonal()
    dinv = torch.where(d != 0, 1.0 / d, torch.zeros_like(d))
    lmax, lmin = 2.5, 2.5 / 30.0  # (bounds of the order of lambda(D^{-1}A); the two paths share them)
    z = op.cheby_precondition(rhs, degree, lmin, lmax)
    x, spare, tm = torch.zeros_like(rhs), _nan_vec(op), _nan_vec(op)
    op.cheby_sweep(rhs, x, spare, dinv, tm, degree, lmin, lmax)
    tol = 5e-5 if prec == "fp32" else 1e-10
    assert rel_err(to_np(z), to_np(x)) < tol


def test_tune_per_q_sipg_chebyshev_schedule(cuda, tmp_path, monkeypatch):
    """per-q SIPG: sfk_tune also picks the fused Chebyshev step's cells per
    block and order (smoother schedule), the step on it matches the oracle,
    set_schedule drops it, and the tune cache restores it"""
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("SFK_TUNE_CACHE", str(path))
    degree, nc = 3, (4, 3, 5)
    cm, fm = sine_metric(degree, nc, 0.1)
    N = (degree + 1) ** 3 * int(np.prod(nc))
    rng = np.random.default_rng(4510)
    src, rhs = rng.standard_normal(N), rng.standard_normal(N)
    diag, tm = rng.uniform(0.5, 1.5, N), rng.standard_normal(N)
    ref = ob.dg_sipg_general(degree, nc, cm, fm, src)
    r_tm = 0.45 * tm + 0.6 * diag * (ref - rhs)

    def make():
        op = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
        op.set_cell_metric(cm)
        op.set_face_metric(fm)
        return op

    def check(op):
        d_new, d_tm = to_dev(np.zeros(N)), to_dev(tm)
        op.cheby_step(to_dev(rhs), to_dev(src), to_dev(diag), d_new, d_tm, 0.6, 0.45)
        assert rel_err(to_np(d_tm), r_tm) < TOL_FP64
        assert rel_err(to_np(d_new), src - r_tm) < TOL_FP64

    op = make()
    sched, _ = op.tune(repeats=2)
    sm = op.smoother_schedule
    assert 1 <= sm[0] <= op.max_cells_per_block and 0 <= sm[1] <= 2
    check(op)
    f = path.read_text().split()
    assert int(f[6]) == -1 and (int(f[7]), int(f[8])) == (sm[0], sm[1])
    op2 = make()
    op2.tune(repeats=2)  # from the cache line
    assert op2.smoother_schedule == sm
    check(op2)
    op2.set_schedule(*sched)
    assert op2.smoother_schedule == op2.schedule
    check(op2)
