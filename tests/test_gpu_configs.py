"""The benchmark workloads at their full sizes, on their own inputs: every
operator against the CPU oracle (FP64 to 1e-12, FP32 to 2e-5, relative to
the largest entry), the degree-5 Chebyshev sweep against the oracle's
recurrence, and the CG-Chebyshev solves to an FP64 relative residual of
1e-10 recomputed independently."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP32, TOL_FP64, rel_err, sine_metric, to_dev, to_np

pytestmark = pytest.mark.gpu

sfk = pytest.importorskip("sum_factorization_kernels_amd")


def _zero_boundary(x, g):
    v = x.reshape(g[2], g[1], g[0])
    v[0], v[-1], v[:, 0], v[:, -1], v[:, :, 0], v[:, :, -1] = 0, 0, 0, 0, 0, 0
    return x


def test_q4_16cube_100_applies(cuda):
    """continuous Q4, 16^3 cells of [0,1]^3, Dirichlet, input uniform [-1,1)
    from seed 0 with boundary entries zeroed, 100 repeated applies"""
    import torch
    nc, ij = (16, 16, 16), (16.0, 16.0, 16.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, 4, nc, inv_jac=ij)
    assert op.vector_size == 65 ** 3
    x = _zero_boundary(np.random.default_rng(0).uniform(-1, 1, op.vector_size), (65, 65, 65))
    ref = ob.cont_laplacian(4, nc, ij, x)
    d_x = to_dev(x)
    y = torch.empty_like(d_x)
    worst = 0.0
    for _ in range(100):
        op.apply(d_x, y)
        worst = max(worst, rel_err(to_np(y), ref))
    assert worst < TOL_FP64


def test_sipg_p5_32cube_deformed(cuda):
    """DG-SIPG p5 on 32^3 cells deformed by x + 0.1 sin(pi x)sin(pi y)sin(pi z),
    input standard normal from seed 1: matvec (FP64, FP32) and one degree-5
    point-Jacobi Chebyshev sweep against the oracle"""
    import torch
    degree, nc = 5, (32, 32, 32)
    cm, fm = sine_metric(degree, nc, 0.1)
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    op.set_sine_deformation(0.1)  # device generator (pinned against sine_metric)
    x = np.random.default_rng(1).standard_normal(op.vector_size)
    ref = ob.dg_sipg_general(degree, nc, cm, fm, x)
    assert rel_err(to_np(op.apply(to_dev(x))), ref) < TOL_FP64
    op32 = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q, precision=sfk.FP32)
    op32.set_sine_deformation(0.1)
    assert rel_err(to_np(op32.apply(to_dev(x, torch.float32))), ref) < TOL_FP32
    # the schedules a benchmark runs: the autotuner's pick on this device
    # and the cells-per-block / block order of the recorded runs (3, 2)
    for sched in ("tuned", (3, 2, -1)):
        for o in (op, op32):
            if sched == "tuned":
                o.tune(repeats=2)
            else:
                o.set_schedule(*sched)
        assert rel_err(to_np(op.apply(to_dev(x))), ref) < TOL_FP64, sched
        assert rel_err(to_np(op32.apply(to_dev(x, torch.float32))), ref) < TOL_FP32, sched
    # Chebyshev(5) sweep from x with rhs = 1 on [lmax/30, lmax]
    dinv = 1.0 / op.diagonal()
    rhs = np.ones(op.vector_size)
    lmax = 3.0
    xs, spare, tmp = to_dev(x), to_dev(np.zeros_like(x)), to_dev(np.zeros_like(x))
    op.cheby_sweep(to_dev(rhs), xs, spare, dinv, tmp, 5, lmax / 30.0, lmax)
    theta, delta = 0.5 * (lmax + lmax / 30.0), 0.5 * (lmax - lmax / 30.0)
    sigma = theta / delta
    d = to_np(dinv)
    xr, tr, rho_old = x.copy(), np.zeros_like(x), 1.0 / sigma
    for j in range(5):
        if j == 0:
            c_np, c_tm = 1.0 / theta, 0.0
        else:
            rho = 1.0 / (2.0 * sigma - rho_old)
            c_np, c_tm = 2.0 * rho / delta, rho * rho_old
            rho_old = rho
        tr = c_tm * tr + c_np * d * (ob.dg_sipg_general(degree, nc, cm, fm, xr) - rhs)
        xr = xr - tr
    assert rel_err(to_np(xs), xr) < 1e-11


def test_q6_64cube_matvec_and_solves(cuda):
    """continuous Q6 on 64^3 cells: FP64 and FP32 matvec against the oracle,
    then the FP64 and the mixed-precision CG-Chebyshev(6) solves of the load
    vector of f = 1, each to an FP64 relative residual of 1e-10 recomputed
    here with the FP64 operator"""
    import torch
    nc, ij = (64, 64, 64), (64.0, 64.0, 64.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, 6, nc, inv_jac=ij)
    s32 = sfk.Operator(sfk.CONT_LAPLACE, 6, nc, inv_jac=ij, precision=sfk.FP32)
    g = (385, 385, 385)
    x = _zero_boundary(np.random.default_rng(3).uniform(-1, 1, op.vector_size), g)
    ref = ob.cont_laplacian(6, nc, ij, x)
    assert rel_err(to_np(op.apply(to_dev(x))), ref) < TOL_FP64
    assert rel_err(to_np(s32.apply(to_dev(x, torch.float32))), ref) < TOL_FP32
    # the autotuner's schedules (tile shape, z cells, block order: what a
    # benchmark runs), then the solves on them
    op.tune(repeats=2)
    s32.tune(repeats=2)
    assert rel_err(to_np(op.apply(to_dev(x))), ref) < TOL_FP64, op.schedule
    assert rel_err(to_np(s32.apply(to_dev(x, torch.float32))), ref) < TOL_FP32, s32.schedule
    del ref, x
    b = op.rhs_constant(1.0)
    bn = float(torch.linalg.norm(b))
    for smoother in (op, s32):
        sol = torch.zeros_like(b)
        st = sfk.solve_cg_cheby(op, smoother, b, sol, 1e-10, 3000, 6, 12)
        r = b - op.apply(sol)
        assert float(torch.linalg.norm(r)) / bn <= 1e-10
        assert st.final_residual / bn <= 1e-10


def test_varcoef_q8_24cube_8rhs(cuda):
    """-div(k grad u), continuous Q8 on 24^3 cells, k = exp(g) with
    g ~ 0.5 N(0,1) per q-point (seed 5), 8 right-hand sides in one launch,
    FP32 against the FP64 oracle (first, a middle and the last right-hand
    side); the coefficient gradient summed over the 8 pairs against the oracle"""
    import torch
    degree, nc, ij = 8, (24, 24, 24), (24.0, 24.0, 24.0)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=ij, precision=sfk.FP32,
                      flags=sfk.FLAG_VARIABLE_COEF, n_rhs=8)
    k = np.exp(0.5 * np.random.default_rng(5).standard_normal(op.n_cell_q))
    op.set_coefficient(k)
    N = op.vector_size
    rng = np.random.default_rng(6)
    u = rng.standard_normal(8 * N)
    refs = {r: ob.cont_laplacian(degree, nc, ij, u[r * N:(r + 1) * N].copy(), coef=k) for r in (0, 3, 7)}
    # the default schedule, the autotuner's pick on this device, and the
    # schedule of the recorded runs (3 cells per workgroup, block order 1,
    # colored stores)
    for sched in ("default", "tuned", (3, 1, 1)):
        if sched == "tuned":
            op.tune(repeats=2)
        elif sched != "default":
            op.set_schedule(*sched)
        out = to_np(op.apply(to_dev(u, torch.float32)))
        for r, ref in refs.items():
            assert rel_err(out[r * N:(r + 1) * N], ref) < TOL_FP32, (sched, r)
    lam = rng.standard_normal(8 * N)
    g = to_np(op.grad_coefficient(to_dev(u, torch.float32), to_dev(lam, torch.float32)))
    gref = ob.cont_laplacian_coef_grad(degree, nc, ij, u, lam, n_rhs=8)
    assert rel_err(g, gref) < TOL_FP32
