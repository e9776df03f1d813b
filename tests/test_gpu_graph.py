"""Chebyshev smoother sweeps and HIP-graph capture: the sweep equals the
sequence of fused steps with the Chebyshev coefficients, and graph replays
(Chebyshev sweep, LSRK(5,4) time step) equal the eager calls bit for bit on
every replay, with eager work and synchronisations between replays."""
import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP64, rel_err, sine_metric, to_dev, to_np

pytestmark = pytest.mark.gpu

sfk = pytest.importorskip("sum_factorization_kernels_amd")


def _cheby_coeffs(degree, lmin, lmax):
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    out, rho_old = [(1.0 / theta, 0.0)], 1.0 / sigma
    for _ in range(1, degree):
        rho = 1.0 / (2.0 * sigma - rho_old)
        out.append((2.0 * rho / delta, rho * rho_old))
        rho_old = rho
    return out


def _sipg_op(degree, nc):
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    cm, fm = sine_metric(degree, nc, 0.1)
    op.set_cell_metric(cm)
    op.set_face_metric(fm)
    return op


@pytest.mark.parametrize("kind,degree", [("sipg", 3), ("sipg", 5), ("cont", 4), ("cont_tp", 4), ("cont_tp", 6)])
@pytest.mark.parametrize("cheby_degree", [1, 4, 5])
def test_cheby_sweep_equals_fused_steps(cuda, kind, degree, cheby_degree):
    import torch
    nc = (3, 2, 4)
    if kind == "sipg":
        op = _sipg_op(degree, nc)
    else:
        op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=(3.0, 2.0, 4.0))
        # colored scatter: deterministic sums (atomics are not); the
        # tensor-product kernel (fused update) is deterministic as well
        op.set_schedule(0, 1, 1 if kind == "cont" else 2, 0, 1)
    rng = np.random.default_rng(6000 + degree)
    N = op.vector_size
    diag = op.diagonal()
    dinv = torch.where(diag != 0, 1.0 / diag, torch.zeros_like(diag))
    rhs = to_dev(rng.standard_normal(N))
    if kind == "cont":
        rhs = op.rhs_constant(1.0)
    x0 = to_dev(rng.standard_normal(N)) if kind == "sipg" else torch.zeros(N, dtype=torch.float64, device="cuda")
    lmin, lmax = 0.05, 2.5
    # the sweep (the work vector's initial contents are never read: NaN)
    x, spare, tmp = x0.clone(), torch.empty_like(x0), torch.full_like(x0, float("nan"))
    op.cheby_sweep(rhs, x, spare, dinv, tmp, cheby_degree, lmin, lmax)
    # the same steps one by one
    cur, nxt, tm = x0.clone(), torch.empty_like(x0), torch.full_like(x0, float("nan"))
    for c_np, c_tm in _cheby_coeffs(cheby_degree, lmin, lmax):
        op.cheby_step(rhs, cur, dinv, nxt, tm, c_np, c_tm)
        cur, nxt = nxt, cur
    torch.cuda.synchronize()
    assert torch.equal(x, cur)
    if kind == "sipg":
        # and against the oracle: x_{j+1} = x_j - tm_j with tm from the recurrence
        xr, tr = to_np(x0), np.zeros(N)
        cm, fm = sine_metric(degree, nc, 0.1)
        for c_np, c_tm in _cheby_coeffs(cheby_degree, lmin, lmax):
            tr = c_tm * tr + c_np * to_np(dinv) * (ob.dg_sipg_general(degree, nc, cm, fm, xr) - to_np(rhs))
            xr = xr - tr
        assert rel_err(to_np(x), xr) < 1e-11
    else:
        # against the oracle's operator with the same recurrence
        xr, tr = to_np(x0), np.zeros(N)
        ij = (3.0, 2.0, 4.0)
        for c_np, c_tm in _cheby_coeffs(cheby_degree, lmin, lmax):
            tr = c_tm * tr + c_np * to_np(dinv) * (ob.cont_laplacian(degree, nc, ij, xr.copy()) - to_np(rhs))
            xr = xr - tr
        assert rel_err(to_np(x), xr) < 1e-11


def test_graph_cheby_sweep_replays(cuda):
    import torch
    degree, nc = 4, (4, 3, 3)
    op = _sipg_op(degree, nc)
    st = sfk.Stream()
    rng = np.random.default_rng(6100)
    N = op.vector_size
    diag = op.diagonal()
    dinv = 1.0 / diag
    rhs = to_dev(rng.standard_normal(N))
    x0 = to_dev(rng.standard_normal(N))
    x, spare, tmp = x0.clone(), torch.empty_like(x0), torch.zeros_like(x0)
    torch.cuda.synchronize()
    # eager reference (also the warm-up before capture)
    op.cheby_sweep(rhs, x, spare, dinv, tmp, 5, 0.1, 2.0, stream=st.ptr)
    st.synchronize()
    ref = x.clone()
    x.copy_(x0)
    tmp.zero_()
    torch.cuda.synchronize()
    with sfk.Graph(st) as g:
        op.cheby_sweep(rhs, x, spare, dinv, tmp, 5, 0.1, 2.0, stream=st.ptr)
    for rep in range(4):
        x.copy_(x0)  # eager work between replays
        tmp.fill_(float("nan"))  # never read: the sweep's first step has c_tm = 0
        torch.cuda.synchronize()
        g.launch()
        st.synchronize()
        assert torch.equal(x, ref), rep


def test_graph_lsrk_time_step_replays(cuda):
    """five fused LSRK(5,4) stages (evaluation_dg_advect.h:190-214) captured
    once, replayed for several time steps against the eager stages"""
    import torch
    degree, nc = 3, (4, 4, 2)
    op = sfk.Operator(sfk.DG_ADVECT, degree, nc, inv_jac=(4.0, 4.0, 2.0), velocity=(1.0, 0.15, -0.05))
    a = [970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
         2251764453980. / 15575788980749., 26877169314380. / 34165994151039.]
    b = [1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
         -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
         5198255086312. / 14908931495163.]
    dt = 1e-3
    N = op.vector_size
    rng = np.random.default_rng(6200)

    def time_step(sn, so, stmp, stream):
        op.rk_stage(sn, sn, so, stmp, a[0] * dt, (b[0] - a[0]) * dt, stream=stream)
        op.rk_stage(so, stmp, sn, stmp, a[1] * dt, (b[1] - a[1]) * dt, stream=stream)
        op.rk_stage(sn, stmp, so, stmp, a[2] * dt, (b[2] - a[2]) * dt, stream=stream)
        op.rk_stage(so, stmp, stmp, sn, a[3] * dt, (b[3] - a[3]) * dt, stream=stream)
        op.rk_stage(stmp, sn, sn, stmp, b[4] * dt, 0.0, stream=stream)

    st = sfk.Stream()
    u0 = to_dev(rng.standard_normal(N))
    e_new, e_old, e_tmp = u0.clone(), torch.zeros_like(u0), torch.zeros_like(u0)
    g_new, g_old, g_tmp = u0.clone(), torch.zeros_like(u0), torch.zeros_like(u0)
    torch.cuda.synchronize()
    time_step(g_new, g_old, g_tmp, st.ptr)  # warm-up, then reset
    st.synchronize()
    g_new.copy_(u0)
    torch.cuda.synchronize()
    with sfk.Graph(st) as g:
        time_step(g_new, g_old, g_tmp, st.ptr)
    for step in range(4):
        time_step(e_new, e_old, e_tmp, None)
        torch.cuda.synchronize()
        g.launch()
        st.synchronize()
        assert torch.equal(g_new, e_new), step
    # and the first step against the oracle's stage sequence
    x = to_np(u0)
    xo, xt = np.zeros(N), np.zeros(N)
    ij, vel = (4.0, 4.0, 2.0), (1.0, 0.15, -0.05)

    def stage(src, upd, c_np, c_tm):
        r = ob.dg_advect(degree, nc, ij, vel, src)
        return upd + c_np * r, (upd + c_np * r) + c_tm * r

    xo, xt = stage(x, x, a[0] * dt, (b[0] - a[0]) * dt)
    x, xt = stage(xo, xt, a[1] * dt, (b[1] - a[1]) * dt)
    xo, xt = stage(x, xt, a[2] * dt, (b[2] - a[2]) * dt)
    xt2, x = stage(xo, xt, a[3] * dt, (b[3] - a[3]) * dt)
    x_final, _ = stage(xt2, x, b[4] * dt, 0.0)
    ref_after_one = x_final
    e1, o1, t1 = u0.clone(), torch.zeros_like(u0), torch.zeros_like(u0)
    time_step(e1, o1, t1, None)
    assert rel_err(to_np(e1), ref_after_one) < TOL_FP64
