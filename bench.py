"""Benchmark of the MI355X sum-factorization operators.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras]

Headline workload (one JSON line on stdout): DG-SIPG Laplacian, degree 5
(6^3 dofs per cell), 32^3 cells of [0,1]^3 deformed by
x + 0.1 sin(pi x) sin(pi y) sin(pi z) (1,1,1) with the per-quadrature-point
metric stored (cell: 6 values per q-point, faces: 7 per face q-point), FP64,
input standard normal from seed 1; a "step" is one matvec, value = GDoF/s.
The same line carries the FP32 matvec, one degree-5 point-Jacobi Chebyshev
sweep, and (unless --no-extras) the other workloads: continuous Q4 16^3
(100 repeated applies), the mixed-precision CG-Chebyshev solve of the
continuous Q6 Laplacian on 64^3 cells against the all-FP64 solve, DG
advection p=3 on 48^3 (forward, adjoint, velocity gradient), the
variable-coefficient Q8 24^3 8-RHS apply and the precomputed-matrix DG
Laplacian p=3 on the reference driver's 56 x 50 x 55 grid.

Timing: HIP events (torch.cuda.Event) on the stream every kernel is launched
on (torch's current stream, passed through the C ABI), inputs resident in
HBM.  roofline.achieved = algorithmic bytes per matvec / mean matvec time.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / steps  # seconds per step


def sipg_bytes_per_dof(degree, elem):
    """algorithmic bytes per dof of the per-q-metric SIPG matvec: src + dst,
    the cell metric (6 per q-point) and the face metric (3 faces per cell,
    7 values per face q-point), all at the operator precision"""
    n = degree + 1
    return elem * (2 + 6 + 3 * 7 * n * n / n ** 3)


def power_lmax(op, dinv, n, iters=15):
    import torch
    v = torch.randn(n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(7))
    lam = 0.0
    for _ in range(iters):
        w = op.apply(v) * dinv
        lam = float(torch.dot(w, v) / torch.dot(v, v))
        v = w / torch.linalg.norm(w)
    return lam


def cpu_baseline(degree, n_cells, target_seconds=12.0):
    """the reference's own DG-SIPG evaluator (Cartesian, FP64) compiled from
    the reference sources into oracle/_ref, timed on this host with all
    cores (test_dg_integrate protocol: best of repeats of n_tests matvecs)"""
    import ctypes
    here = os.path.join(ROOT, "oracle", "_ref")
    flags = open("/proc/cpuinfo").read()
    name = "libsfkref_bench_avx512.so" if " avx512f" in flags else "libsfkref_bench_avx2.so"
    path = os.path.join(here, name)
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.refbench_dg_laplacian.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    cores = len(os.sched_getaffinity(0))
    nc = (ctypes.c_uint * 3)(*n_cells)
    sec = ctypes.c_double()
    # calibrate: 3 matvecs
    lib.refbench_dg_laplacian(degree, nc, 3, 1, cores, 0, ctypes.byref(sec))
    per = max(sec.value / 3, 1e-4)
    n_tests = max(5, int(target_seconds / 3 / per))
    lib.refbench_dg_laplacian(degree, nc, n_tests, 3, cores, 0, ctypes.byref(sec))
    ndof = (degree + 1) ** 3 * int(np.prod(n_cells))
    return {"value": round(ndof * n_tests / sec.value / 1e9, 4), "unit": "GDoF/s", "cores": cores,
            "kind": "reference",
            "sample": f"reference EvaluationDGLaplacian<3,{degree},double> do_matvec (Cartesian: the reference has "
                      f"no deformed geometry), {n_cells[0]}x{n_cells[1]}x{n_cells[2]} cells, "
                      f"{lib.refbench_n_lanes()}-lane SIMD, best of 3 x {n_tests} matvecs, {cores} threads"}


def read_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary:
    a line 'traffic_bytes_per_launch: <value>'"""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-headline", action="store_true", help="headline matvec only (profiling)")
    ap.add_argument("--only-precomputed", action="store_true",
                    help="precomputed-matrix DG p3 56x50x55 FP64 matvec only, cells per workgroup 4 (profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="plain copy kernel of known size (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the operators would shard by cells, but the library has "
              "no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import sum_factorization_kernels_amd as sfk
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    extra = {}
    if args.copy_probe:
        n = 64 * 1024 * 1024  # 512 MiB of FP64 read + 512 MiB written per launch
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 5}))
        return

    if args.only_precomputed:
        op = sfk.Operator(sfk.DG_PRECOMP, 3, (56, 50, 55))
        op.set_schedule(4, 1, -1)
        x = torch.randn(op.vector_size, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        t = timed(lambda: op.apply(x, y), args.steps, args.warmup)
        log(f"precomputed DG p3 56x50x55 FP64: {t * 1e6:.1f} us, {op.vector_size / t / 1e9:.2f} GDoF/s")
        return

    # ------------------------------------------------------------ headline
    degree, nc = 5, (32, 32, 32)
    log("setup: DG-SIPG p=5 32^3 deformed (device metric)")
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    op.set_sine_deformation(0.1)
    N = op.vector_size
    g = torch.Generator("cuda").manual_seed(1)
    u = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    v = torch.empty_like(u)
    # workgroup schedule (cells per workgroup, block order) chosen by the
    # library's autotuner before the timed region (profiling runs,
    # --only-headline, keep the default schedule so that every dispatch of
    # the kernel is the measured one)
    if args.only_headline:
        sched = op.schedule
    else:
        sched, _ = op.tune(repeats=10)
    log(f"schedule (cells_per_block, block_order, scatter): {sched}")
    t64 = timed(lambda: op.apply(u, v), args.steps, args.warmup)
    gdofs = N / t64 / 1e9
    bytes_mv = sipg_bytes_per_dof(degree, 8) * N
    achieved = bytes_mv / t64 / 1e9
    log(f"SIPG p5 FP64 matvec: {t64 * 1e6:.1f} us, {gdofs:.2f} GDoF/s, {achieved:.0f} GB/s")
    if not args.only_headline:
        op32 = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q, precision=sfk.FP32)
        op32.set_sine_deformation(0.1)
        sched32, _ = op32.tune(repeats=10)
        extra["sipg_p5_deformed_fp32_schedule"] = list(sched32)
        u32, v32 = u.float(), torch.empty(N, dtype=torch.float32, device="cuda")
        t32 = timed(lambda: op32.apply(u32, v32), args.steps, args.warmup)
        extra["sipg_p5_deformed_fp32_gdofs"] = round(N / t32 / 1e9, 3)
        extra["sipg_p5_deformed_fp32_hbm_gbs"] = round(sipg_bytes_per_dof(degree, 4) * N / t32 / 1e9, 1)
        log(f"SIPG p5 FP32 matvec: {t32 * 1e6:.1f} us, {N / t32 / 1e9:.2f} GDoF/s")
        # one degree-5 point-Jacobi Chebyshev sweep with the operator diagonal
        diag = op.diagonal()
        dinv = 1.0 / diag
        lmax = 1.1 * power_lmax(op, dinv, N)
        rhs = torch.ones(N, dtype=torch.float64, device="cuda")
        xa, xb, tm = torch.zeros_like(u), torch.empty_like(u), torch.zeros_like(u)
        tch = timed(lambda: op.cheby_sweep(rhs, xa, xb, dinv, tm, 5, lmax / 30.0, lmax),
                    max(10, args.steps // 10), 3)
        extra["sipg_p5_deformed_cheby5_sweep_ms"] = round(tch * 1e3, 4)
        extra["sipg_p5_deformed_cheby_step_gdofs"] = round(5 * N / tch / 1e9, 3)
        log(f"Chebyshev(5) sweep: {tch * 1e3:.3f} ms")
        # the same sweep captured once in a HIP graph and replayed
        st = sfk.Stream()
        torch.cuda.synchronize()
        with sfk.Graph(st) as graph:
            op.cheby_sweep(rhs, xa, xb, dinv, tm, 5, lmax / 30.0, lmax, stream=st.ptr)
        reps = max(10, args.steps // 10)
        graph.launch()
        st.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            graph.launch()
        st.synchronize()
        extra["sipg_p5_deformed_cheby5_sweep_graph_ms"] = round((time.perf_counter() - t0) / reps * 1e3, 4)
        del graph, st
        # Cartesian SIPG (the reference's own operator, correct x sweep)
        opc = sfk.Operator(sfk.DG_SIPG, degree, nc, inv_jac=(32, 32, 32))
        schedc, _ = opc.tune(repeats=5)
        tc = timed(lambda: opc.apply(u, v), args.steps, args.warmup)
        extra["sipg_p5_cartesian_fp64_gdofs"] = round(N / tc / 1e9, 3)
        extra["sipg_p5_cartesian_schedule"] = list(schedc)
        log(f"SIPG p5 Cartesian FP64: {N / tc / 1e9:.2f} GDoF/s")
        del op32, opc, u32, v32, xa, xb, tm, rhs, diag, dinv

    if not args.no_extras and not args.only_headline:
        # c1: continuous Q4 on 16^3, Dirichlet, 100 repeated applies
        o1 = sfk.Operator(sfk.CONT_LAPLACE, 4, (16, 16, 16), inv_jac=(16, 16, 16))
        n1 = o1.vector_size
        x1 = (torch.rand(n1, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
              * 2 - 1)
        y1 = torch.empty_like(x1)
        sched1, _ = o1.tune(repeats=20)
        t1 = timed(lambda: o1.apply(x1, y1), 100, 10)
        extra["q4_16cube_fp64_gdofs"] = round(n1 / t1 / 1e9, 3)
        extra["q4_16cube_schedule"] = list(sched1)
        log(f"Q4 16^3 continuous: {t1 * 1e6:.1f} us, {n1 / t1 / 1e9:.2f} GDoF/s")
        # the 100 applies captured in one HIP graph (launch latency removed)
        st1 = sfk.Stream()
        torch.cuda.synchronize()
        o1.apply(x1, y1, stream=st1.ptr)
        st1.synchronize()
        with sfk.Graph(st1) as g1:
            for _ in range(100):
                o1.apply(x1, y1, stream=st1.ptr)
        g1.launch()
        st1.synchronize()
        t0 = time.perf_counter()
        g1.launch()
        st1.synchronize()
        tg1 = (time.perf_counter() - t0) / 100
        extra["q4_16cube_fp64_graph_gdofs"] = round(n1 / tg1 / 1e9, 3)
        del g1, st1
        # c1 scaled to 128^3 cells (HBM regime)
        o1b = sfk.Operator(sfk.CONT_LAPLACE, 4, (128, 128, 128), inv_jac=(128, 128, 128))
        xb1 = torch.randn(o1b.vector_size, dtype=torch.float64, device="cuda")
        yb1 = torch.empty_like(xb1)
        sched1b, _ = o1b.tune(repeats=3)
        tb1 = timed(lambda: o1b.apply(xb1, yb1), 20, 3)
        extra["q4_128cube_fp64_gdofs"] = round(o1b.vector_size / tb1 / 1e9, 3)
        extra["q4_128cube_fp64_hbm_gbs"] = round(16 * o1b.vector_size / tb1 / 1e9, 1)
        extra["q4_128cube_schedule"] = list(sched1b)
        log(f"Q4 128^3 continuous: {tb1 * 1e3:.3f} ms, {o1b.vector_size / tb1 / 1e9:.2f} GDoF/s")
        del o1b, xb1, yb1
        # c3: mixed-precision CG-Chebyshev on continuous Q6 64^3
        o3 = sfk.Operator(sfk.CONT_LAPLACE, 6, (64, 64, 64), inv_jac=(64, 64, 64))
        s32 = sfk.Operator(sfk.CONT_LAPLACE, 6, (64, 64, 64), inv_jac=(64, 64, 64), precision=sfk.FP32)
        o3.tune(repeats=3)
        s32.tune(repeats=3)
        b3 = o3.rhs_constant(1.0)
        x3 = torch.randn(o3.vector_size, dtype=torch.float64, device="cuda")
        y3 = torch.empty_like(x3)
        t3 = timed(lambda: o3.apply(x3, y3), 20, 3)
        x3f, y3f = x3.float(), torch.empty(o3.vector_size, dtype=torch.float32, device="cuda")
        t3f = timed(lambda: s32.apply(x3f, y3f), 20, 3)
        extra["q6_64cube_fp64_gdofs"] = round(o3.vector_size / t3 / 1e9, 3)
        extra["q6_64cube_fp32_gdofs"] = round(o3.vector_size / t3f / 1e9, 3)
        log(f"Q6 64^3 matvec FP64 {o3.vector_size / t3 / 1e9:.2f} GDoF/s, FP32 {o3.vector_size / t3f / 1e9:.2f}")
        del x3f, y3f
        bn = float(torch.linalg.norm(b3))
        try:
            # both solvers at Chebyshev degree 6, and each at its own best
            # degree of {4, 6, 8} (the degree trades outer FP64 iterations
            # against inner smoother steps differently for the two)
            runs = {}
            for deg in (4, 6, 8):
                for name, sm in (("mixed", s32), ("fp64", o3)):
                    x3.zero_()
                    runs[(name, deg)] = sfk.solve_cg_cheby(o3, sm, b3, x3, 1e-10, 3000, deg, 12)
            st_m, st_d = runs[("mixed", 6)], runs[("fp64", 6)]
            best_m = min((st.seconds, deg) for (name, deg), st in runs.items() if name == "mixed")
            best_d = min((st.seconds, deg) for (name, deg), st in runs.items() if name == "fp64")
            extra["cg_cheby_q6_64cube"] = {
                "mixed_seconds": round(st_m.seconds, 4), "fp64_seconds": round(st_d.seconds, 4),
                "speedup": round(st_d.seconds / st_m.seconds, 3),
                "mixed_rel_residual": st_m.final_residual / bn, "fp64_rel_residual": st_d.final_residual / bn,
                "mixed_iterations": st_m.iterations, "fp64_iterations": st_d.iterations,
                "mixed_matvecs_fp64_fp32": [st_m.matvecs_outer, st_m.matvecs_inner],
                "fp64_matvecs": st_d.matvecs_outer + st_d.matvecs_inner,
                "lambda_max": round(st_d.lambda_max, 6),
                "seconds_by_degree": {f"{name}_deg{deg}": round(st.seconds, 4) for (name, deg), st in runs.items()},
                "best_mixed": {"degree": best_m[1], "seconds": round(best_m[0], 4)},
                "best_fp64": {"degree": best_d[1], "seconds": round(best_d[0], 4)},
                "speedup_best_vs_best": round(best_d[0] / best_m[0], 3)}
            log(f"CG-Chebyshev Q6 64^3 (degree 6): mixed {st_m.seconds:.3f} s ({st_m.iterations} it), "
                f"FP64 {st_d.seconds:.3f} s ({st_d.iterations} it); best: mixed {best_m[0]:.3f} s (degree "
                f"{best_m[1]}), FP64 {best_d[0]:.3f} s (degree {best_d[1]})")
        except sfk.SfkError as e:
            extra["cg_cheby_q6_64cube"] = {"error": str(e)}
            log(f"CG-Chebyshev Q6 64^3 failed: {e}")
        del o3, s32, b3, x3, y3
        # c4: DG advection p=3, 48^3 periodic, velocity (1, 0.5+0.25 sin 2 pi x, 0.25) at q-points
        o4 = sfk.Operator(sfk.DG_ADVECT, 3, (48, 48, 48), inv_jac=(48, 48, 48), flags=sfk.FLAG_VARIABLE_VELOCITY)
        cv, fv = advection_velocity(3, (48, 48, 48))
        o4.set_velocity(cv, fv)
        sched4, _ = o4.tune(repeats=5)
        u4 = torch.randn(o4.vector_size, dtype=torch.float64, device="cuda",
                         generator=torch.Generator("cuda").manual_seed(4))
        v4 = torch.empty_like(u4)
        t4 = timed(lambda: o4.apply(u4, v4), 50, 5)
        t4a = timed(lambda: o4.apply_transpose(u4, v4), 50, 5)
        t4g = timed(lambda: o4.grad_velocity(u4, v4), 20, 3)
        extra["advect_p3_48cube"] = {"forward_gdofs": round(o4.vector_size / t4 / 1e9, 3),
                                     "adjoint_gdofs": round(o4.vector_size / t4a / 1e9, 3),
                                     "velocity_grad_ms": round(t4g * 1e3, 4), "schedule": list(sched4)}
        log(f"advection p3 48^3: fwd {o4.vector_size / t4 / 1e9:.2f} GDoF/s, adj {o4.vector_size / t4a / 1e9:.2f}")
        del o4, u4, v4
        # c5: variable-coefficient Q8 24^3, 8 RHS in one launch, FP32
        o5 = sfk.Operator(sfk.CONT_LAPLACE, 8, (24, 24, 24), inv_jac=(24, 24, 24), precision=sfk.FP32,
                          flags=sfk.FLAG_VARIABLE_COEF, n_rhs=8)
        k5 = np.exp(0.5 * np.random.default_rng(5).standard_normal(o5.n_cell_q))
        o5.set_coefficient(k5)
        sched5, _ = o5.tune(repeats=5)
        log(f"var-coef Q8 autotuned schedule: {sched5}")
        x5 = torch.randn(8 * o5.vector_size, dtype=torch.float32, device="cuda")
        y5 = torch.empty_like(x5)
        t5 = timed(lambda: o5.apply(x5, y5), 20, 3)
        t5g = timed(lambda: o5.grad_coefficient(x5, y5), 10, 2)
        extra["varcoef_q8_24cube_8rhs_fp32"] = {"gdofs": round(8 * o5.vector_size / t5 / 1e9, 3),
                                                "schedule": list(sched5),
                                                "apply_ms": round(t5 * 1e3, 4),
                                                "coef_grad_ms": round(t5g * 1e3, 4)}
        log(f"var-coef Q8 24^3 x8 FP32: {8 * o5.vector_size / t5 / 1e9:.2f} GDoF/s")
        del o5, x5, y5

        # precomputed-matrix DG Laplacian, degree 3, the reference driver's
        # default grid (test_dg_precomputed.cc: 10M dofs -> 56 x 50 x 55 cells)
        pc = {}
        for name, prec, dt in (("fp64", sfk.FP64, torch.float64), ("fp32", sfk.FP32, torch.float32)):
            o6 = sfk.Operator(sfk.DG_PRECOMP, 3, (56, 50, 55), precision=prec)
            sched6, _ = o6.tune(repeats=5)
            x6 = torch.randn(o6.vector_size, dtype=dt, device="cuda")
            y6 = torch.empty_like(x6)
            t6 = timed(lambda: o6.apply(x6, y6), 50, 5)
            pc[name] = {"gdofs": round(o6.vector_size / t6 / 1e9, 3), "apply_us": round(t6 * 1e6, 2),
                        "hbm_gbs": round(2 * x6.element_size() * o6.vector_size / t6 / 1e9, 1),
                        "schedule": list(sched6)}
            log(f"precomputed DG p3 56x50x55 {name}: {o6.vector_size / t6 / 1e9:.2f} GDoF/s")
            del o6, x6, y6
        extra["precomputed_dg_p3_56x50x55"] = pc

    cpu = None
    if not args.no_cpu and not args.only_headline:
        log("cpu baseline: reference DG-SIPG p=5 on the host cores")
        cpu = cpu_baseline(degree, nc)

    traffic = read_traffic(os.path.join(ROOT, "profiles", "sipg_general_p5_32cube_fp64_pmc.txt"))
    line = {
        "metric": "GDoF/s per matvec (FP64 vs FP32) + CG-Chebyshev time to 1e-10 residual",
        "value": round(gdofs, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t64 * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: standard normal input from seed 1, sine-deformed mesh metric generated on the device",
        "config": {"workload": "DG-SIPG Laplacian matvec, degree 5, 32^3 cells deformed by "
                               "x+0.1 sin(pi x)sin(pi y)sin(pi z), per-q metric stored, penalty (p+1)^2/h, FP64",
                   "n_dofs": N, "degree": degree, "n_cells": list(nc),
                   "schedule": {"cells_per_block": sched[0], "block_order": sched[1]}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


def advection_velocity(degree, nc):
    """velocity (1, 0.5 + 0.25 sin(2 pi x), 0.25) at the cell and face Gauss
    points of the uniform mesh of [0,1]^3 (layouts of sfk_op_set_velocity)"""
    import sum_factorization_kernels_amd as sfk
    n = degree + 1
    xg, _, _ = sfk.quadrature(n)
    ncell = int(np.prod(nc))
    ix = np.arange(ncell) % nc[0]
    x = (ix[:, None] + np.tile(xg, n * n)[None, :]) / nc[0]  # x of cell q-point (x fastest)
    cv = np.empty((ncell, 3, n ** 3))
    cv[:, 0, :] = 1.0
    cv[:, 1, :] = 0.5 + 0.25 * np.sin(2 * np.pi * x)
    cv[:, 2, :] = 0.25
    fv = np.empty((3, ncell, n * n))
    fv[0] = 1.0  # a.n on x-faces: a_x
    xf = (ix[:, None] + np.tile(xg, n)[None, :]) / nc[0]  # y-faces: lo = x
    fv[1] = 0.5 + 0.25 * np.sin(2 * np.pi * xf)
    fv[2] = 0.25  # a.n on z-faces: a_z
    return np.ascontiguousarray(cv).ravel(), np.ascontiguousarray(fv).ravel()


if __name__ == "__main__":
    main()
