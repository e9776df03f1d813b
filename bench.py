"""Benchmark of the MI355X sum-factorization operators.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu]
                    [--only WORKLOAD [--schedule CB,ORDER,SCATTER,SHAPE,ZC]] [--copy-probe]

Headline workload (one JSON line on stdout): the continuous Q_6 Laplacian
matvec on 64^3 affine Cartesian cells of [0,1]^3 (385^3 = 57,066,625 nodes,
homogeneous Dirichlet), FP64, input standard normal; a "step" is one matvec,
value = GDoF/s (nodes per second).  This is the mixed-precision solve's
configuration: the same line carries its FP32 matvec and the CG-Chebyshev
times to an FP64 relative residual of 1e-10 (all-FP64 against the FP32
smoother inside FP64 CG).  Under "extra.configs" every other configuration
is measured with its own roofline: continuous Q4 on 16^3 cells (100
repeated applies) and on 128^3 cells, the DG-SIPG p=5 32^3 deformed matvec
and Chebyshev sweep, DG advection p=3 48^3 (FP64 and FP32 forward, adjoint,
velocity gradient), the variable-coefficient Q8 24^3 8-RHS apply with its
coefficient gradient, and the precomputed-matrix DG Laplacian p=3.

Timing: HIP events (torch.cuda.Event) on the stream every kernel is launched
on (torch's current stream, passed through the C ABI), inputs resident in
HBM; before the W warm-up steps, 50 ms of untimed calls bring the GPU to
its clocks under sustained load (settle()), so that K steps measure the
steady state also when K and W are small.  roofline.achieved = algorithmic bytes per launch / mean launch time;
the algorithmic bytes count every vector and stored array the operation must
touch once, at its stored precision (the rule of the reference's drivers,
legacy/src/test_dg_integrate.cc:127,183,235).
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
FP32_PEAK_TFLOPS = 157.3

WORKLOADS = ("q6-fp64", "q6-fp32", "q4-128", "q4-16", "sipg-p5", "advect-p3", "advect-p3-fp32", "varcoef-q8",
             "precomp-p3", "cg-q6")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def settle(fn, seconds=0.05):
    """untimed calls of fn for about `seconds` of sustained GPU load: the
    MI355X raises its clocks only under load, and the first ~60 launches
    after any pause run up to 20 % slower (Q6 64^3 FP64 matvec: 339, 305,
    294, 285 us over consecutive batches of 20), so the timed steps follow
    this, the warm-up steps and the autotuner's trials"""
    import torch
    t_end = time.perf_counter() + seconds
    while True:
        for _ in range(8):
            fn()
        torch.cuda.synchronize()
        if time.perf_counter() >= t_end:
            return


def timed(fn, steps, warmup):
    import torch
    settle(fn)
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


def roofline(bytes_per_launch, seconds, traffic=None, flops=None, fp32=False):
    """HBM roofline block of one kernel: algorithmic bytes / launch time"""
    gbs = bytes_per_launch / seconds / 1e9
    r = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": int(bytes_per_launch)}
    if flops is not None:
        # FLOPs of the reference's own formula for the operation (an
        # effective rate: the kernel's algorithm may need fewer)
        tf = flops / seconds / 1e12
        r["ref_formula_tflops"] = round(tf, 2)
        r["ref_formula_tflops_frac"] = round(tf / (FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS), 4)
    return r


def sipg_bytes_per_dof(degree, elem):
    """algorithmic bytes per dof of the per-q-metric SIPG matvec: src + dst,
    the cell metric (6 per q-point) and the face metric (3 faces per cell,
    7 values per face q-point), all at the operator precision"""
    n = degree + 1
    return elem * (2 + 6 + 3 * 7 * n * n / n ** 3)


def advect_bytes_per_dof(degree, elem):
    """forward/adjoint advection with stored velocity: src + dst, the cell
    velocity (3 values per q-point) and the face a.n (3 faces per cell)"""
    n = degree + 1
    return elem * (2 + 3 + 3 * n * n / n ** 3)


def cell_laplacian_flops(degree):
    """FLOPs per cell of the reference's cell Laplacian formula
    (legacy/src/test_program.cc:63-97, even-odd sweeps and q-point operator)"""
    ops_interpolate = 2 * ((degree + 1) // 2) * 2 + degree + 1 + 2 * ((degree - 1) * (degree + 1) // 2)
    return 4 * 3 * ops_interpolate * (degree + 1) ** 2 + 3 * 2 * 3 * (degree + 1) ** 3


def power_lmax(op, dinv, n, iters=15):
    import torch
    v = torch.randn(n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(7))
    lam = 0.0
    for _ in range(iters):
        w = op.apply(v) * dinv
        lam = float(torch.dot(w, v) / torch.dot(v, v))
        v = w / torch.linalg.norm(w)
    return lam


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary under
    profiles/: a line 'traffic_bytes_per_launch: <value>'"""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


# ------------------------------------------------------------ CPU baselines
def _ref_lib(stem):
    import ctypes
    flags = open("/proc/cpuinfo").read()
    isa = "avx512" if " avx512f" in flags else "avx2"
    path = os.path.join(ROOT, "oracle", "_ref", f"{stem}_{isa}.so")
    return ctypes.CDLL(path) if os.path.exists(path) else None


def cpu_cell_laplacian(degree, n_cells, threads, target_seconds=8.0):
    """the reference's cell Laplacian (EvaluationCellLaplacian with
    READ_VECTOR, test_program.cc protocol: one evaluator per thread) over the
    cells of the continuous workload, compiled from the reference sources into
    oracle/_ref; GDoF/s counted in continuous nodes like the GPU number"""
    import ctypes
    lib = _ref_lib("libsfkref_cell")
    if lib is None:
        return None
    lib.refbench_cell_laplacian.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double)]
    lanes = lib.refcell_n_lanes()
    cells = int(np.prod(n_cells))
    # all cells over `threads` evaluators; one core: an eighth of the cells
    share = cells if threads > 1 else cells // 8
    nb = max(1, share // (lanes * threads))
    sec = ctypes.c_double()
    lib.refbench_cell_laplacian(degree, nb, 2, 1, threads, ctypes.byref(sec))
    per = max(sec.value / 2, 1e-4)
    n_tests = max(3, int(target_seconds / 3 / per))
    lib.refbench_cell_laplacian(degree, nb, n_tests, 3, threads, ctypes.byref(sec))
    t = sec.value / n_tests
    done = nb * lanes * threads
    nodes = np.prod([c * degree + 1 for c in n_cells]) * done / cells
    return {"value": round(nodes / t / 1e9, 4), "unit": "GDoF/s", "cores": threads, "kind": "reference",
            "sample": f"reference EvaluationCellLaplacian<3,{degree},double>::matrix_vector_product with READ_VECTOR "
                      f"(element-local: the reference has no assembled continuous operator), {done} of the "
                      f"{cells} cells ({lanes}-lane SIMD), one evaluator per thread, best of 3 x {n_tests} "
                      f"products, {threads} thread(s); nodes/s scaled to the continuous grid of those cells"}


def cpu_dg_laplacian(degree, n_cells, threads, target_seconds=5.0):
    """the reference's DG-SIPG evaluator (Cartesian, FP64), all host cores"""
    import ctypes
    lib = _ref_lib("libsfkref_bench")
    if lib is None:
        return None
    lib.refbench_dg_laplacian.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    nc = (ctypes.c_uint * 3)(*n_cells)
    sec = ctypes.c_double()
    lib.refbench_dg_laplacian(degree, nc, 3, 1, threads, 0, ctypes.byref(sec))
    per = max(sec.value / 3, 1e-4)
    n_tests = max(5, int(target_seconds / 3 / per))
    lib.refbench_dg_laplacian(degree, nc, n_tests, 3, threads, 0, ctypes.byref(sec))
    ndof = (degree + 1) ** 3 * int(np.prod(n_cells))
    return {"value": round(ndof * n_tests / sec.value / 1e9, 4), "unit": "GDoF/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference EvaluationDGLaplacian<3,{degree},double>::do_matvec (Cartesian), "
                      f"{n_cells[0]}x{n_cells[1]}x{n_cells[2]} cells, best of 3 x {n_tests} matvecs"}


def cpu_dg_advect(degree, n_cells, threads, target_seconds=5.0):
    """the reference's DG advection evaluator (constant velocity, FP64)"""
    import ctypes
    lib = _ref_lib("libsfkref_bench")
    if lib is None or not hasattr(lib, "refbench_dg_advect"):
        return None
    lib.refbench_dg_advect.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    nc = (ctypes.c_uint * 3)(*n_cells)
    sec = ctypes.c_double()
    lib.refbench_dg_advect(degree, nc, 3, 1, threads, ctypes.byref(sec))
    per = max(sec.value / 3, 1e-4)
    n_tests = max(5, int(target_seconds / 3 / per))
    lib.refbench_dg_advect(degree, nc, n_tests, 3, threads, ctypes.byref(sec))
    ndof = (degree + 1) ** 3 * int(np.prod(n_cells))
    return {"value": round(ndof * n_tests / sec.value / 1e9, 4), "unit": "GDoF/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference EvaluationDGAdvection<3,{degree},double>::do_matvec (constant velocity), "
                      f"{n_cells[0]}x{n_cells[1]}x{n_cells[2]} cells, best of 3 x {n_tests} matvecs"}


# ------------------------------------------------------------ workloads
def cont_op(degree, n, prec, tune=True, repeats=10, **kw):
    import sum_factorization_kernels_amd as sfk
    nc = (n, n, n)
    op = sfk.Operator(sfk.CONT_LAPLACE, degree, nc, inv_jac=(float(n),) * 3, precision=prec, **kw)
    sched = op.tune(repeats=repeats)[0] if tune else op.schedule
    return op, sched


def run_cont(degree, n, prec, steps, warmup, tune=True, seed=3):
    """continuous Laplacian matvec: (seconds, op, schedule, x, y)"""
    import torch
    import sum_factorization_kernels_amd as sfk
    op, sched = cont_op(degree, n, prec, tune)
    dt = torch.float64 if prec == sfk.FP64 else torch.float32
    x = torch.randn(op.vector_size, dtype=torch.float64, device="cuda",
                    generator=torch.Generator("cuda").manual_seed(seed)).to(dt)
    y = torch.empty_like(x)
    t = timed(lambda: op.apply(x, y), steps, warmup)
    return t, op, sched, x, y


def cont_entry(name, degree, n, prec, t, sched, profile=None):
    import sum_factorization_kernels_amd as sfk
    N = (n * degree + 1) ** 3
    elem = 8 if prec == sfk.FP64 else 4
    cells = n ** 3
    e = {"time_us": round(t * 1e6, 2), "gdofs": round(N / t / 1e9, 3), "n_nodes": N,
         "schedule": list(sched),
         "roofline": roofline(2 * elem * N, t, read_traffic(profile) if profile else None,
                              flops=cells * cell_laplacian_flops(degree), fp32=elem == 4)}
    log(f"{name}: {t * 1e6:.1f} us, {N / t / 1e9:.2f} GDoF/s, {2 * elem * N / t / 1e9:.0f} GB/s")
    return e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=WORKLOADS, help="one workload's kernel (profiling)")
    ap.add_argument("--schedule", help="with --only: the schedule to run (the five integers of a bench line's "
                                       "'schedule' entry), default: the operator's default schedule")
    ap.add_argument("--copy-probe", action="store_true",
                    help="plain copy kernel of known size (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the operators would shard by cells in z-slabs, but the library "
              "has no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import sum_factorization_kernels_amd as sfk
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    if args.copy_probe:
        n = 64 * 1024 * 1024  # 512 MiB of FP64 read + 512 MiB written per launch
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n, "launches": 5}))
        return
    if args.only:
        return only(args)

    # ------------------------------------------------------------ headline
    log("headline: continuous Q6 on 64^3 cells, FP64 matvec")
    t64, o3, sched, x3, y3 = run_cont(6, 64, sfk.FP64, args.steps, args.warmup)
    N = o3.vector_size
    head = cont_entry("Q6 64^3 FP64", 6, 64, sfk.FP64, t64, sched, "cont_tp_q6_64cube_fp64_pmc.txt")
    configs = {"c3_q6_64cube": {"fp64": head}}
    t32, s32, sched32, _, _ = run_cont(6, 64, sfk.FP32, args.steps, args.warmup)
    configs["c3_q6_64cube"]["fp32"] = cont_entry("Q6 64^3 FP32", 6, 64, sfk.FP32, t32, sched32,
                                                 "cont_tp_q6_64cube_fp32_pmc.txt")
    configs["c3_q6_64cube"]["fp32_over_fp64_matvec_speedup"] = round(t64 / t32, 3)
    # CG-Chebyshev to 1e-10: both smoothers at degree 6 (the configuration)
    # and each at its best degree of {4, 6, 8, 12, 16}
    b3 = o3.rhs_constant(1.0)
    bn = float(torch.linalg.norm(b3))
    try:
        runs = {}
        for deg in ((6,) if args.no_extras else (4, 6, 8, 12, 16)):
            for name, sm in (("mixed", s32), ("fp64", o3)):
                x3.zero_()
                runs[(name, deg)] = sfk.solve_cg_cheby(o3, sm, b3, x3, 1e-10, 3000, deg, 12)
        st_m, st_d = runs[("mixed", 6)], runs[("fp64", 6)]
        best_m = min((st.seconds, deg) for (name, deg), st in runs.items() if name == "mixed")
        best_d = min((st.seconds, deg) for (name, deg), st in runs.items() if name == "fp64")
        cg = {"mixed_seconds": round(st_m.seconds, 4), "fp64_seconds": round(st_d.seconds, 4),
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
        configs["c3_q6_64cube"]["cg_cheby"] = cg
        log(f"CG-Chebyshev Q6 64^3 (degree 6): mixed {st_m.seconds:.3f} s ({st_m.iterations} it), "
            f"FP64 {st_d.seconds:.3f} s ({st_d.iterations} it): {cg['speedup']}x; best {cg['speedup_best_vs_best']}x")
    except sfk.SfkError as e:
        configs["c3_q6_64cube"]["cg_cheby"] = {"error": str(e)}
        log(f"CG-Chebyshev Q6 64^3 failed: {e}")
    del s32, b3, x3, y3, o3

    if not args.no_extras:
        extras(args, configs)

    cpu, cpu_more = None, {}
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        log(f"cpu baseline: reference cell Laplacian Q6 on the 64^3 cells, {cores} threads and 1 thread")
        cpu = cpu_cell_laplacian(6, (64, 64, 64), cores)
        cpu_more["cell_laplacian_q6_64cube_1core"] = cpu_cell_laplacian(6, (64, 64, 64), 1, 4.0)
        if not args.no_extras:
            cpu_more["dg_sipg_p5_32cube_cartesian"] = cpu_dg_laplacian(5, (32, 32, 32), cores)
            cpu_more["dg_advect_p3_48cube"] = cpu_dg_advect(3, (48, 48, 48), cores)

    line = {
        "metric": "GDoF/s per matvec (FP64 vs FP32) + CG-Chebyshev time to 1e-10 residual",
        "value": round(N / t64 / 1e9, 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t64 * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: standard normal input (torch generator, seed 3); CG right-hand side f = 1",
        "config": {"workload": "continuous Q_6 Laplacian matvec, 64^3 affine Cartesian cells of [0,1]^3, "
                               "homogeneous Dirichlet, FP64 (mixed-precision CG-Chebyshev configuration)",
                   "n_nodes": N, "degree": 6, "n_cells": [64, 64, 64],
                   "schedule": {"cells_per_block": sched[0], "block_order": sched[1], "kernel": sched[2],
                                "tile_shape": sched[3], "z_cells": sched[4]}},
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "extra": {"configs": configs, "cpu_reference": cpu_more},
    }
    line["roofline"] = {k: v for k, v in head["roofline"].items() if k in
                        ("bound", "achieved", "peak", "unit", "frac", "traffic")}
    print(json.dumps(line), flush=True)


def extras(args, configs):
    import torch
    import sum_factorization_kernels_amd as sfk
    # c1: continuous Q4 on 16^3, Dirichlet, 100 repeated applies
    o1, sched1 = cont_op(4, 16, sfk.FP64, repeats=20)
    x1 = torch.rand(o1.vector_size, dtype=torch.float64, device="cuda",
                    generator=torch.Generator("cuda").manual_seed(0)) * 2 - 1
    y1 = torch.empty_like(x1)
    t1 = timed(lambda: o1.apply(x1, y1), 100, 10)
    c1 = cont_entry("Q4 16^3 FP64", 4, 16, sfk.FP64, t1, sched1)
    # the 100 applies captured in one HIP graph (launch overhead removed)
    st1 = sfk.Stream()
    torch.cuda.synchronize()
    with sfk.Graph(st1) as g1:
        for _ in range(100):
            o1.apply(x1, y1, stream=st1.ptr)
    g1.launch()
    st1.synchronize()
    t0 = time.perf_counter()
    g1.launch()
    st1.synchronize()
    tg1 = (time.perf_counter() - t0) / 100
    c1["graph_time_us"] = round(tg1 * 1e6, 2)
    c1["graph_gdofs"] = round(o1.vector_size / tg1 / 1e9, 3)
    c1["note"] = "4.39 MB per apply: resident in the 256 MB Infinity Cache, latency-bound; see c1_scaled"
    configs["c1_q4_16cube"] = c1
    del g1, st1, o1, x1, y1
    # c1 scaled to 128^3 cells (HBM regime)
    tb, ob, schedb, _, _ = run_cont(4, 128, sfk.FP64, args.steps, args.warmup)
    configs["c1_scaled_q4_128cube"] = cont_entry("Q4 128^3 FP64", 4, 128, sfk.FP64, tb, schedb,
                                                 "cont_tp_q4_128cube_fp64_pmc.txt")
    del ob
    torch.cuda.empty_cache()

    # c2: DG-SIPG p=5 32^3 deformed, per-q metric, FP64 (+ FP32, Chebyshev)
    degree, nc = 5, (32, 32, 32)
    op = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q)
    op.set_sine_deformation(0.1)
    N = op.vector_size
    u = torch.randn(N, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    v = torch.empty_like(u)
    sched, _ = op.tune(repeats=10)
    t = timed(lambda: op.apply(u, v), args.steps, args.warmup)
    c2 = {"fp64": {"time_us": round(t * 1e6, 2), "gdofs": round(N / t / 1e9, 3), "schedule": list(sched),
                   "roofline": roofline(sipg_bytes_per_dof(degree, 8) * N, t,
                                        read_traffic("sipg_general_p5_32cube_fp64_pmc.txt"))}}
    log(f"SIPG p5 32^3 deformed FP64: {t * 1e6:.1f} us, {N / t / 1e9:.2f} GDoF/s")
    op32 = sfk.Operator(sfk.DG_SIPG, degree, nc, geometry=sfk.GEOM_PER_Q, precision=sfk.FP32)
    op32.set_sine_deformation(0.1)
    sched32, _ = op32.tune(repeats=10)
    u32, v32 = u.float(), torch.empty(N, dtype=torch.float32, device="cuda")
    t32 = timed(lambda: op32.apply(u32, v32), args.steps, args.warmup)
    c2["fp32"] = {"time_us": round(t32 * 1e6, 2), "gdofs": round(N / t32 / 1e9, 3), "schedule": list(sched32),
                  "roofline": roofline(sipg_bytes_per_dof(degree, 4) * N, t32)}
    # one degree-5 point-Jacobi Chebyshev sweep with the operator diagonal
    diag = op.diagonal()
    dinv = 1.0 / diag
    lmax = 1.1 * power_lmax(op, dinv, N)
    rhs = torch.ones(N, dtype=torch.float64, device="cuda")
    xa, xb, tm = torch.zeros_like(u), torch.empty_like(u), torch.zeros_like(u)
    tch = timed(lambda: op.cheby_sweep(rhs, xa, xb, dinv, tm, 5, lmax / 30.0, lmax), max(10, args.steps // 5), 3)
    # a fused step reads src, rhs, diag, tmp and the metric, writes dst, tmp
    step_bytes = (sipg_bytes_per_dof(degree, 8) + 4 * 8) * N
    c2["cheby5_sweep"] = {"time_ms": round(tch * 1e3, 4), "step_gdofs": round(5 * N / tch / 1e9, 3),
                          "roofline": roofline(5 * step_bytes, tch)}
    log(f"SIPG Chebyshev(5) sweep: {tch * 1e3:.3f} ms")
    # Cartesian SIPG (the reference's own operator, correct x sweep)
    opc = sfk.Operator(sfk.DG_SIPG, degree, nc, inv_jac=(32.0, 32.0, 32.0))
    schedc, _ = opc.tune(repeats=5)
    tc = timed(lambda: opc.apply(u, v), args.steps, args.warmup)
    c2["cartesian_fp64"] = {"time_us": round(tc * 1e6, 2), "gdofs": round(N / tc / 1e9, 3),
                            "schedule": list(schedc), "roofline": roofline(16 * N, tc, flops=158.3 * N)}
    configs["c2_sipg_p5_32cube_deformed"] = c2
    del op, op32, opc, u, v, u32, v32, xa, xb, tm, rhs, diag, dinv
    torch.cuda.empty_cache()

    # c4: DG advection p=3, 48^3 periodic, velocity (1, 0.5+0.25 sin 2 pi x, 0.25) at q-points
    c4 = {}
    cv, fv = advection_velocity(3, (48, 48, 48))
    for name, prec, dt, elem in (("fp64", sfk.FP64, torch.float64, 8), ("fp32", sfk.FP32, torch.float32, 4)):
        o4 = sfk.Operator(sfk.DG_ADVECT, 3, (48, 48, 48), inv_jac=(48.0,) * 3, precision=prec,
                          flags=sfk.FLAG_VARIABLE_VELOCITY)
        o4.set_velocity(cv, fv)
        sched4, _ = o4.tune(repeats=5)
        N4 = o4.vector_size
        u4 = torch.randn(N4, dtype=torch.float64, device="cuda",
                         generator=torch.Generator("cuda").manual_seed(4)).to(dt)
        v4 = torch.empty_like(u4)
        t4 = timed(lambda: o4.apply(u4, v4), args.steps, args.warmup)
        t4a = timed(lambda: o4.apply_transpose(u4, v4), args.steps, args.warmup)
        t4g = timed(lambda: o4.grad_velocity(u4, v4), 20, 3)
        bpd = advect_bytes_per_dof(3, elem)
        c4[name] = {"schedule": list(sched4),
                    "forward": {"time_us": round(t4 * 1e6, 2), "gdofs": round(N4 / t4 / 1e9, 3),
                                "roofline": roofline(bpd * N4, t4, read_traffic(f"advect_p3_48cube_{name}_pmc.txt"))},
                    "adjoint": {"time_us": round(t4a * 1e6, 2), "gdofs": round(N4 / t4a / 1e9, 3),
                                "roofline": roofline(bpd * N4, t4a)},
                    # reads u and lambda, writes the cell (3 per q-point) and face (3 per face point) gradients
                    "velocity_grad": {"time_us": round(t4g * 1e6, 2),
                                      "roofline": roofline((2 + 3 + 3 * 16 / 64) * elem * N4, t4g)}}
        log(f"advection p3 48^3 {name}: fwd {N4 / t4 / 1e9:.2f}, adj {N4 / t4a / 1e9:.2f} GDoF/s, "
            f"grad {t4g * 1e6:.1f} us")
        del o4, u4, v4
    configs["c4_advect_p3_48cube"] = c4
    torch.cuda.empty_cache()

    # c5: variable-coefficient Q8 24^3, 8 RHS in one launch, FP32
    o5 = sfk.Operator(sfk.CONT_LAPLACE, 8, (24, 24, 24), inv_jac=(24.0,) * 3, precision=sfk.FP32,
                      flags=sfk.FLAG_VARIABLE_COEF, n_rhs=8)
    k5 = np.exp(0.5 * np.random.default_rng(5).standard_normal(o5.n_cell_q))
    o5.set_coefficient(k5)
    sched5, _ = o5.tune(repeats=5)
    N5 = o5.vector_size
    x5 = torch.randn(8 * N5, dtype=torch.float32, device="cuda")
    y5 = torch.empty_like(x5)
    t5 = timed(lambda: o5.apply(x5, y5), 20, 3)
    t5g = timed(lambda: o5.grad_coefficient(x5, y5), 10, 2)
    nq5 = o5.n_cell_q
    configs["c5_varcoef_q8_24cube_8rhs_fp32"] = {
        "gdofs": round(8 * N5 / t5 / 1e9, 3), "schedule": list(sched5), "apply_ms": round(t5 * 1e3, 4),
        "roofline": roofline(8 * 2 * 4 * N5 + 4 * nq5, t5, read_traffic("cont_varcoef_q8_24cube_8rhs_fp32_pmc.txt"),
                             flops=8 * 13824 * 99873, fp32=True),
        "coef_grad_ms": round(t5g * 1e3, 4),
        "coef_grad_roofline": roofline(8 * 2 * 4 * N5 + 4 * nq5, t5g)}
    log(f"var-coef Q8 24^3 x8 FP32: {t5 * 1e3:.3f} ms, {8 * N5 / t5 / 1e9:.2f} GDoF/s")
    del o5, x5, y5

    # precomputed-matrix DG Laplacian, degree 3, the reference driver's
    # default grid (test_dg_precomputed.cc: 10M dofs -> 56 x 50 x 55 cells)
    pc = {}
    for name, prec, dt in (("fp64", sfk.FP64, torch.float64), ("fp32", sfk.FP32, torch.float32)):
        o6 = sfk.Operator(sfk.DG_PRECOMP, 3, (56, 50, 55), precision=prec)
        sched6, _ = o6.tune(repeats=5)
        x6 = torch.randn(o6.vector_size, dtype=dt, device="cuda")
        y6 = torch.empty_like(x6)
        t6 = timed(lambda: o6.apply(x6, y6), args.steps, args.warmup)
        tr6 = read_traffic("precomputed_dg_p3_56x50x55_fp64.txt") if name == "fp64" else None
        pc[name] = {"gdofs": round(o6.vector_size / t6 / 1e9, 3), "apply_us": round(t6 * 1e6, 2),
                    "schedule": list(sched6), "roofline": roofline(2 * x6.element_size() * o6.vector_size, t6, tr6)}
        log(f"precomputed DG p3 56x50x55 {name}: {o6.vector_size / t6 / 1e9:.2f} GDoF/s")
        del o6, x6, y6
    configs["precomputed_dg_p3_56x50x55"] = pc


def only(args):
    """one workload's kernel, for rocprofv3 passes (every dispatch of the
    kernel is a measured one): the schedule given with --schedule (the
    bench line's, so that profile and line describe the same launches) or
    the operator's default"""
    import torch
    import sum_factorization_kernels_amd as sfk
    w = args.only

    def schedule(op):
        if args.schedule:
            op.set_schedule(*[int(v) for v in args.schedule.split(",")])
        return op.schedule

    if w in ("q6-fp64", "q6-fp32", "q4-128", "q4-16"):
        degree, n = (6, 64) if w.startswith("q6") else (4, 128 if w == "q4-128" else 16)
        prec = sfk.FP32 if w == "q6-fp32" else sfk.FP64
        op, _ = cont_op(degree, n, prec, tune=False)
        sched = schedule(op)
        dt = torch.float64 if prec == sfk.FP64 else torch.float32
        x = torch.randn(op.vector_size, dtype=torch.float64, device="cuda").to(dt)
        y = torch.empty_like(x)
        t = timed(lambda: op.apply(x, y), args.steps, args.warmup)
        log(f"{w}: schedule {sched}: {t * 1e6:.1f} us, {op.vector_size / t / 1e9:.2f} GDoF/s")
    elif w == "sipg-p5":
        op = sfk.Operator(sfk.DG_SIPG, 5, (32, 32, 32), geometry=sfk.GEOM_PER_Q)
        op.set_sine_deformation(0.1)
        sched = schedule(op)
        u = torch.randn(op.vector_size, dtype=torch.float64, device="cuda")
        v = torch.empty_like(u)
        t = timed(lambda: op.apply(u, v), args.steps, args.warmup)
        log(f"{w}: schedule {sched}: {t * 1e6:.1f} us, {op.vector_size / t / 1e9:.2f} GDoF/s")
    elif w in ("advect-p3", "advect-p3-fp32"):
        prec, dt = (sfk.FP32, torch.float32) if w.endswith("fp32") else (sfk.FP64, torch.float64)
        o4 = sfk.Operator(sfk.DG_ADVECT, 3, (48, 48, 48), inv_jac=(48.0,) * 3, precision=prec,
                          flags=sfk.FLAG_VARIABLE_VELOCITY)
        o4.set_velocity(*advection_velocity(3, (48, 48, 48)))
        sched = schedule(o4)
        u4 = torch.randn(o4.vector_size, dtype=torch.float64, device="cuda").to(dt)
        v4 = torch.empty_like(u4)
        t = timed(lambda: o4.apply(u4, v4), args.steps, args.warmup)
        ta = timed(lambda: o4.apply_transpose(u4, v4), args.steps, args.warmup)
        log(f"{w}: schedule {sched}: forward {t * 1e6:.1f} us, adjoint {ta * 1e6:.1f} us")
    elif w == "varcoef-q8":
        o5 = sfk.Operator(sfk.CONT_LAPLACE, 8, (24, 24, 24), inv_jac=(24.0,) * 3, precision=sfk.FP32,
                          flags=sfk.FLAG_VARIABLE_COEF, n_rhs=8)
        o5.set_coefficient(np.exp(0.5 * np.random.default_rng(5).standard_normal(o5.n_cell_q)))
        sched = schedule(o5)
        x5 = torch.randn(8 * o5.vector_size, dtype=torch.float32, device="cuda")
        y5 = torch.empty_like(x5)
        t = timed(lambda: o5.apply(x5, y5), args.steps, args.warmup)
        log(f"{w}: schedule {sched}: {t * 1e3:.3f} ms")
    elif w == "cg-q6":
        # one FP64 and one mixed CG-Chebyshev(6) solve of the Q6 64^3 problem
        # to 1e-10, operators tuned first (as in the full run)
        o3, _ = cont_op(6, 64, sfk.FP64)
        s32, _ = cont_op(6, 64, sfk.FP32)
        b = o3.rhs_constant(1.0)
        x = torch.zeros_like(b)
        for name, sm in (("fp64", o3), ("mixed", s32)):
            x.zero_()
            st = sfk.solve_cg_cheby(o3, sm, b, x, 1e-10, 3000, 6, 12)
            log(f"{w} {name}: {st.seconds:.3f} s, {st.iterations} iterations, smoother schedule "
                f"{sm.smoother_schedule}")
    elif w == "precomp-p3":
        op = sfk.Operator(sfk.DG_PRECOMP, 3, (56, 50, 55))
        sched = schedule(op)
        x = torch.randn(op.vector_size, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        t = timed(lambda: op.apply(x, y), args.steps, args.warmup)
        log(f"{w}: schedule {sched}: {t * 1e6:.1f} us, {op.vector_size / t / 1e9:.2f} GDoF/s")


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
