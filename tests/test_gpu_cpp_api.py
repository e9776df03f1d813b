"""The reference-side C++ API (sum_factorization_kernels_amd/include/
sfk_evaluators.h: the reference's evaluator classes over the C ABI) and the
C++ drivers, run as programs: tests/cpp/evaluators_check.cc drives
do_matvec, do_chebyshev, do_time_step (with and without use_graph),
verify_derivative, compute_max_error, matrix_vector_product and the adjoint
API, and its outputs are compared with the oracle and with the reference's
known answers (tests/golden/advect_kat.npz).  The program is compiled with
plain g++ against the header and libsfk.so (no HIP toolchain), which the
CPU test checks as well."""
import os
import subprocess

import numpy as np
import pytest

import oracle_bindings as ob
from helpers import TOL_FP32, TOL_FP64, rel_err

sfk = pytest.importorskip("sum_factorization_kernels_amd")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sum_factorization_kernels_amd")
SRC = os.path.join(ROOT, "tests", "cpp", "evaluators_check.cc")
DRIVERS = os.path.join(PKG, "drivers")


def _compile(exe):
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(PKG, "include"), SRC, "-o", exe, "-L" + PKG, "-lsfk", "-Wl,-rpath," + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_api_compiles_with_plain_gxx(tmp_path):
    if not os.path.exists(os.path.join(PKG, "libsfk.so")):
        pytest.skip("libsfk.so not built")
    _compile(str(tmp_path / "evaluators_check"))


def _pattern(n, a):
    i = np.arange(n)
    return np.sin(a * (i % 1013) + 0.1 * (i // 1013))


@pytest.fixture(scope="module")
def outputs(tmp_path_factory):
    d = tmp_path_factory.mktemp("cpp_api")
    exe = str(d / "evaluators_check")
    _compile(exe)
    k = np.load(os.path.join(ROOT, "tests", "golden", "advect_kat.npz"))
    nc = [str(int(x)) for x in k["n_cells"]]
    steps = str(int(k["p3"][3]))
    print("running", exe, flush=True)
    r = subprocess.run([exe, str(d)] + nc + [steps], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    scal = {}
    for line in open(d / "scalars.txt"):
        parts = line.split()
        # "label value" or "label name value name value ..."
        scal[parts[0]] = [float(parts[1])] if len(parts) == 2 else [float(v) for v in parts[2::2]]
    return d, scal


def _load(d, name):
    return np.fromfile(d / f"{name}.bin", dtype=np.float64)


@pytest.mark.gpu
def test_dg_laplacian_do_matvec_and_do_chebyshev(cuda, outputs):
    d, _ = outputs
    nc, ij = (4, 3, 2), (4.0, 2.0, 4.0 / 3.0)  # the reference driver's Jacobian
    n = 64 * int(np.prod(nc))
    x = _pattern(n, 0.37)
    assert rel_err(_load(d, "dg_matvec"), ob.dg_sipg(3, nc, ij, x)) < TOL_FP64
    old, rhs, diag, tm = _pattern(n, 0.21), _pattern(n, 0.53), 1.5 + 0.5 * _pattern(n, 0.11), _pattern(n, 0.71)
    new_ref, tm_ref = ob.dg_sipg_cheby(3, nc, ij, rhs, old, diag, tm, 0.5, 0.5)
    assert rel_err(_load(d, "dg_cheby_new"), new_ref) < TOL_FP64
    assert rel_err(_load(d, "dg_cheby_tmp"), tm_ref) < TOL_FP64


@pytest.mark.gpu
def test_cell_laplacian_matvec_and_verification(cuda, outputs):
    d, scal = outputs
    x = _pattern(7 * 125, 0.29)
    ref = ob.cell_laplacian(4, 0, x, (4.0, 2.0, 4.0 / 3.0))  # Hermite-like basis, as the reference class
    assert rel_err(_load(d, "cell_matvec"), ref) < TOL_FP64
    # boundary-integral check for the basis in use (exact up to rounding)
    assert scal["cell_verification"][0] < 1e-11


@pytest.mark.gpu
def test_advection_time_steps_match_reference_kat(cuda, outputs):
    """verify_derivative and compute_max_error after the reference's 20
    steps (evaluation_dg_advect.h:190-283) equal the reference's values,
    with and without the HIP-graph replay of the five stages"""
    _, scal = outputs
    k = np.load(os.path.join(ROOT, "tests", "golden", "advect_kat.npz"))
    vd, err, t_end, _ = k["p3"]
    for graph in (0, 1):
        v, t, e = scal[f"advect_graph{graph}"]
        assert abs(v - vd) <= 1e-10 * max(1.0, vd), (graph, v, vd)
        assert abs(t - t_end) < 1e-12
        assert abs(e - err) <= 1e-9 * max(1.0, err), (graph, e, err)


@pytest.mark.gpu
def test_advection_adjoint_api(cuda, outputs):
    d, _ = outputs
    nc, ij, n = (4, 4, 4), (2.0, 2.0, 2.0), 4
    ncell = 64
    a = (1.0, 0.15, -0.05)
    cv = np.empty((ncell, 3, n ** 3))
    fv = np.empty((3, ncell * n * n))
    for k in range(3):
        cv[:, k, :] = a[k]
        fv[k] = a[k]
    cv, fv = cv.ravel(), fv.ravel()
    u, lam = _pattern(ncell * 64, 0.43), _pattern(ncell * 64, 0.17)
    assert rel_err(_load(d, "advect_transpose"), ob.dg_advect_var_transpose(3, nc, ij, cv, fv, lam)) < TOL_FP64
    gc, gf = ob.dg_advect_velocity_grad(3, nc, ij, fv, u, lam)
    assert rel_err(_load(d, "advect_grad_cell"), gc) < TOL_FP64
    assert rel_err(_load(d, "advect_grad_face"), gf) < TOL_FP64


@pytest.mark.gpu
def test_precomputed_verification(cuda, outputs):
    _, scal = outputs
    assert scal["precomputed_verification"][0] < 1e-12


@pytest.mark.gpu
def test_continuous_laplacian_class(cuda, outputs):
    d, _ = outputs
    nc = (5, 4, 3)
    ij = (5.0, 4.0, 3.0)
    N = int(np.prod([c * 4 + 1 for c in nc]))
    nq = int(np.prod(nc)) * 125
    kq = np.exp(0.5 * _pattern(nq, 0.31))
    u, lam = _pattern(2 * N, 0.19), _pattern(2 * N, 0.61)
    out = _load(d, "cont_matvec")
    tr = _load(d, "cont_transpose")
    for r in range(2):
        sl = slice(r * N, (r + 1) * N)
        assert rel_err(out[sl], ob.cont_laplacian(4, nc, ij, u[sl].copy(), coef=kq)) < TOL_FP64
        assert rel_err(tr[sl], ob.cont_laplacian(4, nc, ij, lam[sl].copy(), coef=kq)) < TOL_FP64
    assert rel_err(_load(d, "cont_grad"), ob.cont_laplacian_coef_grad(4, nc, ij, u, lam, n_rhs=2)) < TOL_FP64
    x32 = _pattern(N, 0.23).astype(np.float32).astype(np.float64)
    assert rel_err(_load(d, "cont_fp32_matvec"), ob.cont_laplacian(4, nc, ij, x32)) < TOL_FP32


# (driver, arguments, a line of its report)
DRIVER_RUNS = [
    ("test_dg_integrate", ["3", "200000", "3"], "MV degree  3"),
    ("test_dg_precomputed", ["3", "200000", "3"], "Verification"),
    ("test_program", ["4", "3", "100000"], "DoFs/s"),
    ("test_dg_advect", ["100000", "2", "3", "0"], "Error against analytic solution"),
    ("test_dg_advect", ["100000", "2", "3", "1"], "(graph)"),
    # the reference drivers' top degrees (test_dg_integrate.cc:23, test_dg_advect.cc:19)
    ("test_dg_integrate", ["12", "300000", "2"], "MV degree 12"),
    ("test_dg_advect", ["200000", "2", "14", "0"], "Error against analytic solution"),
    ("test_program", ["15", "2", "200000"], "DoFs/s"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,args,marker", DRIVER_RUNS)
def test_drivers_run(cuda, name, args, marker):
    exe = os.path.join(DRIVERS, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built (sum_factorization_kernels_amd/build.py builds the drivers)")
    print("running", name, *args, flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert marker in r.stdout, r.stdout
