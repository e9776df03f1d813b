"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference harness (oracle/_ref/libsfkref.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module.  Vectors use the canonical
layout (cell-major, x-fastest) unless a name says "interleaved".
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsfkref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle oracle`")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.sfko_jacobi_polynomial.restype = ctypes.c_double
        lib.sfko_lagrange_value.restype = ctypes.c_double
        lib.sfko_lagrange_derivative.restype = ctypes.c_double
        lib.sfko_hermite_value.restype = ctypes.c_double
        lib.sfko_hermite_derivative.restype = ctypes.c_double
        for name in ("sfko_lagrange_value", "sfko_lagrange_derivative"):
            getattr(lib, name).argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double]
        for name in ("sfko_hermite_value", "sfko_hermite_derivative"):
            getattr(lib, name).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        lib.sfko_jacobi_polynomial.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _oracle = lib
    return _oracle


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        lib.ref_hermite_value.restype = ctypes.c_double
        lib.ref_hermite_derivative.restype = ctypes.c_double
        lib.ref_hermite_value.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        lib.ref_hermite_derivative.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        lib.ref_apply_1d.argtypes = [ctypes.c_int] * 6 + [_dp] * 6
        lib.ref_dg_laplacian.argtypes = [ctypes.c_int, _up, _dp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                         ctypes.c_double, ctypes.c_double, _dp, _dp]
        lib.ref_dg_advect.argtypes = [ctypes.c_int, _up, _dp, ctypes.c_int, _dp, _dp, _dp,
                                      ctypes.c_double, ctypes.c_double, _dp, _dp]
        lib.ref_advect_kat.argtypes = [ctypes.c_int, _up, ctypes.c_int, _dp, _dp, _dp, _dp]
        lib.ref_dg_precomputed.argtypes = [ctypes.c_int, _up, _dp, _dp, _dp, _dp, _dp]
        _ref = lib
    return _ref


# ----------------------------------------------------------------- oracle
def gauss_points(n):
    out = np.zeros(n)
    oracle().sfko_gauss_points(n, _ptr(out))
    return out


def gauss_weights(n):
    out = np.zeros(n)
    oracle().sfko_gauss_weights(n, _ptr(out))
    return out


def gll_points(n):
    out = np.zeros(n)
    oracle().sfko_gll_points(n, _ptr(out))
    return out


def eo_tables(degree, basis):
    n = degree + 1
    h = (n + 1) // 2
    v, g, inv = np.zeros(n * h), np.zeros(n * h), np.zeros(n * h)
    f, hd = np.zeros(2 * n), np.zeros(1)
    rc = oracle().sfko_export_eo_tables(degree, basis, _ptr(v), _ptr(g), _ptr(inv), _ptr(f), _ptr(hd))
    assert rc == 0
    return dict(values_eo=v, gradients_eo=g, inv_values_eo=inv, face_eo=f, hermite_d=hd[0])


def apply_1d_eo(nn, stride, typ, contract_over_rows, add_into, do_dg, coeffs, inp, out, add=None,
                dg=None, face=None):
    oracle().sfko_apply_1d_eo(nn, stride, typ, int(contract_over_rows), int(add_into), do_dg,
                              _ptr(coeffs), _ptr(inp), _ptr(out), _ptr(add), _ptr(dg), _ptr(face))


def cell_laplacian(degree, basis, src, inv_jac):
    n3 = (degree + 1) ** 3
    dst = np.zeros_like(src)
    oracle().sfko_cell_laplacian(degree, basis, ctypes.c_long(src.size // n3), _d3(inv_jac),
                                 _ptr(src), _ptr(dst))
    return dst


def dg_sipg(degree, nc, inv_jac, src):
    dst = np.zeros_like(src)
    oracle().sfko_dg_sipg(degree, _i3(nc), _d3(inv_jac), _ptr(src), _ptr(dst))
    return dst


def dg_sipg_cheby(degree, nc, inv_jac, rhs, sol_old, diag, tm, c_np, c_tm):
    sol_new = np.zeros_like(sol_old)
    tm = tm.copy()
    oracle().sfko_dg_sipg_cheby(degree, _i3(nc), _d3(inv_jac), _ptr(rhs), _ptr(sol_old), _ptr(diag),
                                _ptr(sol_new), _ptr(tm), ctypes.c_double(c_np), ctypes.c_double(c_tm))
    return sol_new, tm


def dg_advect(degree, nc, inv_jac, vel, src):
    dst = np.zeros_like(src)
    oracle().sfko_dg_advect(degree, _i3(nc), _d3(inv_jac), _d3(vel), _ptr(src), _ptr(dst))
    return dst


def dg_advect_stage(degree, nc, inv_jac, vel, src, upd, tm, c_np, c_tm):
    vnp = np.zeros_like(src)
    tm = tm.copy()
    oracle().sfko_dg_advect_stage(degree, _i3(nc), _d3(inv_jac), _d3(vel), _ptr(src), _ptr(upd),
                                  _ptr(vnp), _ptr(tm), ctypes.c_double(c_np), ctypes.c_double(c_tm))
    return vnp, tm


def to_canonical(degree, nc, lanes, x):
    out = np.zeros_like(x)
    oracle().sfko_interleaved_to_canonical(degree, _i3(nc), lanes, _ptr(x), _ptr(out))
    return out


def to_interleaved(degree, nc, lanes, x):
    out = np.zeros_like(x)
    oracle().sfko_canonical_to_interleaved(degree, _i3(nc), lanes, _ptr(x), _ptr(out))
    return out


# -------------------------------------------------------------- reference
def ref_lanes():
    return ref().ref_n_lanes()


def ref_tables(kind, degree):
    n = degree + 1
    h = (n + 1) // 2
    v, g, inv = np.zeros(n * h), np.zeros(n * h), np.zeros(n * h)
    f, hd = np.zeros(2 * n), np.zeros(1)
    rc = ref().ref_tables(kind, degree, _ptr(v), _ptr(g), _ptr(f), _ptr(inv), _ptr(hd))
    assert rc == 0
    return dict(values_eo=v, gradients_eo=g, inv_values_eo=inv, face_eo=f, hermite_d=hd[0])


def ref_cell_laplacian(degree, src_interleaved, inv_jac):
    L = ref_lanes()
    n3 = (degree + 1) ** 3
    nb = src_interleaved.size // (n3 * L)
    dst = np.zeros_like(src_interleaved)
    rc = ref().ref_cell_laplacian(degree, nb, _ptr(np.asarray(inv_jac, dtype=np.float64)),
                                  _ptr(src_interleaved), _ptr(dst))
    assert rc == 0
    return dst


def ref_dg_laplacian(degree, nc, inv_jac, sol_old, rhs=None, diag=None, tm=None, c_np=0.0, c_tm=0.0,
                     mode=0):
    ncu = (ctypes.c_uint * 3)(*nc)
    out_new = np.zeros_like(sol_old)
    out_tm = np.zeros_like(sol_old)
    ij = None if inv_jac is None else np.asarray(inv_jac, dtype=np.float64)
    rc = ref().ref_dg_laplacian(degree, ncu, _ptr(ij), mode, _ptr(sol_old), _ptr(rhs), _ptr(diag),
                                _ptr(tm), c_np, c_tm, _ptr(out_new), _ptr(out_tm))
    assert rc == 0
    return out_new, out_tm


def ref_dg_advect(degree, nc, inv_jac, src, upd=None, tm=None, c_np=0.0, c_tm=0.0, mode=0):
    ncu = (ctypes.c_uint * 3)(*nc)
    out_np = np.zeros_like(src)
    out_tm = np.zeros_like(src)
    ij = None if inv_jac is None else np.asarray(inv_jac, dtype=np.float64)
    rc = ref().ref_dg_advect(degree, ncu, _ptr(ij), mode, _ptr(src), _ptr(upd), _ptr(tm), c_np, c_tm,
                             _ptr(out_np), _ptr(out_tm))
    assert rc == 0
    return out_np, out_tm


def ref_dg_precomputed(degree, nc, src_interleaved=None):
    """reference PrecomputedDGLaplacian: (dst interleaved or None, mass_eo,
    laplace_eo, (value_outer_1, value_outer_2))"""
    n = degree + 1
    h = (n + 1) // 2
    ncu = (ctypes.c_uint * 3)(*nc)
    m, lap, outer = np.zeros(n * h), np.zeros(n * h), np.zeros(2)
    dst = None if src_interleaved is None else np.zeros_like(src_interleaved)
    rc = ref().ref_dg_precomputed(degree, ncu, _ptr(src_interleaved), _ptr(dst), _ptr(m), _ptr(lap), _ptr(outer))
    assert rc == 0
    return dst, m, lap, (outer[0], outer[1])


# ------------------------------------------------------- extended oracle
def precomp_1d(degree, ref_coupling=False):
    """1D mass and Laplace(+penalty) matrices [n, n] and the coupling (v1, v2)"""
    n = degree + 1
    m, lap, v = np.zeros(n * n), np.zeros(n * n), np.zeros(2)
    oracle().sfko_precomp_1d(degree, int(ref_coupling), _ptr(m), _ptr(lap), _ptr(v))
    return m.reshape(n, n), lap.reshape(n, n), (v[0], v[1])


def dg_precomputed(degree, nc, src, inv_jac=(1.0, 1.0, 1.0), ref_coupling=False):
    dst = np.zeros_like(src)
    oracle().sfko_dg_precomputed(degree, _i3(nc), _d3(inv_jac), int(ref_coupling), _ptr(src), _ptr(dst))
    return dst


def cont_laplacian(degree, nc, inv_jac, src, coef=None):
    dst = np.zeros_like(src)
    oracle().sfko_cont_laplacian(degree, _i3(nc), _d3(inv_jac), _ptr(coef), _ptr(src), _ptr(dst))
    return dst


def cont_laplacian_coef_grad(degree, nc, inv_jac, u_all, lam_all, n_rhs=1):
    n3 = (degree + 1) ** 3
    g = np.zeros(int(np.prod(nc)) * n3)
    oracle().sfko_cont_laplacian_coef_grad(degree, _i3(nc), _d3(inv_jac), n_rhs, _ptr(u_all), _ptr(lam_all),
                                           _ptr(g))
    return g


def cartesian_metric(degree, nc, inv_jac):
    n = degree + 1
    ncell = int(np.prod(nc))
    cm = np.zeros(ncell * 6 * n ** 3)
    fm = np.zeros(3 * ncell * 7 * n * n)
    oracle().sfko_cartesian_metric(degree, _i3(nc), _d3(inv_jac), _ptr(cm), _ptr(fm))
    return cm, fm


def dg_sipg_general(degree, nc, cm, fm, src):
    dst = np.zeros_like(src)
    oracle().sfko_dg_sipg_general(degree, _i3(nc), _ptr(cm), _ptr(fm), _ptr(src), _ptr(dst))
    return dst


def dg_advect_var(degree, nc, inv_jac, cv, fv, src):
    dst = np.zeros_like(src)
    oracle().sfko_dg_advect_var(degree, _i3(nc), _d3(inv_jac), _ptr(cv), _ptr(fv), _ptr(src), _ptr(dst))
    return dst


def dg_advect_var_transpose(degree, nc, inv_jac, cv, fv, src):
    dst = np.zeros_like(src)
    oracle().sfko_dg_advect_var_transpose(degree, _i3(nc), _d3(inv_jac), _ptr(cv), _ptr(fv), _ptr(src),
                                          _ptr(dst))
    return dst


def dg_mass_inverse(degree, nc, inv_jac, src):
    dst = np.zeros_like(src)
    oracle().sfko_dg_mass_inverse(degree, _i3(nc), _d3(inv_jac), _ptr(src), _ptr(dst))
    return dst


def dg_advect_velocity_grad(degree, nc, inv_jac, fv, u, lam):
    n = degree + 1
    ncell = int(np.prod(nc))
    gc = np.zeros(ncell * 3 * n ** 3)
    gf = np.zeros(3 * ncell * n * n)
    oracle().sfko_dg_advect_velocity_grad(degree, _i3(nc), _d3(inv_jac), _ptr(fv), _ptr(u), _ptr(lam), _ptr(gc),
                                          _ptr(gf))
    return gc, gf
