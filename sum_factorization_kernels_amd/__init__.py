"""MI355X (gfx950) matrix-free sum-factorization operators.

Python access to libsfk.so (C ABI in include/sfk.h) through ctypes; device
memory comes from torch tensors (plumbing only: every operator runs in the
library's HIP kernels).  Loading fails loudly when the library is missing;
there is no CPU fallback.

    import torch, sum_factorization_kernels_amd as sfk
    op = sfk.Operator(kind=sfk.DG_SIPG, degree=5, n_cells=(32, 32, 32),
                      inv_jac=(32, 32, 32))
    u = torch.randn(op.vector_size, dtype=torch.float64, device="cuda")
    v = op.apply(u)
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libsfk.so")
HEADER = os.path.join(os.path.dirname(PKG_DIR), "include", "sfk.h")

CELL_LAPLACE, CONT_LAPLACE, DG_SIPG, DG_ADVECT, DG_PRECOMP = 0, 1, 2, 3, 4
FP64, FP32 = 0, 1
BASIS_HERMITE, BASIS_GLL = 0, 1
GEOM_CARTESIAN, GEOM_PER_Q = 0, 1
FLAG_REF_X_SWEEP, FLAG_VARIABLE_COEF, FLAG_VARIABLE_VELOCITY, FLAG_REF_PRECOMP_COUPLING = 1, 2, 4, 8

STATUS = {0: "SFK_OK", 1: "SFK_ERR_INVALID_ARG", 2: "SFK_ERR_UNSUPPORTED", 3: "SFK_ERR_HIP",
          4: "SFK_ERR_ALLOC", 5: "SFK_ERR_NONFINITE", 6: "SFK_ERR_NOT_READY", 7: "SFK_ERR_NO_CONVERGENCE"}


class SfkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class OpDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("degree", ctypes.c_int), ("precision", ctypes.c_int),
                ("basis", ctypes.c_int), ("geometry", ctypes.c_int), ("flags", ctypes.c_uint),
                ("n_cells", ctypes.c_int * 3), ("inv_jac", ctypes.c_double * 3),
                ("velocity", ctypes.c_double * 3), ("n_rhs", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double), ("lambda_max", ctypes.c_double),
                ("lambda_min", ctypes.c_double), ("seconds", ctypes.c_double),
                ("matvecs_outer", ctypes.c_int), ("matvecs_inner", ctypes.c_int)]


class Schedule(ctypes.Structure):
    _fields_ = [("cells_per_block", ctypes.c_int), ("block_order", ctypes.c_int),
                ("cont_scatter", ctypes.c_int), ("tile_shape", ctypes.c_int), ("z_cells", ctypes.c_int)]

    def astuple(self):
        return (self.cells_per_block, self.block_order, self.cont_scatter, self.tile_shape, self.z_cells)


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ll = ctypes.c_longlong


def lib():
    """Load libsfk.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not found: build it with "
                               "`python -m sum_factorization_kernels_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.sfk_last_error.restype = ctypes.c_char_p
        L.sfk_version.restype = ctypes.c_char_p
        L.sfk_op_create.argtypes = [ctypes.POINTER(OpDesc), ctypes.POINTER(_vp)]
        L.sfk_op_destroy.argtypes = [_vp]
        L.sfk_op_vector_size.argtypes = [_vp, ctypes.POINTER(_ll)]
        L.sfk_op_qdata_sizes.argtypes = [_vp, ctypes.POINTER(_ll), ctypes.POINTER(_ll)]
        for name in ("sfk_op_set_cell_metric", "sfk_op_set_face_metric", "sfk_op_set_coefficient"):
            getattr(L, name).argtypes = [_vp, _dp, ctypes.c_size_t]
        L.sfk_op_set_velocity.argtypes = [_vp, _dp, ctypes.c_size_t, _dp, ctypes.c_size_t]
        L.sfk_op_set_coefficient_device.argtypes = [_vp, _vp, ctypes.c_size_t, _vp]
        L.sfk_op_set_velocity_device.argtypes = [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _vp]
        L.sfk_op_set_sine_deformation.argtypes = [_vp, ctypes.c_double]
        L.sfk_apply.argtypes = [_vp, _vp, _vp, _vp]
        L.sfk_apply_transpose.argtypes = [_vp, _vp, _vp, _vp]
        L.sfk_cheby_step.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _vp]
        L.sfk_rk_stage.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _vp]
        L.sfk_compute_diagonal.argtypes = [_vp, _vp, _vp]
        L.sfk_op_rhs_constant.argtypes = [_vp, ctypes.c_double, _vp, _vp]
        L.sfk_grad_velocity.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.sfk_grad_coefficient.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.sfk_solve_cg_cheby.argtypes = [_vp, _vp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.POINTER(SolveStats), _vp]
        L.sfk_apply_1d.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _ll, _ll, ctypes.c_int,
                                   _vp, _vp, _vp]
        L.sfk_tables.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.sfk_quadrature.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.sfk_layout_to_canonical.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _vp,
                                              _vp, _vp]
        L.sfk_layout_to_interleaved.argtypes = L.sfk_layout_to_canonical.argtypes
        L.sfk_vec_fill.argtypes = [ctypes.c_int, _vp, _ll, ctypes.c_double, _vp]
        L.sfk_vec_dot.argtypes = [ctypes.c_int, _vp, _vp, _ll, _dp, _vp]
        L.sfk_op_get_schedule.argtypes = [_vp, ctypes.POINTER(Schedule)]
        L.sfk_op_set_schedule.argtypes = [_vp, ctypes.POINTER(Schedule)]
        L.sfk_op_get_smoother_schedule.argtypes = [_vp, ctypes.POINTER(Schedule)]
        L.sfk_op_max_cells_per_block.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
        L.sfk_op_tile_shape.argtypes = [_vp, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
        L.sfk_tune.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(Schedule), _dp, _vp]
        L.sfk_cheby_sweep.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, _vp]
        L.sfk_cheby_precondition.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             _vp]
        L.sfk_stream_create.argtypes = [ctypes.POINTER(_vp)]
        L.sfk_stream_destroy.argtypes = [_vp]
        L.sfk_graph_begin.argtypes = [_vp]
        L.sfk_graph_end.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.sfk_graph_launch.argtypes = [_vp, _vp]
        L.sfk_graph_destroy.argtypes = [_vp]
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise SfkError(status, lib().sfk_last_error().decode())


def _np_ptr(a):
    import numpy as np
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(stream)


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def tables(degree, basis):
    """Host copies of the 1D tables (dense S, D, Sinv, weights, face values)."""
    import numpy as np
    n = degree + 1
    S, D, Si = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
    w, f, hd = np.zeros(n), np.zeros(4 * n), np.zeros(1)
    check(lib().sfk_tables(degree, basis, _np_ptr(S), _np_ptr(D), _np_ptr(Si), _np_ptr(w), _np_ptr(f),
                           _np_ptr(hd)))
    return dict(S=S.reshape(n, n), D=D.reshape(n, n), Sinv=Si.reshape(n, n), w=w,
                fv0=f[:n], fv1=f[n:2 * n], fd0=f[2 * n:3 * n], fd1=f[3 * n:], hermite_d=hd[0])


def quadrature(n):
    import numpy as np
    p, w, g = np.zeros(n), np.zeros(n), np.zeros(n)
    check(lib().sfk_quadrature(n, _np_ptr(p), _np_ptr(w), _np_ptr(g)))
    return p, w, g


class Operator:
    """Owning wrapper of an sfk_op handle."""

    def __init__(self, kind, degree, n_cells, precision=FP64, basis=None, geometry=GEOM_CARTESIAN,
                 flags=0, inv_jac=(1.0, 1.0, 1.0), velocity=(0.0, 0.0, 0.0), n_rhs=1):
        if basis is None:
            basis = BASIS_HERMITE if kind in (CELL_LAPLACE, DG_SIPG, DG_PRECOMP) else BASIS_GLL
        d = OpDesc()
        d.kind, d.degree, d.precision, d.basis, d.geometry, d.flags = kind, degree, precision, basis, geometry, flags
        d.n_cells[:] = list(n_cells)
        d.inv_jac[:] = [float(x) for x in inv_jac]
        d.velocity[:] = [float(x) for x in velocity]
        d.n_rhs = n_rhs
        self.desc = d
        self.handle = ctypes.c_void_p()
        check(lib().sfk_op_create(ctypes.byref(d), ctypes.byref(self.handle)))
        n = _ll()
        check(lib().sfk_op_vector_size(self.handle, ctypes.byref(n)))
        self.vector_size = n.value
        ncq, nfq = _ll(), _ll()
        check(lib().sfk_op_qdata_sizes(self.handle, ctypes.byref(ncq), ctypes.byref(nfq)))
        self.n_cell_q, self.n_face_q = ncq.value, nfq.value
        self.kind, self.degree, self.precision, self.n_rhs = kind, degree, precision, n_rhs

    @property
    def dtype(self):
        import torch
        return torch.float32 if self.precision == FP32 else torch.float64

    def close(self):
        if getattr(self, "handle", None):
            lib().sfk_op_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_vec(self, t, n=None):
        n = self.vector_size * self.n_rhs if n is None else n
        if not (t.is_cuda and t.is_contiguous() and t.dtype == self.dtype and t.numel() == n):
            raise ValueError(f"expected a contiguous {self.dtype} CUDA tensor with {n} elements")

    def new_vector(self):
        import torch
        return torch.zeros(self.vector_size * self.n_rhs, dtype=self.dtype, device="cuda")

    def set_cell_metric(self, a):
        check(lib().sfk_op_set_cell_metric(self.handle, _np_ptr(a), a.size))

    def set_face_metric(self, a):
        check(lib().sfk_op_set_face_metric(self.handle, _np_ptr(a), a.size))

    def set_coefficient(self, a):
        check(lib().sfk_op_set_coefficient(self.handle, _np_ptr(a), a.size))

    def set_velocity(self, cell, face):
        check(lib().sfk_op_set_velocity(self.handle, _np_ptr(cell), cell.size, _np_ptr(face), face.size))

    def _check_data(self, t, n):
        if not (t.is_cuda and t.is_contiguous() and t.dtype == self.dtype and t.numel() == n):
            raise ValueError(f"expected a contiguous CUDA tensor of {n} {self.dtype} values")

    def set_coefficient_device(self, k, stream=None):
        """k: device tensor [cell][q] in the operator's precision"""
        self._check_data(k, self.n_cell_q)
        check(lib().sfk_op_set_coefficient_device(self.handle, _dptr(k), k.numel(), _stream(stream)))

    def set_velocity_device(self, cell, face, stream=None):
        """cell: [cell][3][q], face: [d][cell][q] device tensors in the operator's precision"""
        self._check_data(cell, 3 * self.n_cell_q)
        self._check_data(face, 3 * self.n_face_q)
        check(lib().sfk_op_set_velocity_device(self.handle, _dptr(cell), cell.numel(), _dptr(face), face.numel(),
                                               _stream(stream)))

    def set_sine_deformation(self, amp):
        check(lib().sfk_op_set_sine_deformation(self.handle, float(amp)))

    def apply(self, src, dst=None, stream=None):
        self._check_vec(src)
        dst = self.new_vector() if dst is None else dst
        self._check_vec(dst)
        check(lib().sfk_apply(self.handle, _dptr(src), _dptr(dst), _stream(stream)))
        return dst

    def apply_transpose(self, src, dst=None, stream=None):
        self._check_vec(src)
        dst = self.new_vector() if dst is None else dst
        self._check_vec(dst)
        check(lib().sfk_apply_transpose(self.handle, _dptr(src), _dptr(dst), _stream(stream)))
        return dst

    def cheby_step(self, rhs, sol_old, diag, sol_new, tmp, c_np, c_tm, stream=None):
        for t in (rhs, sol_old, diag, sol_new, tmp):
            self._check_vec(t)
        check(lib().sfk_cheby_step(self.handle, _dptr(rhs), _dptr(sol_old), _dptr(diag), _dptr(sol_new),
                                   _dptr(tmp), float(c_np), float(c_tm), _stream(stream)))

    def rk_stage(self, src, upd, vec_np, vec_tm, c_np, c_tm, stream=None):
        for t in (src, upd, vec_np):
            self._check_vec(t)
        if vec_tm is not None:
            self._check_vec(vec_tm)
        check(lib().sfk_rk_stage(self.handle, _dptr(src), _dptr(upd), _dptr(vec_np), _dptr(vec_tm),
                                 float(c_np), float(c_tm), _stream(stream)))

    def diagonal(self, stream=None):
        d = self.new_vector()[: self.vector_size]
        check(lib().sfk_compute_diagonal(self.handle, _dptr(d), _stream(stream)))
        return d

    def rhs_constant(self, f=1.0, stream=None):
        """b_i = integral of f phi_i (continuous Laplacian, Dirichlet rows 0)"""
        b = self.new_vector()[: self.vector_size]
        check(lib().sfk_op_rhs_constant(self.handle, float(f), _dptr(b), _stream(stream)))
        return b

    def grad_velocity(self, u, lam, stream=None):
        import torch
        self._check_vec(u)
        self._check_vec(lam)
        gc = torch.zeros(3 * self.n_cell_q, dtype=self.dtype, device="cuda")
        gf = torch.zeros(3 * self.n_face_q, dtype=self.dtype, device="cuda")
        check(lib().sfk_grad_velocity(self.handle, _dptr(u), _dptr(lam), _dptr(gc), _dptr(gf), _stream(stream)))
        return gc, gf

    @property
    def schedule(self):
        """(cells_per_block, block_order, cont_scatter, tile_shape, z_cells) of the operator kernel"""
        sc = Schedule()
        check(lib().sfk_op_get_schedule(self.handle, ctypes.byref(sc)))
        return sc.astuple()

    @property
    def smoother_schedule(self):
        """schedule of the tensor-product kernel's fused Chebyshev launches
        (cheby_step, the smoother of solve_cg_cheby): the one tune() picked
        for them, else the matvec schedule"""
        sc = Schedule()
        check(lib().sfk_op_get_smoother_schedule(self.handle, ctypes.byref(sc)))
        return sc.astuple()

    def set_schedule(self, cells_per_block=0, block_order=1, cont_scatter=-1, tile_shape=0, z_cells=0):
        sc = Schedule(int(cells_per_block), int(block_order), int(cont_scatter), int(tile_shape), int(z_cells))
        check(lib().sfk_op_set_schedule(self.handle, ctypes.byref(sc)))

    def tile_shape(self, shape):
        """(count, cells in x, cells in y) of tile shape `shape` of the continuous tensor-product kernel"""
        c, bx, by = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().sfk_op_tile_shape(self.handle, int(shape), ctypes.byref(c), ctypes.byref(bx), ctypes.byref(by)))
        return c.value, bx.value, by.value

    @property
    def max_cells_per_block(self):
        m = ctypes.c_int()
        check(lib().sfk_op_max_cells_per_block(self.handle, ctypes.byref(m)))
        return m.value

    def tune(self, repeats=10, stream=None):
        """Time every candidate schedule, keep the fastest; returns (schedule, seconds per apply)."""
        sc = Schedule()
        t = ctypes.c_double()
        check(lib().sfk_tune(self.handle, int(repeats), ctypes.byref(sc), ctypes.byref(t), _stream(stream)))
        return sc.astuple(), t.value

    def cheby_sweep(self, rhs, x, spare, diag_inv, tmp, degree, lambda_min, lambda_max, stream=None):
        """`degree` fused Chebyshev steps on [lambda_min, lambda_max]; x updated in place"""
        for t in (rhs, x, spare, diag_inv, tmp):
            self._check_vec(t)
        check(lib().sfk_cheby_sweep(self.handle, _dptr(rhs), _dptr(x), _dptr(spare), _dptr(diag_inv), _dptr(tmp),
                                    int(degree), float(lambda_min), float(lambda_max), _stream(stream)))

    def cheby_precondition(self, rhs, degree, lambda_min, lambda_max, stream=None):
        """z = P(rhs): the solver's Chebyshev preconditioner (three-term form on
        the tensor-product kernel); equals cheby_sweep from x = 0"""
        self._check_vec(rhs)
        z, spare = self.new_vector(), self.new_vector()
        check(lib().sfk_cheby_precondition(self.handle, _dptr(rhs), _dptr(z), _dptr(spare), int(degree),
                                           float(lambda_min), float(lambda_max), _stream(stream)))
        return z

    def grad_coefficient(self, u, lam, stream=None):
        import torch
        self._check_vec(u)
        self._check_vec(lam)
        g = torch.zeros(self.n_cell_q, dtype=self.dtype, device="cuda")
        check(lib().sfk_grad_coefficient(self.handle, _dptr(u), _dptr(lam), _dptr(g), _stream(stream)))
        return g


class Stream:
    """A non-default HIP stream (default priority) owned by the library."""

    def __init__(self):
        self.handle = _vp()
        check(lib().sfk_stream_create(ctypes.byref(self.handle)))

    @property
    def ptr(self):
        return self.handle.value

    def synchronize(self):
        check(lib().sfk_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().sfk_stream_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """sfk calls issued on `stream` inside `with Graph(stream) as g:` are
    captured; g.launch() replays them."""

    def __init__(self, stream):
        self.stream = stream
        self.handle = _vp()

    def __enter__(self):
        check(lib().sfk_graph_begin(self.stream.handle))
        return self

    def __exit__(self, exc_type, exc, tb):
        status = lib().sfk_graph_end(self.stream.handle, ctypes.byref(self.handle))
        if exc_type is None:
            check(status)
        return False

    def launch(self):
        check(lib().sfk_graph_launch(self.handle, self.stream.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().sfk_graph_destroy(self.handle)
        except Exception:
            pass


def _check_dev(t, dtypes, n=None, min_n=None, name="tensor"):
    """a contiguous CUDA tensor of one of `dtypes` with n (or at least min_n)
    elements: the C ABI takes raw device pointers and cannot check them"""
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous() or t.dtype not in dtypes:
        raise ValueError(f"{name}: expected a contiguous CUDA tensor of dtype {' or '.join(map(str, dtypes))}")
    if n is not None and t.numel() != n:
        raise ValueError(f"{name}: expected {n} elements, got {t.numel()}")
    if min_n is not None and t.numel() < min_n:
        raise ValueError(f"{name}: expected at least {min_n} elements, got {t.numel()}")


def solve_cg_cheby(op, smoother, b, x, rel_tol=1e-10, max_iter=1000, cheby_degree=6, lanczos_steps=12,
                   stream=None):
    """FP64 PCG with a Chebyshev point-Jacobi preconditioner on `smoother`."""
    import torch
    _check_dev(b, (torch.float64,), op.vector_size, name="b")
    _check_dev(x, (torch.float64,), op.vector_size, name="x")
    st = SolveStats()
    check(lib().sfk_solve_cg_cheby(op.handle, smoother.handle, ctypes.cast(b.data_ptr(), _dp),
                                   ctypes.cast(x.data_ptr(), _dp), rel_tol, max_iter, cheby_degree,
                                   lanczos_steps, ctypes.byref(st), _stream(stream)))
    return st


def _layout_size(degree, n_cells):
    return int(n_cells[0]) * int(n_cells[1]) * int(n_cells[2]) * (int(degree) + 1) ** 3


def layout_to_canonical(degree, n_cells, lanes, x, stream=None):
    import torch
    _check_dev(x, (torch.float64,), _layout_size(degree, n_cells), name="x")
    out = torch.empty_like(x)
    nc = (ctypes.c_int * 3)(*n_cells)
    check(lib().sfk_layout_to_canonical(degree, nc, lanes, _dptr(x), _dptr(out), _stream(stream)))
    return out


def layout_to_interleaved(degree, n_cells, lanes, x, stream=None):
    import torch
    _check_dev(x, (torch.float64,), _layout_size(degree, n_cells), name="x")
    out = torch.empty_like(x)
    nc = (ctypes.c_int * 3)(*n_cells)
    check(lib().sfk_layout_to_interleaved(degree, nc, lanes, _dptr(x), _dptr(out), _stream(stream)))
    return out


def apply_1d(matrix, sym, forward, x, n_lines, line_stride, stride, stream=None):
    import numpy as np
    import torch
    nn = matrix.shape[0]
    if n_lines < 0 or line_stride < 0 or stride < 1:
        raise ValueError("apply_1d: n_lines, line_stride >= 0 and stride >= 1 expected")
    extent = (n_lines - 1) * line_stride + (nn - 1) * stride + 1 if n_lines > 0 else 0
    _check_dev(x, (torch.float64,), min_n=extent, name="x")
    out = torch.zeros_like(x)
    m = np.ascontiguousarray(matrix, dtype=np.float64).ravel()
    check(lib().sfk_apply_1d(nn, _np_ptr(m), sym, int(forward), n_lines, line_stride, stride, _dptr(x),
                             _dptr(out), _stream(stream)))
    return out


def vec_dot(x, y, stream=None):
    import torch
    _check_dev(x, (torch.float32, torch.float64), name="x")
    _check_dev(y, (x.dtype,), x.numel(), name="y")
    r = ctypes.c_double()
    prec = FP32 if x.dtype == torch.float32 else FP64
    check(lib().sfk_vec_dot(prec, _dptr(x), _dptr(y), x.numel(), ctypes.byref(r), _stream(stream)))
    return r.value
