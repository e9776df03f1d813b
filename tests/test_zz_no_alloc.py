"""The allocation-free inner product (sfk_vec_dot_noalloc) against the
allocating sfk_vec_dot, through the C ABI: two calls in a row (the second on
a shorter vector, so stale partial sums of the first would show), FP64 and
FP32, the empty vector, and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sfk = pytest.importorskip("sum_factorization_kernels_amd")

MAX_N = 4096 * 65536


def _lib():
    L = sfk.lib()
    L.sfk_vec_dot_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                      ctypes.c_void_p, ctypes.c_void_p]
    return L


def _dot_alloc(prec, x, y, n):
    out = ctypes.c_double()
    sfk.check(sfk.lib().sfk_vec_dot(prec, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), n,
                                    ctypes.byref(out), ctypes.c_void_p(0)))
    return out.value


def _dot_noalloc(prec, x, y, n, res):
    import torch
    sfk.check(_lib().sfk_vec_dot_noalloc(prec, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), n,
                                         ctypes.c_void_p(res.data_ptr()), ctypes.c_void_p(0)))
    torch.cuda.synchronize()
    return float(res.cpu()[0])


@pytest.mark.parametrize("prec", [sfk.FP64, sfk.FP32])
def test_noalloc_dot_matches_allocating_dot(cuda, prec):
    import torch
    dt = torch.float64 if prec == sfk.FP64 else torch.float32
    g = torch.Generator("cuda").manual_seed(9001)
    res = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    # first call: 3.1M elements (48 chunks); second: a shorter vector whose
    # last chunk is partial
    for n in (3_100_000, 70_001, 1):
        x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g).to(dt)
        y = torch.randn(n, dtype=torch.float64, device="cuda", generator=g).to(dt)
        ref = _dot_alloc(prec, x, y, n)
        got = _dot_noalloc(prec, x, y, n, res)
        exact = float(np.dot(x.double().cpu().numpy(), y.double().cpu().numpy()))
        scale = float((x.double().abs() * y.double().abs()).sum())
        assert abs(got - ref) <= 1e-12 * scale, (n, got, ref)
        assert abs(got - exact) <= 1e-12 * scale, (n, got, exact)
        # the same inputs give the same bits
        assert _dot_noalloc(prec, x, y, n, res) == got


def test_noalloc_dot_empty_and_limit(cuda):
    import torch
    x = torch.ones(4, dtype=torch.float64, device="cuda")
    res = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    assert _dot_noalloc(sfk.FP64, x, x, 0, res) == 0.0
    # beyond the static array: rejected before anything is launched (the
    # pointers are never read)
    res.fill_(float("nan"))
    torch.cuda.synchronize()
    st = _lib().sfk_vec_dot_noalloc(sfk.FP64, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                    MAX_N + 1, ctypes.c_void_p(res.data_ptr()), ctypes.c_void_p(0))
    assert st == 1  # SFK_ERR_INVALID_ARG
    assert b"exceeds the maximum" in sfk.lib().sfk_last_error()
    torch.cuda.synchronize()
    assert torch.isnan(res).all()  # nothing written
