"""The allocation-free conservation entry point (aero_lbm_get_conservation_rect_noalloc)
against aero_lbm_get_conservation_rect: identical sums (the same kernels and
summation order, only the buffers differ), on two calls in a row, the second
on a smaller context, so the fixed device scratch is reused with fewer partials."""
import ctypes
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu


def bind(lib):
    fn = lib.aero_lbm_get_conservation_rect_noalloc
    fn.argtypes = [ctypes.c_int] * 3 + [ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    return fn


def sums(lib, fn, dims, key):
    out = (ctypes.c_double * 4)()
    A.check(getattr(lib, fn)(*dims, key, out), fn)
    return np.array(out[:])


def make(lib, n, key, mode):
    cfg, pk, _ = W.config1(n)
    A.check(lib.aero_lbm_init_rect(n, n, n, 11, 4), "init")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(W.reference_overrides(A.KERNEL_FAST))), "overrides")
    lib.aero_lbm_set_realtime_solver_mode(mode)
    ctx = A.Context(n, n, n, key=key)
    ctx.step(pk)
    for _ in range(4):
        ctx.step(None)
    return ctx


@pytest.mark.parametrize("mode", [A.SOLVER_D3Q27_FP16_INPLACE, A.SOLVER_CLASSIC_D3Q27])
def test_noalloc_conservation_equals_allocating_entry(mode):
    lib = A.load()
    bind(lib)
    # 84^3 cells need more workgroups than the 2048 partials (grid-stride loop), 20^3 far fewer
    for n, key in ((84, 91), (20, 92)):
        print(f"noalloc conservation mode={mode} n={n}", flush=True)
        ctx = make(lib, n, key, mode)
        want = sums(lib, "aero_lbm_get_conservation_rect", (n, n, n), key)
        got = sums(lib, "aero_lbm_get_conservation_rect_noalloc", (n, n, n), key)
        again = sums(lib, "aero_lbm_get_conservation_rect_noalloc", (n, n, n), key)
        assert np.array_equal(got, want), (got, want)
        assert np.array_equal(again, want)
        assert abs(want[0]) > 0.5 * (n ** 3) * 0.8
        ctx.release()


def test_noalloc_conservation_errors():
    lib = A.load()
    fn = bind(lib)
    out = (ctypes.c_double * 4)()
    assert fn(8, 8, 8, 123456, out) == 0
    assert "no initialised context" in A.last_error()
    assert fn(8, 8, 8, 123456, None) == 0
    assert "null output" in A.last_error()
