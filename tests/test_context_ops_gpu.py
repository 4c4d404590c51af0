"""Context operations of the core C ABI on the GPU, against the oracle's restatement of
the reference (tests/test_oracle_context_ops.py pins it):

* aero_lbm_patch_d3q27_f16_static_cells_rect on an in-place context, after an even
  and after an odd step (the two A-A layouts), with cells turning solid and fluid,
  neighbours of each other, a repeated entry and an unchanged one: the FP32 in-place
  path equals the FP32 oracle given the same patch within 1e-4 of the inflow speed
  (fast kernels: FMA and hardware reciprocal), FP16 storage within 1e-3;
* aero_lbm_step_rect_with_sparse_overlays: bit-identical to the packet with the
  overlay values written in, on the exact kernels, and to the oracle;
* aero_lbm_shift_context and aero_lbm_exchange_halo(_layers_rect) on the exact
  kernels: bit-identical to the oracle through the following steps; in-place
  contexts are left alone by the exchange, as the reference's exchange leaves its
  in-place buffers (native/src/aero_lbm_jni.cpp:7736-7739)."""
import ctypes
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

U_IN = 0.05


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def setup(lib, dims, cfg, variant, mode):
    A.check(lib.aero_lbm_init_rect(*dims, 11, 4), "init_rect")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(W.reference_overrides(variant))), "overrides")
    lib.aero_lbm_set_realtime_solver_mode(mode)


def channel_case(dims, seed, thermal=False):
    cfg = W.channel_config(U_IN, 50.0, sgs=True, buoyancy=thermal)
    solid = W.solid_boxes(*dims, 0.12, seed=seed, edge=(2, 4), keep_clear_x=2)
    t = None
    if thermal:
        t = W.hot_blob(*dims, 0.02, 3.0, seed=seed + 1, axis=2)
        t[solid.astype(bool)] = 0.0
    pk = W.make_packet(solid, u0=(U_IN, 0.01, 0.0), perturb=1e-3, seed_u=seed + 2, temperature=t)
    return cfg, pk, solid, t


def patch_list(solid):
    """Cells turning solid and fluid (neighbours among them), a repeated entry and an
    unchanged one, including cells on the periodic y / z faces."""
    nx, ny, nz = solid.shape
    s = solid.astype(bool)
    cells, vals = [], []

    def add(x, y, z, v):
        cells.append((x * ny + y) * nz + z)
        vals.append(v)

    fluid = np.argwhere(~s)
    solids = np.argwhere(s)
    # a 2 x 2 x 1 block of fluid cells turning solid, in the middle of the flow
    x0, y0, z0 = nx // 2, ny // 2, nz // 2
    for dx in (0, 1):
        for dy in (0, 1):
            if not s[x0 + dx, y0 + dy, z0]:
                add(x0 + dx, y0 + dy, z0, 1)
    # solid cells on box surfaces turning fluid (with a solid neighbour left in place)
    picked = 0
    for x, y, z in solids:
        if 2 < x < nx - 3 and picked < 6:
            add(x, y, z, 0)
            picked += 1
    # fluid cells next to a solid turning solid
    picked = 0
    for x, y, z in fluid:
        if 2 < x < nx - 3 and picked < 5 and s[x, (y + 1) % ny, z]:
            add(x, y, z, 1)
            picked += 1
    # on the periodic faces
    add(nx // 3, 0, 0, 1 - int(s[nx // 3, 0, 0]))
    add(nx // 3 + 1, ny - 1, nz - 1, 1 - int(s[nx // 3 + 1, ny - 1, nz - 1]))
    # repeated: a fluid cell set solid then fluid again (restarts), an unchanged entry
    fx, fy, fz = fluid[len(fluid) // 3]
    add(fx, fy, fz, 1)
    add(fx, fy, fz, 0)
    gx, gy, gz = fluid[len(fluid) // 2]
    add(gx, gy, gz, 0)
    return np.array(cells, np.int32), np.array(vals, np.uint8)


def patch_gpu(lib, dims, key, cells, vals, fans=None):
    fans = np.zeros(len(cells), np.uint8) if fans is None else fans
    return lib.aero_lbm_patch_d3q27_f16_static_cells_rect(
        *dims, key, len(cells), cells.ctypes.data_as(A._I), vals.ctypes.data_as(A._U8), fans.ctypes.data_as(A._U8))


@pytest.mark.parametrize("before", [3, 4])
@pytest.mark.parametrize("mode,bar", [(A.SOLVER_D3Q27_FP32_INPLACE, 1e-4), (A.SOLVER_D3Q27_FP16_INPLACE, 1e-3)])
def test_patch_in_place_equals_oracle(oracle_mod, mode, bar, before):
    lib = A.load()
    dims = (20, 14, 16)
    cfg, pk, solid, _ = channel_case(dims, seed=21)
    setup(lib, dims, cfg, A.KERNEL_FAST, mode)
    ctx = A.Context(*dims, key=81)
    o = oracle_mod.Oracle(*dims, cfg, precision=64)
    ref = oracle_mod.Oracle(*dims, cfg, precision=64)  # the same run without the patch
    cells, vals = patch_list(solid)
    log(f"patch mode {mode} after {before} steps: {len(cells)} entries")
    try:
        for i in range(before):
            ctx.step(pk if i == 0 else None)
            o.step(pk if i == 0 else None)
            ref.step(pk if i == 0 else None)
        A.check(patch_gpu(lib, dims, ctx.key, cells, vals), "patch")
        o.patch_cells(cells, vals)
        for _ in range(7):
            ctx.step(None)
            o.step(None)
            ref.step(None)
        g = ctx.flow().reshape(-1, 4)
        r = o.flow().reshape(-1, 4)
        err = float(np.max(np.abs(g[:, :3] - r[:, :3]))) / U_IN
        effect = float(np.max(np.abs(r[:, :3] - ref.flow().reshape(-1, 4)[:, :3]))) / U_IN
        log(f"  |u_gpu - u_oracle| / U_in = {err:.2e}; the patch moved the flow by {effect:.2e}")
        assert err <= bar, err
        assert effect > 20 * bar  # the comparison sees the patch
        now_solid = np.zeros(solid.size, bool)
        now_solid[:] = solid.reshape(-1) != 0
        for c, v in zip(cells, vals):
            now_solid[c] = v != 0
        assert np.all(g[now_solid] == 0.0)
    finally:
        ctx.release()
        lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_CLASSIC_D3Q27)


def test_patch_refusals():
    lib = A.load()
    dims = (12, 10, 8)
    cfg, pk, solid, _ = channel_case(dims, seed=5)
    setup(lib, dims, cfg, A.KERNEL_EXACT, A.SOLVER_CLASSIC_D3Q27)
    exact = A.Context(*dims, key=82)
    exact.step(pk)
    cells = np.array([5], np.int32)
    vals = np.array([1], np.uint8)
    assert patch_gpu(lib, dims, exact.key, cells, vals) == 0
    assert b"in-place" in lib.aero_lbm_last_error()
    exact.release()
    setup(lib, dims, cfg, A.KERNEL_FAST, A.SOLVER_D3Q27_FP16_INPLACE)
    ctx = A.Context(*dims, key=83)
    assert patch_gpu(lib, dims, ctx.key, cells, vals) == 0  # not stepped yet
    ctx.step(pk)
    assert patch_gpu(lib, dims, 9999, cells, vals) == 0
    assert b"context not found" in lib.aero_lbm_last_error()
    assert patch_gpu(lib, (12, 10, 9), ctx.key, cells, vals) == 0
    assert b"invalid dimensions" in lib.aero_lbm_last_error()
    # fan direction codes while fan forcing is on
    cfg_fan = W.channel_config(U_IN, 50.0, sgs=True)
    cfg_fan.flags &= ~A.FLAG_DISABLE_FAN_FORCING
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg_fan)), "config")
    fluid = int(np.flatnonzero(solid.reshape(-1) == 0)[3])
    assert patch_gpu(lib, dims, ctx.key, np.array([fluid], np.int32), np.array([0], np.uint8),
                     np.array([2], np.uint8)) == 0
    assert b"fan" in lib.aero_lbm_last_error()
    # with fan forcing off the codes are moot; an empty patch is a no-op
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
    A.check(patch_gpu(lib, dims, ctx.key, np.array([fluid], np.int32), np.array([0], np.uint8),
                      np.array([2], np.uint8)), "patch")
    assert lib.aero_lbm_patch_d3q27_f16_static_cells_rect(*dims, ctx.key, 0, None, None, None) == 1
    assert lib.aero_lbm_context_compact_initialized(ctx.key) == 0
    ctx.release()
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_CLASSIC_D3Q27)


def test_sparse_overlays_equal_the_merged_packet(oracle_mod):
    lib = A.load()
    dims = (16, 12, 10)
    cfg, pk, solid, t = channel_case(dims, seed=31, thermal=True)
    setup(lib, dims, cfg, A.KERNEL_EXACT, A.SOLVER_CLASSIC_D3Q27)
    rng = np.random.default_rng(4)
    cells = np.sort(rng.choice(solid.size, 40, replace=False)).astype(np.int32)
    vals = np.zeros((len(cells), 6), np.float32)
    vals[:, 0] = rng.uniform(0.02, 0.08, len(cells))
    vals[:, 1:3] = rng.uniform(-0.01, 0.01, (len(cells), 2))
    vals[:, 3] = rng.uniform(-0.002, 0.002, len(cells))
    vals[:, 4] = rng.uniform(-0.5, 0.5, len(cells))
    vals[:, 5] = rng.uniform(-0.004, 0.004, len(cells))
    merged = pk.reshape(-1, 11).copy()
    merged[cells, 5:9] = vals[:, :4]
    merged[cells, 9] = vals[:, 5]
    merged[cells, 10] = vals[:, 4]
    merged = merged.reshape(-1)
    a, b = A.Context(*dims, key=84), A.Context(*dims, key=85)
    o = oracle_mod.Oracle(*dims, cfg, precision=32)
    out = np.zeros(solid.size * 4, np.float32)
    try:
        for ctx in (a, b):
            ctx.set_temperature(t)
        o.set_temperature(t)
        A.check(lib.aero_lbm_step_rect_with_sparse_overlays(A.fptr(pk), *dims, a.key, cells.ctypes.data_as(A._I),
                                                            A.fptr(vals.reshape(-1)), len(cells), A.fptr(out)),
                "overlay step")
        b.step(merged)
        o.step(merged)
        for _ in range(3):
            a.step(None)
            b.step(None)
            o.step(None)
        assert np.array_equal(a.flow(), b.flow())
        assert np.array_equal(a.temperature(), b.temperature())
        assert np.array_equal(a.flow(), o.flow())
        assert np.array_equal(a.temperature(), o.temperature())
        assert lib.aero_lbm_step_rect_with_sparse_overlays(A.fptr(pk), *dims, a.key, None, None, 3, None) == 0
        assert b"missing overlay data" in lib.aero_lbm_last_error()
        assert lib.aero_lbm_step_rect_with_sparse_overlays(A.fptr(pk), *dims, a.key, None, None, -1, None) == 0
    finally:
        a.release()
        b.release()


@pytest.mark.parametrize("shift", [(2, -1, 1), (-3, 0, 0)])
def test_shift_context_equals_oracle(oracle_mod, shift):
    lib = A.load()
    n = 16
    dims = (n, n, n)
    cfg, pk, solid, t = channel_case(dims, seed=41, thermal=True)
    setup(lib, dims, cfg, A.KERNEL_EXACT, A.SOLVER_CLASSIC_D3Q27)
    ctx = A.Context(*dims, key=86)
    o = oracle_mod.Oracle(*dims, cfg, precision=32)
    try:
        ctx.set_temperature(t)
        o.set_temperature(t)
        for i in range(3):
            ctx.step(pk if i == 0 else None)
            o.step(pk if i == 0 else None)
        log(f"shift {shift}")
        A.check(lib.aero_lbm_shift_context(n, ctx.key, *shift), "shift")
        o.shift(*shift)
        for _ in range(4):
            ctx.step(None)
            o.step(None)
        assert np.array_equal(ctx.flow(), o.flow())
        assert np.array_equal(ctx.temperature(), o.temperature())
        assert lib.aero_lbm_shift_context(n, ctx.key, n, 0, 0) == 0
        assert lib.aero_lbm_shift_context(n + 1, ctx.key, 1, 0, 0) == 0
    finally:
        ctx.release()


@pytest.mark.parametrize("offset,layers", [((1, 0, 0), 2), ((0, -1, 1), 0)])
def test_halo_exchange_equals_oracle(oracle_mod, offset, layers):
    lib = A.load()
    n = 16
    dims = (n, n, n)
    cfg, pk1, _, t1 = channel_case(dims, seed=51, thermal=True)
    _, pk2, _, t2 = channel_case(dims, seed=52, thermal=True)
    setup(lib, dims, cfg, A.KERNEL_EXACT, A.SOLVER_CLASSIC_D3Q27)
    a, b = A.Context(*dims, key=87), A.Context(*dims, key=88)
    oa, ob = oracle_mod.Oracle(*dims, cfg, precision=32), oracle_mod.Oracle(*dims, cfg, precision=32)
    try:
        for ctx, o, pk, t in ((a, oa, pk1, t1), (b, ob, pk2, t2)):
            ctx.set_temperature(t)
            o.set_temperature(t)
            for i in range(3):
                ctx.step(pk if i == 0 else None)
                o.step(pk if i == 0 else None)
        log(f"halo exchange offset {offset} layers {layers or 'grid/4'}")
        if layers:
            A.check(lib.aero_lbm_exchange_halo_layers_rect(*dims, layers, a.key, b.key, *offset), "halo")
        else:
            A.check(lib.aero_lbm_exchange_halo(n, a.key, b.key, *offset), "halo")
        oa.exchange_halo(ob, offset, layers or max(1, n // 4))
        for _ in range(3):
            for ctx, o in ((a, oa), (b, ob)):
                ctx.step(None)
                o.step(None)
        for ctx, o in ((a, oa), (b, ob)):
            assert np.array_equal(ctx.flow(), o.flow())
            assert np.array_equal(ctx.temperature(), o.temperature())
        for bad in ((2, 0, 0), (0, 0, 0)):
            assert lib.aero_lbm_exchange_halo_layers_rect(*dims, 2, a.key, b.key, *bad) == 0
        assert lib.aero_lbm_exchange_halo_layers_rect(*dims, 8, a.key, b.key, 1, 0, 0) == 0
        assert lib.aero_lbm_exchange_halo_layers_rect(*dims, 2, a.key, 4242, 1, 0, 0) == 0
        assert b"[missing]" in lib.aero_lbm_last_error()
    finally:
        a.release()
        b.release()


def test_halo_exchange_leaves_in_place_contexts():
    lib = A.load()
    n = 16
    dims = (n, n, n)
    cfg, pk1, _, _ = channel_case(dims, seed=61)
    _, pk2, _, _ = channel_case(dims, seed=62)
    setup(lib, dims, cfg, A.KERNEL_FAST, A.SOLVER_D3Q27_FP16_INPLACE)
    a, b = A.Context(*dims, key=89), A.Context(*dims, key=90)
    try:
        a.step(pk1)
        b.step(pk2)
        fa, fb = a.flow(), b.flow()
        A.check(lib.aero_lbm_exchange_halo_layers_rect(*dims, 2, a.key, b.key, 1, 0, 0), "halo")
        assert np.array_equal(a.flow(), fa) and np.array_equal(b.flow(), fb)
        assert lib.aero_lbm_shift_context(n, a.key, 1, 0, 0) == 0
        assert b"exact" in lib.aero_lbm_last_error()
    finally:
        a.release()
        b.release()
        lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_CLASSIC_D3Q27)
