"""Batched stepping (aero_lbm_batch_*) against the per-context path.

Bricks of mixed sizes, each with its own inflow and solids, advance together in
the batched launches (graph replay and direct enqueue) and, as twins, one context
at a time through aero_lbm_step_rect_cached with their own configuration set
before every step. The same kernels run on the same parameters, so the flow fields
must agree bit for bit, and the device-quantised atlas must equal the host
quantisation of the per-brick readout.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def brick_setup(n, seed, speed, angle_deg, thermal):
    cfg = W.channel_config(speed, 40.0, sgs=True, buoyancy=thermal)
    a = math.radians(angle_deg)
    v = (speed * math.cos(a), speed * math.sin(a), 0.0)
    cfg.initial_velocity[0], cfg.initial_velocity[1] = v[0], v[1]
    cfg.x_min.velocity[0], cfg.x_min.velocity[1] = v[0], v[1]
    solid = W.solid_boxes(n, n, n, 0.06, seed=seed, edge=(2, 5), keep_clear_x=2)
    t = W.hot_blob(n, n, n, 0.02, max(2.0, n / 8), seed=seed + 1, axis=2) if thermal else None
    if t is not None:
        t[solid.astype(bool)] = 0.0
    pk = W.make_packet(solid, u0=v, perturb=1e-3, seed_u=seed + 2, temperature=t)
    return cfg, pk, t


SPECS = [(16, 1, 0.03, 0.0, False), (32, 2, 0.05, 30.0, True), (16, 3, 0.08, -60.0, True),
         (24, 4, 0.02, 90.0, False), (32, 5, 0.04, 170.0, True)]


def make_contexts(lib, mode, base_key, shell_index=None):
    keys, cfgs = [], []
    lib.aero_lbm_set_realtime_solver_mode(mode)
    for i, (n, seed, speed, ang, thermal) in enumerate(SPECS):
        cfg, pk, t = brick_setup(n, seed, speed, ang, thermal)
        A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
        ctx = A.Context(n, n, n, key=base_key + i)
        if t is not None:
            ctx.set_temperature(t)
        ctx.step(pk)
        if shell_index == i:
            ref = np.tile(np.array([0.05, 0.01, 0.0, 0.001], np.float32), n ** 3)
            A.check(lib.aero_lbm_set_d3q27_f16_boundary_shell_rect(n, n, n, ctx.key, A.fptr(ref), 4, 0b001101, 4,
                                                                   0.2), "shell")
        keys.append(ctx.key)
        cfgs.append(cfg)
    lib.aero_lbm_set_realtime_solver_mode(0)
    return keys, cfgs


def flows(lib, keys):
    out = []
    for k, spec in zip(keys, SPECS):
        n = spec[0]
        f = np.zeros(n ** 3 * 4, np.float32)
        A.check(lib.aero_lbm_get_flow_state_rect(n, n, n, k, A.fptr(f)), "flow")
        out.append(f)
    return out


@pytest.fixture(scope="module")
def lib():
    lib = A.load()
    A.check(lib.aero_lbm_init_rect(16, 16, 16, 11, 4), "init")
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(W.reference_overrides(A.KERNEL_FAST))), "overrides")
    return lib


@pytest.mark.parametrize("mode", [A.SOLVER_D3Q27_FP16_INPLACE, A.SOLVER_D3Q27_FP32_INPLACE])
def test_batch_equals_per_context_steps(lib, mode):
    K, cycles = 4, 3
    ka, cfgs = make_contexts(lib, mode, 7000, shell_index=1)
    kb, _ = make_contexts(lib, mode, 7100, shell_index=1)
    kc, _ = make_contexts(lib, mode, 7200, shell_index=1)
    scales = [20.0] * len(SPECS)
    ba = A.Batch(ka, steps_per_publish=K, stride=4, scales=scales, autotune=False)
    bb = A.Batch(kb, steps_per_publish=K, stride=4, scales=scales, autotune=True)
    log(ba.info())
    log(bb.info())
    ba.run(cycles, graph=True)
    bb.run(cycles - 1, graph=False)
    bb.run_per_brick(1)
    # twins of the plain API: each context with its own configuration before each step
    for _ in range(cycles * K):
        for k, cfg, spec in zip(kc, cfgs, SPECS):
            n = spec[0]
            A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
            A.check(lib.aero_lbm_step_rect_cached(n, n, n, k, None), "cached step")
    fa, fb, fc = flows(lib, ka), flows(lib, kb), flows(lib, kc)
    for i in range(len(SPECS)):
        assert np.all(np.isfinite(fa[i]))
        assert np.array_equal(fa[i], fc[i]), i
        assert np.array_equal(fb[i], fc[i]), i
        n = SPECS[i][0]
        # batch a published on the device, batch b's last cycle quantised on the host
        assert np.array_equal(ba.atlas(i, n), bb.atlas(i, n)), i
    ba.release()
    bb.release()
    for k in ka + kb + kc:
        lib.aero_lbm_release_context(k)


def test_batch_rejects_mixed_parity_and_outside_steps(lib):
    ka, cfgs = make_contexts(lib, A.SOLVER_D3Q27_FP16_INPLACE, 7300)
    n = SPECS[0][0]
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfgs[0])), "config")
    A.check(lib.aero_lbm_step_rect_cached(n, n, n, ka[0], None), "cached step")
    with pytest.raises(A.AeroLbmError, match="parity"):
        A.Batch(ka, steps_per_publish=4)
    b = A.Batch(ka[1:], steps_per_publish=2)
    b.run(1)
    n1 = SPECS[1][0]
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfgs[1])), "config")
    A.check(lib.aero_lbm_step_rect_cached(n1, n1, n1, ka[1], None), "cached step")
    with pytest.raises(A.AeroLbmError, match="outside the batch"):
        b.run(1)
    with pytest.raises(A.AeroLbmError, match="even"):
        A.Batch(ka[2:], steps_per_publish=3)
    b.release()
    for k in ka:
        lib.aero_lbm_release_context(k)


def test_batch_rejects_member_stepped_twice_with_new_obstacles(lib):
    """Two outside steps of a member restore its parity, and a new packet rebuilds its
    link mask and rare-cell list; the batch must refuse to launch on its stale tables
    (the context's generation moved) before touching any device memory."""
    ka, cfgs = make_contexts(lib, A.SOLVER_D3Q27_FP16_INPLACE, 7400)
    b = A.Batch(ka, steps_per_publish=2)
    b.run(1)
    n, seed, speed, ang, thermal = SPECS[2]
    solid = W.solid_boxes(n, n, n, 0.2, seed=99, edge=(2, 5), keep_clear_x=2)
    pk = W.make_packet(solid, u0=(speed, 0.0, 0.0), perturb=1e-3, seed_u=5)
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfgs[2])), "config")
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_D3Q27_FP16_INPLACE)
    A.check(lib.aero_lbm_step_rect(A.fptr(pk), n, n, n, ka[2], None), "packet step")
    A.check(lib.aero_lbm_step_rect_cached(n, n, n, ka[2], None), "cached step")
    lib.aero_lbm_set_realtime_solver_mode(0)
    before = flows(lib, ka)
    with pytest.raises(A.AeroLbmError, match="outside the batch"):
        b.run(1)
    after = flows(lib, ka)
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    b.release()
    for k in ka:
        lib.aero_lbm_release_context(k)
