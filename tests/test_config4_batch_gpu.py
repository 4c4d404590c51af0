"""The batched workload at full size on the GPU: 40 x 32^3 + 12 x 64^3 + 4 x 128^3 FP16
bricks (workloads.batch_brick_specs), each with its own solids, inflow and hot blob,
500 steps, the last 496 in graph-replayed cycles of 4 steps plus the publish of every
brick's int16 atlas at stride 4.

Every brick also has a twin context stepped 500 times through the plain per-context
API with its own configuration set before each step; the batched launches run the
same kernels on the same parameters, so each brick's flow must equal its twin's bit
for bit. Against the FP64 oracle (probe fixture, tests/golden/make_golden.py
config4_batch) the FP16 error is reported per brick relative to its inflow speed."""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config4_batch.json")
STEPS, K = 500, 4


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def flow(lib, key, n):
    f = np.zeros(n ** 3 * 4, np.float32)
    A.check(lib.aero_lbm_get_flow_state_rect(n, n, n, key, A.fptr(f)), "flow")
    return f


def test_config4_batched_bricks_500_steps():
    lib = A.load()
    A.check(lib.aero_lbm_init_rect(32, 32, 32, 11, 4), "init_rect")
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(W.reference_overrides(A.KERNEL_FAST))), "overrides")
    specs = W.batch_brick_specs()
    cfgs, batch_keys, twin_keys = [], [], []
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_D3Q27_FP16_INPLACE)
    for i, spec in enumerate(specs):
        n = spec["edge"]
        cfg, pk, t, _ = W.batch_brick_case(spec)
        A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
        for keys, base in ((batch_keys, 40000), (twin_keys, 41000)):
            ctx = A.Context(n, n, n, key=base + i)
            ctx.set_temperature(t)
            ctx.step(pk)  # upload, initialise, step 1
            keys.append(ctx.key)
        cfgs.append(cfg)
    log(f"config4: {len(specs)} bricks, {sum(s['edge'] ** 3 for s in specs)} cells")

    def plain_steps(keys, count):
        for _ in range(count):
            for k, cfg, spec in zip(keys, cfgs, specs):
                n = spec["edge"]
                A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
                A.check(lib.aero_lbm_step_rect_cached(n, n, n, k, None), "cached step")

    batch = None
    try:
        # steps 2-4 through the plain API, then 124 batched cycles of 4
        plain_steps(batch_keys, 3)
        batch = A.Batch(batch_keys, steps_per_publish=K, stride=4, scales=[20.0] * len(specs), autotune=True)
        log(f"config4 batch: {batch.info()}")
        cycles = (STEPS - 4) // K
        batch.run(cycles, graph=True)
        log("config4: twins through the plain API")
        plain_steps(twin_keys, STEPS - 1)
        fixture = None
        if os.path.exists(FIXTURE):
            with open(FIXTURE) as fh:
                fixture = {b["index"]: b for b in json.load(fh)["bricks"]}
        worst = 0.0
        for i, spec in enumerate(specs):
            n = spec["edge"]
            fa, fb = flow(lib, batch_keys[i], n), flow(lib, twin_keys[i], n)
            assert np.all(np.isfinite(fa)), i
            assert np.array_equal(fa, fb), (i, n, float(np.max(np.abs(fa - fb))))
            if fixture is not None:
                b = fixture[i]
                idx = np.array(b["probe_cells"])
                fluid = np.array(b["probe_solid"]) == 0
                ref = np.array(b["probe_flow"])[:, :3]
                got = fa.reshape(-1, 4)[idx, :3].astype(np.float64)
                err = float(np.max(np.abs(got[fluid] - ref[fluid]))) / spec["speed"] if fluid.any() else 0.0
                worst = max(worst, err)
                log(f"brick {i:2d} {n:3d}^3 inflow {spec['speed']:.3f}: FP16 vs FP64 max |du| / U_in = {err:.2e}")
        if fixture is not None:
            log(f"config4: worst FP16 error after {STEPS} steps over all bricks: {worst:.2e} of the inflow speed")
    finally:
        if batch is not None:
            batch.release()
        lib.aero_lbm_set_realtime_solver_mode(0)
        for k in batch_keys + twin_keys:
            lib.aero_lbm_release_context(k)
