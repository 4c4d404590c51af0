"""GPU checks of the fast paths (FP32 A-B, FP32/FP16/BF16/S16 A-A) through the C ABI.

* FP32 A-A and FP32 A-B run the same collision code; the A-A indexing, parity
  flip, bounce-back slots and face store must reproduce the A-B result bit for
  bit after every step count.
* FP32 fast / 16-bit storage against the FP64 oracle: velocity error relative to
  the inflow speed. Bars: 16-bit shifted storage <= 1e-3 of the inflow speed
  (channel 16^3 x 200, Boussinesq 32^3 x 300, the 128^3 workload at 100-300
  steps); FP32 <= 1e-4. IEEE half misses the 16-bit bar by 5-15% and its cases are
  expected failures; the fixed-point storage (S16) meets it.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    return A.load()


def make_ctx(lib, dims, cfg, ov, mode, key):
    A.check(lib.aero_lbm_init_rect(*dims, 11, 4), "init_rect")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "set_config")
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(ov)), "set_overrides")
    lib.aero_lbm_set_realtime_solver_mode(mode)
    return A.Context(*dims, key=key)


def run(lib, dims, cfg, ov, mode, packet, steps, temperature=None, key=31):
    ctx = make_ctx(lib, dims, cfg, ov, mode, key)
    if temperature is not None:
        ctx.set_temperature(temperature)
    ctx.step(packet)
    for _ in range(steps - 1):
        ctx.step(None)
    flow, temp = ctx.flow(), ctx.temperature()
    ctx.release()
    return flow, temp


def variant_cfg(variant):
    cfg = W.channel_config(0.04, 40.0, sgs=True, buoyancy=True)
    if variant == "symmetry":
        for f in (cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
            f.hydrodynamic_kind = A.HYDRO_SYMMETRY
            f.thermal_kind = A.THERMAL_ADIABATIC
    elif variant == "walls":
        cfg.y_min.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
        cfg.y_max.hydrodynamic_kind = A.HYDRO_MOVING_WALL
        cfg.y_max.velocity[0] = 0.03
        cfg.z_min.hydrodynamic_kind = A.HYDRO_PRESSURE
        cfg.z_max.hydrodynamic_kind = A.HYDRO_INHERIT
    elif variant == "sponge":
        cfg.flags &= ~A.FLAG_DISABLE_SPONGE
    return cfg


def variant_inputs(dims):
    solid = W.solid_boxes(*dims, 0.12, seed=21, edge=(1, 3), keep_clear_x=1)
    t = W.hot_blob(*dims, 0.05, 3.0, seed=4, axis=1)
    pk = W.make_packet(solid, u0=(0.04, 0.0, 0.0), perturb=2e-3, seed_u=5, temperature=t,
                       source=0.001 * np.ones(dims, np.float32))
    return pk, t


@pytest.mark.parametrize("variant", ["channel", "symmetry", "walls", "sponge"])
def test_aa_fp32_equals_ab_fp32_bitwise(lib, variant):
    dims = (10, 9, 8)
    cfg = variant_cfg(variant)
    ov = W.reference_overrides(A.KERNEL_FAST)
    ov.boussinesq_beta = 0.05
    pk, t = variant_inputs(dims)
    for steps in (1, 2, 3, 6, 9):
        log(f"aa-vs-ab {variant} steps={steps}")
        fa, ta = run(lib, dims, cfg, ov, A.SOLVER_D3Q27_FP32_INPLACE, pk, steps, t)
        fb, tb = run(lib, dims, cfg, ov, A.SOLVER_CLASSIC_D3Q27, pk, steps, t)
        assert np.array_equal(fa, fb), (steps, np.max(np.abs(fa - fb)))
        assert np.array_equal(ta, tb), steps


@pytest.mark.parametrize("dims", [(6, 5, 70), (5, 4, 130)])
@pytest.mark.parametrize("variant", ["channel", "walls"])
def test_ragged_z_rows(oracle_mod, lib, dims, variant):
    """z rows longer than one wavefront and not a multiple of 64 (partial last wave):
    A-A in place equals A-B bit for bit, and both stay within 1e-4 of the inflow
    speed of the FP32 oracle."""
    cfg = variant_cfg(variant)
    ov = W.reference_overrides(A.KERNEL_FAST)
    ov.boussinesq_beta = 0.05
    pk, t = variant_inputs(dims)
    for steps in (1, 2, 5):
        log(f"ragged {dims} {variant} steps={steps}")
        fa, ta = run(lib, dims, cfg, ov, A.SOLVER_D3Q27_FP32_INPLACE, pk, steps, t)
        fb, tb = run(lib, dims, cfg, ov, A.SOLVER_CLASSIC_D3Q27, pk, steps, t)
        assert np.array_equal(fa, fb), (steps, np.max(np.abs(fa - fb)))
        assert np.array_equal(ta, tb), steps
    o = oracle_mod.Oracle(*dims, cfg, overrides=ov, precision=32)
    o.set_temperature(t)
    o.run(pk, 5)
    ref = o.flow().reshape(-1, 4)
    got = fb.reshape(-1, 4)
    assert np.max(np.abs(got[:, :3] - ref[:, :3])) <= 1e-4 * 0.04
    assert np.max(np.abs(tb.reshape(-1) - o.temperature())) <= 1e-5


def test_fast_population_array_size_limit(lib):
    """The fast kernels address a population array through one buffer resource with
    32-bit byte offsets: FP32 in place takes 27 planes x 4 B per cell, so a brick of
    20 x 1024 x 1001 cells (2.2 GB) is refused with a message and 19 x 1024 x 1000
    (2.10 GB) is accepted and steps: uniform flow through the periodic channel is a
    fixed point, so after two even and two odd in-place steps every cell, the last
    planes' (q = 26, byte offsets near 2^31) included, still reads the inflow."""
    u0 = 0.04
    cfg = W.channel_config(u0, 40.0, sgs=False, buoyancy=False)
    ov = W.reference_overrides(A.KERNEL_FAST)
    for dims, ok in (((20, 1024, 1001), False), ((19, 1024, 1000), True)):
        log(f"size limit {dims} accepted={ok}")
        ctx = make_ctx(lib, dims, cfg, ov, A.SOLVER_D3Q27_FP32_INPLACE, key=41)
        t = np.zeros(ctx.cells, np.float32)
        if ok:
            ctx.set_temperature(t)
            assert np.array_equal(ctx.temperature()[:1000], t[:1000])
            packet = W.make_packet(np.zeros(dims, np.uint8), u0=(u0, 0.0, 0.0), perturb=0.0)
            ctx.step(packet)
            del packet
            for _ in range(3):
                ctx.step(None)
            flow = ctx.flow().reshape(-1, 4)
            assert np.all(np.isfinite(flow))
            err = float(np.max(np.abs(flow[:, :3] - np.array([u0, 0.0, 0.0], np.float32))))
            log(f"  uniform flow after 4 steps: max |u - u0| = {err:.2e}")
            assert err < 1e-5
        else:
            with pytest.raises(A.AeroLbmError, match="too large for the fast kernels"):
                ctx.set_temperature(t)
        ctx.release()


def velocity_error(flow, ref_u, solid, u_in):
    f = flow.reshape(-1, 4)[:, :3].astype(np.float64)
    fluid = solid.reshape(-1) == 0
    return float(np.max(np.abs(f[fluid] - ref_u[fluid]))) / u_in


# Bar: 16-bit shifted storage within 1e-3 of the inflow speed of the FP64 oracle,
# FP32 within 1e-4. Storage rounding dominates the 16-bit figures, which land at
# 8.9e-4 (IEEE half, 11 significant bits of f - w) and 9.6e-4 (16-bit fixed point)
# of the inflow speed on this channel; FP32 ends at 8e-6.
@pytest.mark.parametrize("mode,bar", [(A.SOLVER_CLASSIC_D3Q27, 1e-4),
                                      (A.SOLVER_D3Q27_FP16_INPLACE, 1e-3),
                                      (A.SOLVER_D3Q27_S16_INPLACE, 1e-3),
                                      (A.SOLVER_D3Q27_FP32_INPLACE, 1e-4)])
def test_fast_paths_vs_fp64_oracle_channel(oracle_mod, lib, mode, bar):
    n, steps = 16, 200
    log(f"fast vs fp64 channel mode={mode}")
    cfg, pk, solid = W.config1(n)
    o = oracle_mod.Oracle(n, n, n, cfg, precision=64)
    o.run(pk, steps)
    ref = o.flow().reshape(-1, 4)[:, :3].astype(np.float64)
    flow, _ = run(lib, (n, n, n), cfg, W.reference_overrides(A.KERNEL_FAST), mode, pk, steps)
    err = velocity_error(flow, ref, solid, 0.05)
    log(f"  velocity error / U_in = {err:.3e}")
    assert err <= bar


SMALL_MODES = {"fp32": A.SOLVER_CLASSIC_D3Q27, "fp16": A.SOLVER_D3Q27_FP16_INPLACE,
               "s16": A.SOLVER_D3Q27_S16_INPLACE, "bf16": A.SOLVER_D3Q27_BF16_INPLACE}


@pytest.mark.parametrize("name,bar,tbar", [("fp32", 1e-4, 1e-5),
                                           ("fp16", 1e-3, 1e-4),
                                           ("s16", 1e-3, 1e-4),
                                           # BF16's 8-bit significand is reported, not held to the 16-bit bar
                                           ("bf16", 5e-2, 1e-3)])
def test_config3_small_vs_fp64_fixture(lib, name, bar, tbar):
    """32^3 Boussinesq workload (C=0.1, beta 0.1, gravity -z), 300 steps, against the
    FP64 oracle fixture tests/golden/config3_small.json."""
    gold = json.load(open(os.path.join(HERE, "golden", "config3_small.json")))
    n, steps = gold["n"], gold["steps"]
    cfg, ov, pk, t, solid = W.config3(n, beta=0.1, cs=0.1)
    idx = gold["fp64"]["probe_cells"]
    ref = np.array(gold["fp64"]["probe_flow"])[:, :3]
    fluid = solid.reshape(-1)[idx] == 0
    log(f"config3 small {name}")
    flow, temp = run(lib, (n, n, n), cfg, ov, SMALL_MODES[name], pk, steps, t)
    got = flow.reshape(-1, 4)[idx, :3].astype(np.float64)
    err = float(np.max(np.abs(got[fluid] - ref[fluid]))) / 0.05
    terr = float(np.max(np.abs(temp[idx] - np.array(gold["fp64"]["probe_temperature"]))))
    log(f"  {name}: velocity error / U_in = {err:.3e}, temperature error = {terr:.3e}")
    assert err <= bar
    assert terr <= tbar


def test_fp16_mass_drift_closed_box(lib):
    """Closed box, bounce-back on all faces: mass drift stays at 16-bit rounding level
    (FP16 and fixed-point storage)."""
    n = 16
    cfg = W.channel_config(0.03, 30.0, sgs=True, buoyancy=False)
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
    solid = np.zeros((n, n, n), np.uint8)
    pk = W.make_packet(solid, u0=(0.03, 0.0, 0.0), perturb=5e-3, seed_u=2)
    ov = W.reference_overrides(A.KERNEL_FAST)
    for mode in (A.SOLVER_D3Q27_FP16_INPLACE, A.SOLVER_D3Q27_S16_INPLACE):
        f1, _ = run(lib, (n, n, n), cfg, ov, mode, pk, 1)
        f2, _ = run(lib, (n, n, n), cfg, ov, mode, pk, 400)
        m1 = np.sum(f1.reshape(-1, 4)[:, 3].astype(np.float64))
        m2 = np.sum(f2.reshape(-1, 4)[:, 3].astype(np.float64))
        log(f"  mode {mode}: mass drift per cell {abs(m2 - m1) / n ** 3:.3e}")
        assert abs(m2 - m1) / n ** 3 < 2e-5


def run_config3_128(lib, mode, steps, key=41):
    n = 128
    cfg, ov, pk, t, solid = W.config3(n, beta=0.1, cs=0.1)
    ctx = make_ctx(lib, (n, n, n), cfg, ov, mode, key)
    ctx.set_temperature(t)
    ctx.step(pk)
    ms = ctx.run_timed(steps - 1)
    flow = ctx.flow()
    ctx.release()
    return flow, solid, ms


H128 = {"fp32": A.SOLVER_CLASSIC_D3Q27, "fp16": A.SOLVER_D3Q27_FP16_INPLACE, "s16": A.SOLVER_D3Q27_S16_INPLACE}


@pytest.mark.parametrize("name,bar", [("fp32", 1e-4), ("fp16", 1e-3), ("s16", 1e-3)])
def test_config3_128_horizon_vs_fp64_fixture(lib, name, bar):
    """The 128^3 workload (C=0.1, beta 0.1, gravity -z, 10% solid boxes, hot blob on
    the floor, inflow 0.05) against the FP64 oracle at 4087 seeded probes after 100,
    200 and 300 steps (tests/golden/config3_horizons.json). Up to step 300 the FP32
    oracle stays within 1e-4 of the inflow speed of the FP64 oracle (fixture fields
    fp32_oracle_err_*), so these horizons measure each storage format's own error;
    beyond them the plume's instability amplifies every rounding difference (see
    test_config3_1000_steps_report)."""
    gold = json.load(open(os.path.join(HERE, "golden", "config3_horizons.json")))
    idx = gold["probe_cells"]
    n = gold["n"]
    _, _, _, _, solid = W.config3(n, beta=0.1, cs=0.1)
    fluid = solid.reshape(-1)[idx] == 0
    for h in gold["horizons"]:
        assert gold[f"fp32_oracle_err_{h}"] <= 1e-4, (h, gold[f"fp32_oracle_err_{h}"])
        log(f"config3 128^3 {name} horizon {h}")
        flow, _, ms = run_config3_128(lib, H128[name], h)
        got = flow.reshape(-1, 4)[idx, :3].astype(np.float64)
        ref = np.array(gold[f"fp64_{h}"]["probe_flow"])[:, :3]
        err = float(np.max(np.abs(got[fluid] - ref[fluid]))) / 0.05
        log(f"  {name} step {h}: max velocity error / U_in = {err:.3e} "
            f"(FP32 oracle {gold[f'fp32_oracle_err_{h}']:.2e})")
        assert err <= bar


def test_config3_1000_steps_report(lib):
    """The same workload after 1000 steps against the FP64 oracle fixture
    (tests/golden/config3_full.json). The plume turns unstable after about 300
    steps: two FP32 evaluations differing only in fused multiply-add rounding drift
    apart to 5e-2 of the inflow speed by step 1000, and the FP32 oracle itself ends
    1e-2 away from the FP64 oracle (fixture field fp32_oracle_err). The figures are
    reported; the assertions are what 1000 steps of any correct update satisfy:
    finite velocities, and FP32 within twice the FP32 oracle's own distance."""
    gold = json.load(open(os.path.join(HERE, "golden", "config3_full.json")))
    n, steps = gold["n"], gold["steps"]
    _, _, _, _, solid = W.config3(n, beta=0.1, cs=0.1)
    idx = gold["fp64"]["probe_cells"]
    ref = np.array(gold["fp64"]["probe_flow"])[:, :3]
    fluid = solid.reshape(-1)[idx] == 0
    oracle32 = gold["fp32_oracle_err"]
    errs = {}
    for name, mode in (("fp32", A.SOLVER_CLASSIC_D3Q27), ("fp16", A.SOLVER_D3Q27_FP16_INPLACE),
                       ("s16", A.SOLVER_D3Q27_S16_INPLACE), ("bf16", A.SOLVER_D3Q27_BF16_INPLACE)):
        log(f"config3 1000 steps {name}")
        flow, _, ms = run_config3_128(lib, mode, steps)
        u = flow.reshape(-1, 4)[:, :3]
        assert np.all(np.isfinite(u))
        got = u[idx].astype(np.float64)
        errs[name] = float(np.max(np.abs(got[fluid] - ref[fluid]))) / 0.05
        log(f"  {name}: max velocity error / U_in = {errs[name]:.3e} ({ms:.1f} ms for {steps - 1} steps)")
    log(f"  FP32 oracle vs FP64 oracle: {oracle32:.3e}")
    assert errs["fp32"] <= 2.0 * oracle32


def flow_errors(got, ref, u_in):
    """velocity difference over the inflow speed, density difference over the density"""
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    return (float(np.max(np.abs(got[:, :3] - ref[:, :3]))) / u_in,
            float(np.max(np.abs(got[:, 3] - ref[:, 3]) / (1.0 + ref[:, 3]))))


@pytest.mark.parametrize("mode", [A.SOLVER_CLASSIC_D3Q27, A.SOLVER_D3Q27_FP32_INPLACE])
def test_fast_fp32_config1_vs_fp32_oracle(lib, mode):
    """The fast FP32 kernels (A-B: the bench's FP32 line; A-A FP32) on config 1 (32^3,
    15% solid boxes, SGS at the reference constant) against the FP32 oracle: density
    and velocity within 1e-5 relative (velocity relative to the inflow speed) at
    50, 100 and 200 steps (tests/golden/config1_horizons.json), where the fast kernels
    differ from the FP32 oracle by as much as the FP32 oracle differs from the FP64
    one (4-8e-6). This low-viscosity flow (tau ~0.503) amplifies rounding
    differences: at 400 steps the fast A-B kernel is 3.5e-5 of the inflow speed from
    the FP32 oracle, and by 1000 steps the FP32 oracle itself is 5.8e-3 from the FP64
    one, so only the bit-exact kernel (test_exact_config1_full_1000_steps) meets 1e-5
    there; the 400- and 1000-step distances are reported beside the oracle's."""
    gold = json.load(open(os.path.join(HERE, "golden", "config1_horizons.json")))
    n = gold["n"]
    cfg, pk, solid = W.config1(n)
    idx = gold["probe_cells"]
    ctx = make_ctx(lib, (n, n, n), cfg, W.reference_overrides(A.KERNEL_FAST), mode, key=33)
    ctx.step(pk)
    done = 1
    for h in gold["horizons"]:
        for _ in range(h - done):
            ctx.step(None)
        done = h
        got = ctx.flow().reshape(-1, 4)[idx]
        eu, er = flow_errors(got, gold[f"fp32_{h}"], 0.05)
        ou, orr = flow_errors(gold[f"fp32_{h}"], gold[f"fp64_{h}"], 0.05)
        log(f"  mode {mode} step {h}: velocity {eu:.2e}, density {er:.2e} of the FP32 oracle "
            f"(FP32 vs FP64 oracle: {ou:.2e}, {orr:.2e})")
        if h <= 200:
            assert ou <= 1e-5 and orr <= 1e-5
            assert eu <= 1e-5 and er <= 1e-5, h
        else:
            assert eu <= 1e-3 and er <= 1e-5, h
    full = json.load(open(os.path.join(HERE, "golden", "config1.json")))
    for _ in range(full["steps"] - done):
        ctx.step(None)
    got = ctx.flow().reshape(-1, 4)[full["fp32"]["probe_cells"]]
    eu, er = flow_errors(got, full["fp64"]["probe_flow"], 0.05)
    ou, orr = flow_errors(full["fp32"]["probe_flow"], full["fp64"]["probe_flow"], 0.05)
    log(f"  mode {mode} step 1000 vs FP64 oracle: velocity {eu:.2e}, density {er:.2e} "
        f"(FP32 oracle: {ou:.2e}, {orr:.2e})")
    assert np.all(np.isfinite(got))
    ctx.release()
