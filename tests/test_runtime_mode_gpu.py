"""GPU parity of the reference's runtime (non-benchmark) mode: the mod's default
classic path (ClientL2Solver V0_1 steps its bricks with the benchmark configuration
reset, src/main/java/com/aerodynamics4mc/client/ClientL2Solver.java:84-87,
native/src/aero_lbm_simulation_bridge.cpp:1790-1794).

Runtime mode runs on the exact FP32 kernels and must equal the FP32 oracle bit for
bit: state nudge and temperature reference at the faces, sponge, SGS at the air
viscosity, fan forcing with its noise, turbulent thermal diffusivity, BFECC every
second step, inherit faces as convective outflow (native/src/aero_lbm_jni.cpp:675-993,
1971-2143, 2209-2413, 8037-8064, 8480-8507).
"""
import ctypes
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    return A.load()


def runtime_packet(dims, seed, fan_box=True, blob=0.3, src=True):
    nx, ny, nz = dims
    solid = W.solid_boxes(nx, ny, nz, 0.08, seed=seed, edge=(2, 4), keep_clear_x=2)
    t = W.hot_blob(nx, ny, nz, blob, 3.0, seed=seed + 1, axis=1)
    t[solid.astype(bool)] = 0.0
    source = np.zeros(dims, np.float32)
    if src:
        source[nx // 2 - 1:nx // 2 + 1, 1:3, nz // 2 - 1:nz // 2 + 1] = 0.004
    pk = W.make_packet(solid, u0=(0.06, 0.01, -0.02), perturb=0.02, seed_u=seed + 2, temperature=t,
                       source=source).reshape(nx, ny, nz, 11)
    if fan_box:
        fan = np.zeros(dims, bool)
        fan[3:7, 4:8, 3:7] = True
        fan &= solid == 0
        pk[fan, 1] = 1.0
        pk[fan, 2:5] = (3.0, 1.0, 0.5)  # m/s at 20 m/s per lattice unit: target ~0.16
        half = np.zeros(dims, bool)
        half[10:12, 2:6, 8:12] = True
        half &= solid == 0
        pk[half, 1] = 0.5
        pk[half, 2:5] = (-1.0, 2.0, 0.0)
    return pk.reshape(-1), solid


def runtime_cfg(lib, periodic=False):
    cfg = A.default_config()  # enabled = 0: runtime mode, inherit faces
    if periodic:
        for f in (cfg.y_min, cfg.y_max):
            f.hydrodynamic_kind = A.HYDRO_PERIODIC
            f.thermal_kind = A.THERMAL_PERIODIC
    return cfg


def make(lib, dims, cfg, key=61):
    A.check(lib.aero_lbm_init_rect(*dims, 11, 4), "init_rect")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "set_config")
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(W.reference_overrides(A.KERNEL_FAST))), "overrides")
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_CLASSIC_D3Q27)
    return A.Context(*dims, key=key)


@pytest.mark.parametrize("periodic", [False, True])
def test_runtime_mode_equals_fp32_oracle(oracle_mod, lib, periodic):
    dims = (20, 18, 16)
    cfg = runtime_cfg(lib, periodic)
    pk, _ = runtime_packet(dims, seed=31)
    ctx = make(lib, dims, cfg)
    o = oracle_mod.Oracle(*dims, cfg, precision=32)
    done = 0
    for steps in (1, 2, 3, 6, 11):
        log(f"runtime mode periodic={periodic} steps={steps}")
        while done < steps:
            # the classic brick path hands the packet over every step
            ctx.step(pk)
            o.step(pk)
            done += 1
        g, r = ctx.flow(), o.flow()
        assert np.array_equal(g, r), (steps, float(np.max(np.abs(g - r))))
        tg, tr = ctx.temperature(), o.temperature()
        assert np.array_equal(tg, tr), (steps, float(np.max(np.abs(tg - tr))))
    assert np.abs(ctx.temperature()).max() > 1e-2
    u = ctx.flow().reshape(-1, 4)[:, :3]
    assert np.abs(u).max() > 0.05
    ctx.release()


def test_runtime_mode_payload_change_and_cached_steps(oracle_mod, lib):
    dims = (16, 16, 16)
    cfg = runtime_cfg(lib)
    pk1, _ = runtime_packet(dims, seed=41)
    pk2, _ = runtime_packet(dims, seed=53, blob=-0.2)
    ctx = make(lib, dims, cfg, key=62)
    o = oracle_mod.Oracle(*dims, cfg, precision=32)
    seq = [pk1, None, None, pk2, None, pk2, None, pk1, None]
    for i, pk in enumerate(seq):
        log(f"runtime payload sequence step {i + 1}")
        ctx.step(pk)
        o.step(pk)
    assert np.array_equal(ctx.flow(), o.flow())
    assert np.array_equal(ctx.temperature(), o.temperature())
    ctx.release()


def test_runtime_mode_rejects_fast_storage(lib):
    dims = (8, 8, 8)
    cfg = runtime_cfg(lib)
    A.check(lib.aero_lbm_init_rect(*dims, 11, 4), "init_rect")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "set_config")
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_D3Q27_FP16_INPLACE)
    ctx = A.Context(*dims, key=63)
    pk, _ = runtime_packet(dims, seed=5, fan_box=False)
    with pytest.raises(A.AeroLbmError, match="runtime"):
        ctx.step(pk)
    ctx.release()
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_CLASSIC_D3Q27)
