"""Wind-tunnel solver C ABI (aero_solver_*) on the GPU.

The layer configures the benchmark step for each wind-tunnel mode
(native/src/aero_solver_capi.cpp:209-258) and drives one lattice context with an
11-channel packet built from the solid mask and the inlet or caller flow state
(aero_solver_capi.cpp:130-160, 442-470). With the exact kernel the output must
equal the FP32 oracle driven with that configuration and packet, bit for bit.
"""
import ctypes
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

from wind_tunnel_cases import DIMS, DX, DT, boundary, tunnel_config, tunnel_packet  # noqa: E402


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_flow(oracle_mod, cfg, packet, steps):
    o = oracle_mod.Oracle(*DIMS, cfg, overrides=W.reference_overrides(A.KERNEL_EXACT), precision=32)
    out = np.zeros(np.prod(DIMS) * 4, np.float32)
    scale = float(np.float32(DX) / np.float32(DT))
    o.step(packet, out if steps == 1 else None, scale)
    for s in range(1, steps):
        o.step(None, out if s == steps - 1 else None, scale)
    return out


def solid_box():
    s = np.zeros(DIMS, np.uint8)
    s[8:12, 4:8, 3:6] = 1
    return s


@pytest.fixture()
def exact_lib():
    lib = A.load()
    A.check(lib.aero_lbm_init_rect(*DIMS, 11, 4), "init_rect")
    ov = W.reference_overrides(A.KERNEL_EXACT)
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(ov)), "overrides")
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_CLASSIC_D3Q27)
    yield lib
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(W.reference_overrides(A.KERNEL_FAST))), "overrides")


@pytest.mark.parametrize("mode", [A.SOLVER_WIND_TUNNEL, A.SOLVER_WIND_TUNNEL_PRESSURE_OUTLET, A.SOLVER_CLOSED])
def test_run_wind_tunnel_equals_oracle(oracle_mod, exact_lib, mode):
    lib = exact_lib
    log(f"wind tunnel mode {mode}, {DIMS}, 12 steps")
    solid = solid_box()
    bc = boundary(mode)
    n = solid.size
    out = np.zeros(n * 4, np.float32)
    inp = A.StepInput()
    inp.solid_mask = solid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    inp.prev_flow = None
    inp.grid = A.GridDesc(*DIMS, DX, DT)
    inp.boundary = bc
    inp.steps = 12
    res = A.StepOutput()
    res.flow_out = out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.aero_solver_run_wind_tunnel(ctypes.byref(inp), ctypes.byref(res)) == 1, lib.aero_solver_last_error()
    assert res.status == 1
    # the benchmark configuration the layer left set equals the reference layer's
    got = A.BenchmarkConfig()
    assert lib.aero_lbm_benchmark_get_config(ctypes.byref(got)) == 1
    assert lib.aero_lbm_benchmark_set_config(ctypes.byref(tunnel_config(bc))) == 1
    want = A.BenchmarkConfig()
    assert lib.aero_lbm_benchmark_get_config(ctypes.byref(want)) == 1
    assert bytes(got) == bytes(want)
    ref = oracle_flow(oracle_mod, tunnel_config(bc), tunnel_packet(solid, bc), 12)
    assert np.array_equal(out, ref), np.abs(out - ref).max()
    speed = np.sqrt((out.reshape(-1, 4)[:, :3].astype(np.float32) ** 2).sum(1))
    assert np.isclose(res.max_velocity, speed.max(), rtol=1e-6)
    assert np.all(out.reshape(-1, 4)[solid.reshape(-1) != 0] == 0)


def test_handle_api_flow_state_and_errors(oracle_mod, exact_lib):
    lib = exact_lib
    solid = solid_box()
    n = solid.size
    h = ctypes.c_longlong(0)
    assert lib.aero_solver_create(*DIMS, DX, DT, ctypes.byref(h)) == 1
    u8 = ctypes.POINTER(ctypes.c_uint8)
    fp = ctypes.POINTER(ctypes.c_float)
    assert lib.aero_solver_set_solid_mask(h, solid.ctypes.data_as(u8), n - 1) == 0
    assert b"invalid solid mask buffer" in lib.aero_solver_last_error()
    assert lib.aero_solver_set_solid_mask(h, solid.ctypes.data_as(u8), n) == 1
    rng = np.random.default_rng(3)
    flow = np.zeros((n, 4), np.float32)
    flow[:, 0] = 1.0 + rng.uniform(-0.1, 0.1, n)
    flow[:, 1] = rng.uniform(-0.1, 0.1, n)
    flow[:, 3] = rng.uniform(-0.002, 0.002, n)
    assert lib.aero_solver_set_flow_state(h, flow.ctypes.data_as(fp), n * 4) == 1
    bc = boundary(A.SOLVER_WIND_TUNNEL)
    assert lib.aero_solver_advance_wind_tunnel(h, ctypes.byref(bc), 3) == 1
    out = np.zeros(n * 4, np.float32)
    assert lib.aero_solver_step_wind_tunnel(h, ctypes.byref(bc), 2, out.ctypes.data_as(fp), n * 4 + 4) == 0
    assert b"invalid output flow buffer" in lib.aero_solver_last_error()
    assert lib.aero_solver_step_wind_tunnel(h, ctypes.byref(bc), 0, out.ctypes.data_as(fp), n * 4) == 0
    assert b"steps must be positive" in lib.aero_solver_last_error()
    assert lib.aero_solver_step_wind_tunnel(h, ctypes.byref(bc), 2, out.ctypes.data_as(fp), n * 4) == 1
    # the flow state leaves the packet's obstacle channel as the last rebuild left it
    # (all fluid at creation): the mask set before it only zeroes the solid cells'
    # flow, as in the reference (aero_solver_capi.cpp:442-470 clears packet_dirty)
    ref = oracle_flow(oracle_mod, tunnel_config(bc), tunnel_packet(solid, bc, flow, obstacle=False), 5)
    assert np.array_equal(out, ref), np.abs(out - ref).max()
    bad = boundary(9)
    assert lib.aero_solver_advance_wind_tunnel(h, ctypes.byref(bad), 1) == 0
    assert b"invalid wind-tunnel boundary mode" in lib.aero_solver_last_error()
    assert lib.aero_solver_runtime_info().startswith(b"hip|")
    lib.aero_solver_destroy(h)
    assert lib.aero_solver_advance_wind_tunnel(h, None, 1) == 0
    assert lib.aero_solver_finish() == 1
