"""Invariants of the FP16 in-place path through the wind-tunnel C ABI, as the
reference's own FP16 check states them (native/tools/check_d3q27_fp16.py:156-216):
a 16^3 empty tunnel at 2 m/s (dx 1 m, dt 0.05 s) keeps vx in [1.85, 2.20] m/s and
|vy|, |vz| <= 0.08 m/s over 8 frames with adjacent frames within 0.08 m/s (the A-A
parity flip shows no jump), and a sphere of radius 0.12 n at (n/3, n/2, n/2) gives,
after 10 frames, finite output, zero flow (<= 1e-4) at solid cells, speeds <= 8 m/s
and a mean vx above 0.2 m/s. The FP16 storage is selected the way the reference's
tool selects it: the solver mode left at auto and AERO_LBM_D3Q27_FP16_INPLACE=1
(native/src/aero_lbm_jni.cpp:5730-5764).
"""
import ctypes
import sys

import numpy as np
import pytest

import aero_lbm as A

pytestmark = pytest.mark.gpu

N = 16


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sphere_mask(n):
    x, y, z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = max(1.0, n * 0.12)
    return (((x - n // 3) ** 2 + (y - n // 2) ** 2 + (z - n // 2) ** 2) <= r * r).astype(np.uint8).reshape(-1)


def ok(v, lib, what):
    if not v:
        msg = lib.aero_solver_last_error()
        raise A.AeroLbmError(f"{what} failed: {msg.decode() if msg else ''}")


def frames(lib, mask, count):
    """count single-step frames of the wind tunnel (inlet 2 m/s along x, air)."""
    cells = N ** 3
    h = ctypes.c_longlong(0)
    ok(lib.aero_solver_create(N, N, N, 1.0, 0.05, ctypes.byref(h)), lib, "create")
    try:
        m = np.ascontiguousarray(mask, np.uint8)
        ok(lib.aero_solver_set_solid_mask(h.value, m.ctypes.data_as(A._U8), cells), lib, "mask")
        bc = A.BoundaryDesc(A.SOLVER_WIND_TUNNEL, 2.0, 0.0, 0.0, 0.0, 1.225, 1.5e-5)
        out = np.zeros(cells * 4, np.float32)
        snaps = []
        for _ in range(count):
            ok(lib.aero_solver_step_wind_tunnel(h.value, ctypes.byref(bc), 1, A.fptr(out), out.size),
                           lib, "step")
            snaps.append(out.copy())
        return snaps
    finally:
        lib.aero_solver_destroy(h.value)


@pytest.fixture()
def fp16_lib(monkeypatch):
    lib = A.load()
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_AUTO)
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(A.default_overrides())), "overrides")
    monkeypatch.setenv("AERO_LBM_D3Q27_FP16_INPLACE", "1")
    yield lib
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_AUTO)


def test_uniform_tunnel_band_and_parity(fp16_lib, monkeypatch):
    snaps = frames(fp16_lib, np.zeros(N ** 3, np.uint8), 8)
    for i, v in enumerate(snaps, start=1):
        f = v.reshape(-1, 4)
        assert np.all(np.isfinite(v)), i
        log(f"uniform frame {i}: vx in [{f[:, 0].min():.5f}, {f[:, 0].max():.5f}], "
            f"max |v_perp| {np.abs(f[:, 1:3]).max():.2e}")
        assert f[:, 0].min() >= 1.85 and f[:, 0].max() <= 2.20, i
        assert np.abs(f[:, 1:3]).max() <= 0.08, i
    delta = float(np.abs(snaps[-1] - snaps[-2]).max())
    log(f"adjacent-frame max delta {delta:.3e}")
    assert delta <= 0.08
    # the switch selected the 16-bit storage: the FP32 path ends elsewhere by rounding
    monkeypatch.delenv("AERO_LBM_D3Q27_FP16_INPLACE")
    ref = frames(fp16_lib, np.zeros(N ** 3, np.uint8), 8)[-1]
    diff = float(np.abs(ref - snaps[-1]).max())
    log(f"FP16 vs FP32 storage after 8 frames: {diff:.3e}")
    assert 0.0 < diff < 1e-2


def test_sphere_obstacle(fp16_lib):
    mask = sphere_mask(N)
    v = frames(fp16_lib, mask, 10)[-1]
    f = v.reshape(-1, 4)
    speed = np.sqrt((f[:, :3].astype(np.float64) ** 2).sum(1))
    solid = mask.astype(bool)
    log(f"sphere: {solid.sum()} solid cells, max solid speed {speed[solid].max():.2e}, max speed {speed.max():.3f}, "
        f"mean vx {f[:, 0].mean():.3f}")
    assert np.all(np.isfinite(v))
    assert np.all(speed[solid] <= 1e-4)
    assert speed.max() <= 8.0
    assert f[:, 0].mean() > 0.2
