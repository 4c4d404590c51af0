"""GPU D3Q19 mesoscale step through the C ABI against the mesoscale oracle.

The GPU step follows the reference CPU step's operation order (mesoscale.cpp:
548-738), so its 6-channel state must equal the FP32 oracle's bit for bit;
the FP64 oracle gives the error of the FP32 arithmetic itself.
"""
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


def oracle_for(oracle_mod, cfg, precision=32, tau_min=0.58):
    return oracle_mod.MesoOracle(cfg.nx, cfg.ny, cfg.nz, dx=cfg.dx_m, dt=cfg.dt_s, nu=cfg.molecular_nu_m2_s,
                                 pr=cfg.prandtl_air, prt=cfg.turbulent_prandtl, tau_min=tau_min, precision=precision)


def terrain_case(dims, seed=3, height=6.0, angle=30.0, humidity=0.0):
    F = W.meso_forcing(*dims, seed=seed, max_height=height, angle_deg=angle).reshape(*dims, 25)
    if humidity:
        F[..., 9] = humidity
        F[..., 10] = 0.3   # convective heating
        F[..., 18] = 0.05  # tornado updraft
        F[..., 23] = 1.5   # nested surface delta
    return F.reshape(-1)


@pytest.fixture(scope="module")
def lib():
    return A.load()


@pytest.mark.parametrize("dims", [(24, 12, 20), (17, 9, 13)])
def test_meso_bit_exact_vs_fp32_oracle(oracle_mod, lib, dims):
    cfg = W.meso_config(*dims)
    F = terrain_case(dims, humidity=0.4)
    ctx = A.Mesoscale(cfg)
    o = oracle_for(oracle_mod, cfg)
    for step in range(1, 13):
        log(f"meso {dims} step {step}")
        got = ctx.step(F)
        ref = o.step(F)
        assert np.array_equal(got, ref), (step, float(np.max(np.abs(got - ref))))
    ctx.release()


def test_meso_forcing_change_timed_run_and_state(oracle_mod, lib):
    dims = (20, 10, 16)
    cfg = W.meso_config(*dims)
    F1 = terrain_case(dims, angle=30.0)
    F2 = terrain_case(dims, angle=75.0, humidity=0.2)
    ctx = A.Mesoscale(cfg)
    o = oracle_for(oracle_mod, cfg)
    for _ in range(3):
        assert np.array_equal(ctx.step(F1), o.step(F1))
    log("meso forcing change")
    for _ in range(3):
        assert np.array_equal(ctx.step(F2), o.step(F2))
    ms = ctx.run_timed(F2, 7)
    assert ms > 0.0
    for _ in range(7):
        ref = o.step(F2)
    assert np.array_equal(ctx.state(), ref)
    ctx.release()


def test_meso_uniform_and_varying_planes_bit_exact(oracle_mod, lib):
    """Coefficient planes holding one value everywhere are passed to the kernels as
    parameters; planes that vary are read per cell. Uniform forcing, fully varying
    forcing and back: every step bit-identical to the FP32 oracle."""
    dims = (18, 10, 14)
    cfg = W.meso_config(*dims)
    F1 = terrain_case(dims)
    F2 = W.meso_forcing_varying(*dims, base=terrain_case(dims, humidity=0.2))
    ctx = A.Mesoscale(cfg)
    o = oracle_for(oracle_mod, cfg)
    for k, F in enumerate((F1, F2, F1)):
        log(f"meso plane case {k}")
        for _ in range(3):
            got = ctx.step(F)
            ref = o.step(F)
            assert np.array_equal(got, ref), (k, float(np.max(np.abs(got - ref))))
    ctx.release()


def test_meso_tau_floor_override(oracle_mod, lib):
    """Lattice viscosity 5e-4 (tau = 0.5015) with the tau floor lowered to 0.5015."""
    dims = (16, 10, 16)
    cfg = W.meso_config(*dims, nu_lattice=5e-4)
    tr = A.MesoscaleTransport()
    A.check(lib.aero_lbm_mesoscale_derive_transport(cfg, tr), "derive transport")
    tau = tr.tau_shear_molecular
    assert abs(tau - 0.5015) < 1e-6
    A.check(lib.aero_lbm_mesoscale_set_tau_floor(tau), "tau floor")
    try:
        F = terrain_case(dims)
        ctx = A.Mesoscale(cfg)
        o = oracle_for(oracle_mod, cfg, tau_min=tau)
        for step in range(6):
            assert np.array_equal(ctx.step(F), o.step(F)), step
        ctx.release()
    finally:
        A.check(lib.aero_lbm_mesoscale_set_tau_floor(math.nan), "tau floor reset")


def test_meso_fp32_error_vs_fp64(oracle_mod, lib):
    dims = (32, 16, 32)
    cfg = W.meso_config(*dims)
    F = terrain_case(dims, height=8.0)
    ctx = A.Mesoscale(cfg)
    o64 = oracle_for(oracle_mod, cfg, precision=64)
    for _ in range(60):
        got = ctx.step(F).reshape(-1, 6)
        ref = o64.step(F).reshape(-1, 6)
    err_t = float(np.max(np.abs(got[:, :3] - ref[:, :3])))
    err_w = float(np.max(np.abs(got[:, 3:] - ref[:, 3:])))
    wind = float(np.max(np.abs(ref[:, 3:])))
    log(f"meso FP32 vs FP64 after 60 steps: temperature {err_t:.3e} K, wind {err_w:.3e} m/s (max wind {wind:.3f})")
    assert err_t < 2e-3 and err_w < 1e-3 * max(wind, 1e-3) * 10
    ctx.release()
