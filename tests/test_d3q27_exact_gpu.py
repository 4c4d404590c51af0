"""GPU parity of the exact FP32 path against the FP32 oracle (bit-exact), through the C ABI."""
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
    lib = A.load()
    return lib


def setup(lib, n, cfg, ov=None, mode=A.SOLVER_CLASSIC_D3Q27, in_ch=11, out_ch=4, dims=None):
    nx, ny, nz = dims if dims else (n, n, n)
    A.check(lib.aero_lbm_init_rect(nx, ny, nz, in_ch, out_ch), "init_rect")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "set_config")
    ov = ov if ov is not None else W.reference_overrides(A.KERNEL_EXACT)
    ov.kernel_variant = A.KERNEL_EXACT
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(ov)), "set_overrides")
    lib.aero_lbm_set_realtime_solver_mode(mode)
    return A.Context(nx, ny, nz, key=11, in_ch=in_ch, out_ch=out_ch)


def run_both(oracle_mod, lib, cfg, packet, steps, n=None, dims=None, ov=None, temperature=None):
    dims = dims or (n, n, n)
    ctx = setup(lib, None, cfg, ov=ov, dims=dims)
    o = oracle_mod.Oracle(*dims, cfg, overrides=ov, precision=32)
    if temperature is not None:
        ctx.set_temperature(temperature)
        o.set_temperature(temperature)
    out_g = np.zeros(ctx.cells * 4, np.float32)
    out_o = np.zeros(ctx.cells * 4, np.float32)
    ctx.step(packet, out_g)
    o.step(packet, out_o)
    assert np.array_equal(out_g, out_o), "step-1 output differs"
    for _ in range(steps - 1):
        ctx.step(None)
        o.step(None)
    return ctx, o


@pytest.mark.parametrize("steps", [1, 2, 7])
def test_exact_channel_small(oracle_mod, lib, steps):
    log(f"exact channel 12^3 steps={steps}")
    cfg, pk, _ = W.config1(12)
    ctx, o = run_both(oracle_mod, lib, cfg, pk, steps, n=12)
    g, r = ctx.flow(), o.flow()
    assert np.array_equal(g, r), np.max(np.abs(g - r))
    ctx.release()


def test_exact_boussinesq_blob_gravity_z(oracle_mod, lib):
    log("exact boussinesq 16^3 gravity -z 40 steps")
    n = 16
    cfg, ov, pk, t, _ = W.config3(n, beta=0.1, cs=0.1)
    ov.kernel_variant = A.KERNEL_EXACT
    ctx, o = run_both(oracle_mod, lib, cfg, pk, 40, n=n, ov=ov, temperature=t)
    assert np.array_equal(ctx.flow(), o.flow())
    assert np.array_equal(ctx.temperature(), o.temperature())
    assert np.abs(ctx.temperature()).max() > 1e-3
    ctx.release()


def heated_cavity_case(dims, seed=7):
    """Cavity state for the heated-cavity presets: a solid block, a linear hot-to-cold
    profile along x with noise, small random velocities (reference velocity)."""
    nx, ny, nz = dims
    solid = np.zeros(dims, np.uint8)
    solid[nx // 2 - 2:nx // 2 + 1, ny // 3:ny // 3 + 3, nz // 3:nz // 3 + 3] = 1
    rng = np.random.default_rng(seed)
    x = np.arange(nx, dtype=np.float64)[:, None, None]
    t = (0.5 - (x + 0.5) / nx) + 0.05 * rng.uniform(-1.0, 1.0, dims)
    t = np.broadcast_to(t, dims).astype(np.float32).copy()
    t[solid.astype(bool)] = 0.0
    packet = W.make_packet(solid, u0=(0.0, 0.0, 0.0), perturb=0.02, seed_u=5, temperature=t)
    return packet, t


@pytest.mark.parametrize("three_d,length,steps", [(True, 1.0, 30), (False, 16.0, 30), (True, 16.0, 200)])
def test_exact_heated_cavity_bfecc(oracle_mod, lib, three_d, length, steps):
    dims = (16, 12, 10)
    log(f"exact heated cavity {'3d' if three_d else '2d'} L={length} {dims} {steps} steps (BFECC temperature)")
    cfg = W.heated_cavity_config(three_d, reference_length=length)
    packet, t = heated_cavity_case(dims)
    ov = W.reference_overrides(A.KERNEL_EXACT)
    ctx, o = run_both(oracle_mod, lib, cfg, packet, steps, dims=dims, ov=ov, temperature=t)
    assert np.array_equal(ctx.temperature(), o.temperature())
    assert np.array_equal(ctx.flow(), o.flow())
    # buoyancy has set the cavity in motion
    assert np.abs(ctx.flow().reshape(-1, 4)[:, :3]).max() > 1e-4
    ctx.release()


def last_force(lib, key):
    f = [ctypes.c_float(), ctypes.c_float(), ctypes.c_float()]
    A.check(lib.aero_lbm_get_last_force(key, *[ctypes.byref(v) for v in f]), "get_last_force")
    return np.array([v.value for v in f], np.float32)


@pytest.mark.parametrize("dims,steps", [((48, 21, 3), 60), ((90, 33, 2), 150)])
def test_exact_cylinder_crossflow_flow_and_drag(oracle_mod, lib, dims, steps):
    log(f"exact cylinder crossflow {dims} {steps} steps (parabolic inlet, drag)")
    cfg = W.cylinder_config()
    solid = W.cylinder_mask(*dims)
    packet = W.make_packet(solid, u0=(0.08, 0.0, 0.0), perturb=1e-3, seed_u=3)
    ov = W.reference_overrides(A.KERNEL_EXACT)
    ctx = setup(lib, None, cfg, ov=ov, dims=dims)
    o = oracle_mod.Oracle(*dims, cfg, overrides=ov, precision=32)
    ctx.step(packet)
    o.step(packet)
    fg, fo = last_force(lib, ctx.key), o.force()
    assert np.allclose(fg, fo, rtol=1e-6, atol=1e-9), (fg, fo)
    for _ in range(steps - 1):
        ctx.step(None)
        o.step(None)
    assert np.array_equal(ctx.flow(), o.flow())
    fg, fo = last_force(lib, ctx.key), o.force()
    assert np.allclose(fg, fo, rtol=1e-6, atol=1e-9), (fg, fo)
    assert abs(fo[0]) > 1e-5  # the drag integral is live (its sign is transient this early)
    ctx.release()


@pytest.mark.parametrize("preset", ["heated_cavity", "cylinder"])
def test_presets_reject_fast_kernels(lib, preset):
    dims = (12, 9, 8)
    log(f"{preset} with the fast kernels: step is refused")
    if preset == "heated_cavity":
        cfg = W.heated_cavity_config(True)
        packet, t = heated_cavity_case(dims)
    else:
        cfg = W.cylinder_config()
        packet = W.make_packet(W.cylinder_mask(*dims), u0=(0.08, 0.0, 0.0))
    ov = W.reference_overrides(A.KERNEL_FAST)
    A.check(lib.aero_lbm_init_rect(*dims, 11, 4), "init_rect")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "set_config")
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(ov)), "set_overrides")
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_CLASSIC_D3Q27)
    ctx = A.Context(*dims, key=12)
    with pytest.raises(A.AeroLbmError, match="exact FP32"):
        ctx.step(packet)
    ctx.release()


@pytest.mark.parametrize("variant", ["symmetry", "walls", "pressure", "inherit", "inherit_noconv", "sponge"])
def test_exact_face_kinds(oracle_mod, lib, variant):
    log(f"exact faces {variant}")
    dims = (10, 9, 8)
    cfg = W.channel_config(0.04, 40.0, sgs=True, buoyancy=True)
    if variant == "symmetry":
        for f in (cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
            f.hydrodynamic_kind = A.HYDRO_SYMMETRY
            f.thermal_kind = A.THERMAL_NEUMANN
            f.heat_flux = 0.01
    elif variant == "walls":
        cfg.y_min.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
        cfg.y_max.hydrodynamic_kind = A.HYDRO_MOVING_WALL
        cfg.y_max.velocity[0] = 0.03
        cfg.z_min.thermal_kind = A.THERMAL_DIRICHLET
        cfg.z_min.temperature = 0.2
    elif variant == "pressure":
        cfg.x_max.hydrodynamic_kind = A.HYDRO_PRESSURE
        cfg.x_max.pressure = -0.01
    elif variant in ("inherit", "inherit_noconv"):
        cfg.x_max.hydrodynamic_kind = A.HYDRO_INHERIT
        cfg.z_max.hydrodynamic_kind = A.HYDRO_INHERIT
        if variant == "inherit_noconv":
            cfg.flags |= A.FLAG_DISABLE_CONVECTIVE_OUTFLOW
    elif variant == "sponge":
        cfg.flags &= ~A.FLAG_DISABLE_SPONGE
    solid = W.solid_boxes(*dims, 0.12, seed=21, edge=(1, 3), keep_clear_x=1)
    t = W.hot_blob(*dims, 0.05, 3.0, seed=4, axis=1)
    pk = W.make_packet(solid, u0=(0.04, 0.0, 0.0), perturb=2e-3, seed_u=5, temperature=t,
                       source=0.001 * np.ones(dims, np.float32))
    pk.reshape(-1, 11)[:, 8] = 0.004 * (W.splitmix_array(77, solid.size) - 0.5)
    ctx, o = run_both(oracle_mod, lib, cfg, pk, 25, dims=dims, temperature=t)
    assert np.array_equal(ctx.flow(), o.flow())
    assert np.array_equal(ctx.temperature(), o.temperature())
    ctx.release()


def test_exact_outputs_atlas_sample_scaled(oracle_mod, lib):
    log("exact atlas/sample/scaled 12^3")
    cfg, pk, solid = W.config1(12)
    ctx, o = run_both(oracle_mod, lib, cfg, pk, 5, n=12)
    for stride in (1, 2, 5):
        assert np.array_equal(ctx.atlas(stride), o.atlas(stride)), stride
    fl = o.flow().reshape(12, 12, 12, 4)
    for (x, y, z) in ((0, 0, 0), (11, 5, 3), (6, 6, 6)):
        assert np.array_equal(ctx.sample(x, y, z), fl[x, y, z])
    out_g = np.zeros(ctx.cells * 4, np.float32)
    out_o = np.zeros(ctx.cells * 4, np.float32)
    ctx.step(None, out_g, scale=20.0)
    o.step(None, out_o, scale=20.0)
    assert np.array_equal(out_g, out_o)
    ctx.release()


def test_exact_payload_change_mid_run(oracle_mod, lib):
    log("exact payload change 12^3")
    cfg, pk, solid = W.config1(12)
    ctx, o = run_both(oracle_mod, lib, cfg, pk, 4, n=12)
    solid2 = solid.copy()
    solid2[5:7, 2:9, 2:9] = 1
    solid2[3:4, :, :] = 0
    pk2 = W.make_packet(solid2, seed_u=3)
    for _ in range(4):
        ctx.step(pk2)
        o.step(pk2)
    assert np.array_equal(ctx.flow(), o.flow())
    ctx.release()


def test_exact_config1_full_1000_steps(lib):
    """Plain D3Q27 cumulant 32^3, 15% solids, nu 1e-3, SGS on, 1000 steps, against the
    FP32 oracle fixture (tests/golden/config1.json, made by make_golden.py)."""
    import hashlib
    import json
    import os
    log("exact config1 32^3 x 1000 steps vs golden fixture")
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config1.json")))
    cfg, pk, solid = W.config1(32)
    ctx = setup(lib, 32, cfg)
    ctx.step(pk)
    for _ in range(gold["steps"] - 1):
        ctx.step(None)
    g = ctx.flow()
    idx = gold["fp32"]["probe_cells"]
    probe = np.array(gold["fp32"]["probe_flow"], dtype=np.float32)
    got = g.reshape(-1, 4)[idx]
    # bar: density and velocity within 1e-5 relative of the reference CPU path; this path is bit-exact
    scale = np.maximum(np.abs(probe), 1e-3)
    assert np.max(np.abs(got - probe) / scale) <= 1e-5
    assert hashlib.sha256(g.tobytes()).hexdigest() == gold["fp32"]["flow_sha256"]
    ctx.release()
