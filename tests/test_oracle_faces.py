"""Oracle pins for the benchmark-mode face rules and the plain semi-Lagrangian
temperature advection (ref: native/src/aero_lbm_jni.cpp:1494-1602 face values,
2145-2188 first-order advection of update_temperature_field).

Every check comes from what the mathematics fixes, not from the rules retyped:

* first-order semi-Lagrangian advection with uniform velocity c multiplies one
  periodic Fourier mode by a = 1 - |c| + |c| e^{-ik sgn c} per axis (linear
  interpolation from the departure point), and a mode along two axes by the
  product of the two factors;
* steady Couette flow between a halfway bounce-back wall and a moving-wall face
  (incoming links set to the equilibrium at the wall velocity) is exactly linear,
  vanishes half a cell outside the first fluid cell, and reaches the wall speed tau
  cells beyond the last one: the x-momentum flux through the moving face is
  (U_w - u_top)/6 + (tau - 1) s/6 for the D3Q27 weights, and in steady state it
  equals the viscous flux nu s = (tau - 1/2) s/3, so U_w = u_top + tau s;
* pressure faces (equilibrium at the face density, zero velocity) drive a channel
  whose developed section is the Hagen-Poiseuille parabola u_max = G h^2/(8 rho nu),
  G = -dp/dx with p = rho/3, between walls half a cell outside the fluid;
* on a state that does not vary along a face normal, the symmetry rule (copy the
  same direction from the clamped in-domain cell) gives exactly what a periodic
  face gives.

The convective-outflow rule (f1 + beta (f1 - f2) from the two cells inward of the
cell along the normal, beta = 0.15) is a heuristic extrapolation: a uniform flow is
its fixed point (pinned in test_oracle_d3q27.py), but neither beta nor the choice
of the cell's own column is fixed by a closed form, so it stays parity unpinned.
"""
import math

import numpy as np
import pytest

import aero_lbm as A
import workloads as W


@pytest.fixture(autouse=True)
def single_thread(oracle_mod):
    """These grids hold a few hundred cells: one thread avoids per-step thread
    start-up costs (results do not depend on the thread count)."""
    oracle_mod.set_threads(1)
    yield
    oracle_mod.set_threads(min(16, __import__("os").cpu_count() or 1))


def base_cfg(speed, reynolds):
    cfg = W.channel_config(speed, reynolds, sgs=False, buoyancy=False)
    cfg.initial_velocity[0] = 0.0
    return cfg


def set_face(f, hydro, thermal):
    f.hydrodynamic_kind = hydro
    f.thermal_kind = thermal


def all_periodic(cfg):
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        set_face(f, A.HYDRO_PERIODIC, A.THERMAL_PERIODIC)


# ---------------------------------------------------------------- semi-Lagrangian advection
def advect_once(oracle_mod, t0, u0):
    """One benchmark-mode step (plain semi-Lagrangian temperature update, no
    diffusion, no source) from temperature t0 with uniform velocity u0."""
    cfg = base_cfg(0.05, 50.0)
    all_periodic(cfg)
    ov = W.reference_overrides(A.KERNEL_EXACT)
    ov.thermal_diffusivity = 0.0
    nx, ny, nz = t0.shape
    packet = W.make_packet(np.zeros(t0.shape, np.uint8), u0=u0, perturb=0.0, temperature=t0)
    o = oracle_mod.Oracle(nx, ny, nz, cfg, overrides=ov, precision=64)
    o.set_temperature(t0.reshape(-1).astype(np.float32))
    o.step(packet)
    c = packet.reshape(-1, 11)[0, 5:8].astype(np.float64)  # the float32 velocity advected with
    return o.temperature_r().reshape(t0.shape), c


def lin_factor(k, c):
    return 1.0 - abs(c) + abs(c) * np.exp(-1j * k * math.copysign(1.0, c)) if c != 0.0 else 1.0 + 0j


@pytest.mark.parametrize("axis,c,mode", [(0, 0.5, 1), (1, -0.3, 2), (2, 0.7, 1), (0, 1.0, 3), (2, -0.85, 2)])
def test_semi_lagrangian_single_mode_factor(oracle_mod, axis, c, mode):
    shape = [6, 6, 6]
    shape[axis] = 24
    n = shape[axis]
    k = 2.0 * math.pi * mode / n
    coord = np.arange(n, dtype=np.float64)
    prof = (0.3 * np.cos(k * coord + 0.4)).astype(np.float32)
    bshape = [1, 1, 1]
    bshape[axis] = n
    t0 = np.broadcast_to(prof.reshape(bshape), shape).astype(np.float32).copy()
    u0 = [0.0, 0.0, 0.0]
    u0[axis] = c
    t1, cf = advect_once(oracle_mod, t0, tuple(u0))
    a = lin_factor(k, float(cf[axis]))
    want = np.real(a * np.fft.fft(prof.astype(np.float64))[mode] * 2.0 / n * np.exp(1j * k * coord))
    got = np.moveaxis(t1, axis, 0)[:, 0, 0]
    assert np.max(np.abs(got - want)) < 1e-7
    assert np.max(np.abs(t1 - np.broadcast_to(got.reshape(bshape), shape))) < 1e-12


def test_semi_lagrangian_two_axis_mode_is_product_of_factors(oracle_mod):
    nx, ny, nz = 16, 12, 4
    kx, ky = 2.0 * math.pi / nx, 2.0 * math.pi * 2 / ny
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    phase = kx * x + ky * y
    t0 = np.repeat((0.25 * np.cos(phase))[:, :, None], nz, axis=2).astype(np.float32)
    t1, cf = advect_once(oracle_mod, t0, (0.4, -0.6, 0.0))
    G = lin_factor(kx, float(cf[0])) * lin_factor(ky, float(cf[1]))
    want = np.real(G * np.exp(1j * phase)) * 0.25
    # float32 input: the mode's coefficient is 0.25 up to input rounding
    assert np.max(np.abs(t1[:, :, 0] - want)) < 1e-7
    # a transposed velocity (factors swapped between the axes) is far off
    G_swapped = lin_factor(kx, float(cf[1])) * lin_factor(ky, float(cf[0]))
    assert np.max(np.abs(np.real(G_swapped * np.exp(1j * phase)) * 0.25 - want)) > 1e-2


# ---------------------------------------------------------------- moving wall (Couette)
@pytest.mark.parametrize("uw,reynolds,ny", [(0.02, 0.2, 8), (0.03, 0.5, 10)])
def test_couette_between_bounce_back_and_moving_wall(oracle_mod, uw, reynolds, ny):
    nx, nz = 2, 2
    cfg = base_cfg(uw, reynolds)
    all_periodic(cfg)
    set_face(cfg.y_min, A.HYDRO_BOUNCE_BACK, A.THERMAL_ADIABATIC)
    set_face(cfg.y_max, A.HYDRO_MOVING_WALL, A.THERMAL_ADIABATIC)
    cfg.y_max.velocity[0] = uw
    packet = W.make_packet(np.zeros((nx, ny, nz), np.uint8), u0=(0.0, 0.0, 0.0), perturb=0.0)
    o = oracle_mod.Oracle(nx, ny, nz, cfg, precision=64)
    tau, nu = o.params()["tau_shear"], o.params()["nu"]
    assert nu == pytest.approx(uw / reynolds, rel=1e-6)
    o.run(packet, int(16 * ny * ny / nu))
    u = o.velocity().reshape(nx, ny, nz, 3)
    prof = u[0, :, 0, 0]
    assert np.max(np.abs(u[..., 1:])) < 1e-12 * uw
    assert np.max(np.abs(u - u[:1, :, :1, :])) < 1e-12 * uw
    y = np.arange(ny, dtype=np.float64)
    s = uw / ((ny - 1) + tau + 0.5)
    want = s * (y + 0.5)
    # 1e-6: the balance above is exact at density 1; the steady density differs from
    # 1 by O(U_w^2 / ny^2), which shifts the profile by ~1e-8 of U_w
    assert np.max(np.abs(prof - want)) < 1e-6 * uw
    assert prof[-1] + tau * s == pytest.approx(uw, rel=1e-6)


# ---------------------------------------------------------------- pressure faces (Poiseuille)
@pytest.mark.parametrize("p_in,p_out", [(0.01, -0.01), (-0.008, 0.012)])
def test_pressure_faces_drive_poiseuille_channel(oracle_mod, p_in, p_out):
    nx, ny, nz = 48, 8, 2
    cfg = base_cfg(0.02, 0.2)
    all_periodic(cfg)
    set_face(cfg.x_min, A.HYDRO_PRESSURE, A.THERMAL_ADIABATIC)
    set_face(cfg.x_max, A.HYDRO_PRESSURE, A.THERMAL_ADIABATIC)
    cfg.x_min.pressure, cfg.x_max.pressure = p_in, p_out
    cfg.x_max.velocity[0] = 0.02  # sets the reference speed (nu = 0.1); the pressure rule ignores it
    set_face(cfg.y_min, A.HYDRO_BOUNCE_BACK, A.THERMAL_ADIABATIC)
    set_face(cfg.y_max, A.HYDRO_BOUNCE_BACK, A.THERMAL_ADIABATIC)
    packet = W.make_packet(np.zeros((nx, ny, nz), np.uint8), u0=(0.0, 0.0, 0.0), perturb=0.0)
    o = oracle_mod.Oracle(nx, ny, nz, cfg, precision=64)
    nu = o.params()["nu"]
    o.run(packet, 3000)
    u = o.velocity().reshape(nx, ny, nz, 3)
    rho = o.populations().reshape(27, nx, ny, nz).sum(axis=0)
    mid = nx // 2
    rline = rho[:, :, 0].mean(axis=1)
    G = -(rline[mid + 8] - rline[mid - 8]) / 16.0 / 3.0
    prof = u[mid, :, 0, 0]
    y = np.arange(ny, dtype=np.float64)
    c = np.polyfit(y, prof, 2)
    peak = np.polyval(c, (ny - 1) / 2.0)
    assert math.copysign(1.0, peak) == math.copysign(1.0, p_in - p_out)
    assert np.max(np.abs(np.polyval(c, y) - prof)) < 1e-3 * abs(peak)
    roots = np.sort(np.real(np.roots(c)))
    assert abs(roots[0] + 0.5) < 0.02 and abs(roots[1] - (ny - 0.5)) < 0.02
    assert peak == pytest.approx(G * ny * ny / (8.0 * rline[mid] * nu), rel=1e-2)
    assert np.max(np.abs(u[mid, :, :, 1:])) < 1e-6 * abs(peak)


# ---------------------------------------------------------------- symmetry and outflow faces
@pytest.mark.parametrize("kind", [A.HYDRO_SYMMETRY])
def test_normal_uniform_state_face_rule_equals_periodic(oracle_mod, kind):
    """A state with no variation along y: symmetry faces on y equal periodic y faces
    bit for bit (FP64), step after step."""
    nx, ny, nz = 10, 5, 7
    rng = np.random.default_rng(3)
    solid2 = (rng.uniform(size=(nx, nz)) < 0.12).astype(np.uint8)
    solid = np.repeat(solid2[:, None, :], ny, axis=1)
    u = np.zeros((nx, ny, nz, 3), np.float32)
    u[..., 0] = np.repeat(rng.uniform(0.02, 0.06, (nx, 1, nz)), ny, axis=1)
    u[..., 2] = np.repeat(rng.uniform(-0.02, 0.02, (nx, 1, nz)), ny, axis=1)
    u[solid.astype(bool)] = 0.0
    packet = W.make_packet(solid, u0=(0.0, 0.0, 0.0), perturb=0.0).reshape(nx, ny, nz, 11)
    packet[..., 5:8] = u
    packet = packet.reshape(-1)
    flows = {}
    for name, k in (("face", kind), ("periodic", A.HYDRO_PERIODIC)):
        cfg = W.channel_config(0.05, 50.0, sgs=True, buoyancy=False)
        cfg.initial_velocity[0] = 0.0
        for f in (cfg.y_min, cfg.y_max):
            set_face(f, k, A.THERMAL_PERIODIC if k == A.HYDRO_PERIODIC else A.THERMAL_ADIABATIC)
        o = oracle_mod.Oracle(nx, ny, nz, cfg, precision=64)
        o.run(packet, 12)
        flows[name] = (o.populations(), o.flow())
    assert np.array_equal(flows["face"][0], flows["periodic"][0])
    assert np.array_equal(flows["face"][1], flows["periodic"][1])
    f4 = flows["face"][1].reshape(nx, ny, nz, 4)
    assert np.max(np.abs(f4 - f4[:, :1])) == 0.0
    assert np.max(np.abs(f4[..., 0])) > 0.01
