"""Oracle pins for the reference's runtime (non-benchmark) mode and fan forcing
(ref: native/src/aero_lbm_jni.cpp:675-693 state nudge, 896-993 runtime parameters,
1971-2143 runtime thermal model, 2191-2196 update stride, 2258-2314 fan force and
nudge, 8037-8064 temperature reference, 8480-8507 step order).

What the mathematics fixes, checked on the FP64 oracle:

* the runtime parameters are the physical constants of air at 1 m cells and 0.05 s
  steps: nu0 = 1.5e-5 * 0.05, alpha0 = (1.5e-5 / 0.71) * 0.05, tau = 1/2 + 3 nu0,
  beta = 9.81 / 300 * 20 * 0.05^2;
* from rest, one step toward a new reference velocity U leaves a post-collision
  momentum (1 - s(d)) n(d) U (the collision conserves the nudged momentum), with the
  nudge n(d) = 0.08 (1 - d/8)^2 and the sponge s(d) = 0.03 (1 - d/4)^2 at distance d
  from the nearest face;
* the temperature is updated on even steps only, with time step 2: one periodic
  Fourier mode advected by a uniform velocity c scales by the BFECC factor of linear
  interpolation at displacement 2c plus 2 alpha0 (2 cos k - 2) of diffusion; on the
  odd step the interior keeps its temperature and a face layer relaxes toward the
  payload temperature by n(d);
* in a uniform shear u_x = g y the thermal diffusivity is alpha0 + min(C^2 g, 5 nu0)/0.85
  (|S| = sqrt(2 S_ij S_ij) = g for this strain), and alpha0 alone next to a solid;
* fan forcing: a uniform state in a periodic box with the fan on every cell changes its
  momentum per step by k F, k = 3/2 - omega/2 (the cumulant collision keeps the
  half-force velocity, and the Guo source adds (1 - omega/2) F), with
  F = 0.07 rho (max(0, T - u_par) n - u_perp), T = 0.05 |v_fan|; above the speed
  0.24 the force is divided by 1 + ((|u| - 0.24)/0.04)^2; the noise scales T by
  1 + 0.02 xi with xi in [-1, 1] of mean 0.

The exact noise sequence (a SplitMix64 finalizer of cell and step counter) is the
reference's own choice of hash and is held only to these bounds: parity unpinned.
"""
import math

import numpy as np
import pytest

import aero_lbm as A
import workloads as W


@pytest.fixture(autouse=True)
def few_threads(oracle_mod):
    oracle_mod.set_threads(2)
    yield
    oracle_mod.set_threads(min(16, __import__("os").cpu_count() or 1))


def runtime_cfg(periodic=True):
    cfg = W.channel_config(0.05, 50.0, sgs=True, buoyancy=True)
    cfg.enabled = 0
    cfg.initial_velocity[0] = 0.0
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_PERIODIC if periodic else A.HYDRO_INHERIT
        f.thermal_kind = A.THERMAL_PERIODIC if periodic else A.THERMAL_INHERIT
        f.velocity[0] = 0.0
    return cfg


def dist_to_face(shape):
    idx = np.indices(shape)
    d = np.minimum.reduce([idx[a] for a in range(3)] + [shape[a] - 1 - idx[a] for a in range(3)])
    return d


CX = np.array([q // 9 - 1 for q in range(27)], dtype=np.float64)
CY = np.array([(q // 3) % 3 - 1 for q in range(27)], dtype=np.float64)
CZ = np.array([q % 3 - 1 for q in range(27)], dtype=np.float64)


def test_runtime_parameters_are_air_at_one_metre_and_twentieth_second(oracle_mod):
    o = oracle_mod.Oracle(4, 4, 4, runtime_cfg(), precision=64)
    p = o.params()
    nu0 = 1.5e-5 * 0.05
    assert p["nu"] == pytest.approx(nu0, rel=1e-12)
    assert p["tau_shear"] == pytest.approx(0.5 + 3.0 * nu0, rel=1e-12)
    assert p["tau_normal"] == pytest.approx(0.5 + 3.0 * nu0, rel=1e-12)
    assert p["alpha"] == pytest.approx(1.5e-5 / 0.71 * 0.05, rel=1e-12)
    assert p["beta"] == pytest.approx(9.81 / 300.0 * 20.0 * 0.05 ** 2, rel=1e-12)


def test_state_nudge_and_sponge_profile(oracle_mod):
    n = 20
    shape = (n, n, n)
    solid = np.zeros(shape, np.uint8)
    cfg = runtime_cfg()
    o = oracle_mod.Oracle(n, n, n, cfg, precision=64)
    o.step(W.make_packet(solid, u0=(0.0, 0.0, 0.0), perturb=0.0))   # initialise at rest
    U = (0.02, 0.01, -0.015)
    o.step(W.make_packet(solid, u0=U, perturb=0.0))                 # new reference velocity
    f = o.post_collision()
    j = np.stack([(C[:, None] * f).sum(axis=0) for C in (CX, CY, CZ)], axis=1).reshape(n, n, n, 3)
    d = dist_to_face(shape)
    nudge = np.where(d < 8, 0.08 * ((8 - d) / 8.0) ** 2, 0.0)
    sponge = np.where(d < 4, 0.03 * ((4 - d) / 4.0) ** 2, 0.0)
    # the payload velocity is float32
    want = ((1.0 - sponge) * nudge)[..., None] * np.array(U, np.float32).astype(np.float64)
    assert np.max(np.abs(j - want)) < 1e-15
    assert np.max(np.abs(j[d >= 8])) < 1e-17


def mode_x(n_axis, mode, amp=0.3):
    k = 2.0 * math.pi * mode / n_axis
    return k, (amp * np.cos(k * np.arange(n_axis) + 0.3)).astype(np.float32)


def test_runtime_temperature_stride_two_bfecc_factor_and_face_relaxation(oracle_mod):
    nx, ny, nz = 32, 20, 20
    shape = (nx, ny, nz)
    k, prof = mode_x(nx, 2)
    t0 = np.broadcast_to(prof[:, None, None], shape).astype(np.float32).copy()
    c = 0.15  # displacement 2c = 0.3: the limiter stays inactive on this mode
    solid = np.zeros(shape, np.uint8)
    # runtime mode starts every fluid cell at the payload temperature
    packet = W.make_packet(solid, u0=(c, 0.0, 0.0), perturb=0.0, temperature=t0)
    o = oracle_mod.Oracle(nx, ny, nz, runtime_cfg(), precision=64)
    alpha0 = o.params()["alpha"]
    o.step(packet)
    t1 = o.temperature_r().reshape(shape)
    cf = float(np.float32(c))
    a = 1.0 - 2.0 * cf + 2.0 * cf * np.exp(-1j * k)            # linear interpolation at displacement 2c
    G = a * (3.0 - abs(a) ** 2) / 2.0 + 2.0 * alpha0 * (2.0 * math.cos(k) - 2.0)
    coef = np.fft.fft(prof.astype(np.float64))[2] * 2.0 / nx
    want = np.real(G * coef * np.exp(1j * k * np.arange(nx)))
    assert np.max(np.abs(t1 - want[:, None, None])) < 1e-6 * 0.3
    # odd step: no update; the payload temperature t0 is nudged in within 8 of a face
    o.step(None)
    t2 = o.temperature_r().reshape(shape)
    d = dist_to_face(shape)
    tn = np.where(d < 8, 0.08 * ((8 - d) / 8.0) ** 2, 0.0)
    assert np.max(np.abs(t2 - (t1 * (1.0 - tn) + t0.astype(np.float64) * tn))) < 1e-15
    # even step: updated again (the interior moves)
    o.step(None)
    t3 = o.temperature_r().reshape(shape)
    assert np.max(np.abs(t3 - t2)[d >= 8]) > 1e-3


@pytest.mark.parametrize("g", [2.0e-3, 2.0e-2])
def test_runtime_turbulent_thermal_diffusivity_in_shear(oracle_mod, g):
    nx, ny, nz = 6, 12, 24
    shape = (nx, ny, nz)
    kz = 2.0 * math.pi / nz
    t0 = np.broadcast_to((0.2 * np.cos(kz * np.arange(nz)))[None, None, :], shape).astype(np.float32).copy()
    solid = np.zeros(shape, np.uint8)
    solid[3, 6, 5] = 1
    packet = W.make_packet(solid, u0=(0.0, 0.0, 0.0), perturb=0.0, temperature=t0).reshape(nx, ny, nz, 11)
    packet[..., 5] = (g * np.arange(ny, dtype=np.float32))[None, :, None]
    packet[solid.astype(bool), 5] = 0.0
    o = oracle_mod.Oracle(nx, ny, nz, runtime_cfg(), precision=64)
    p = o.params()
    o.step(packet.reshape(-1))
    t1 = o.temperature_r().reshape(shape)
    t0d = t0.astype(np.float64)
    lap = np.roll(t0d, 1, axis=2) + np.roll(t0d, -1, axis=2) - 2.0 * t0d
    # |S| = |du_x/dy| from the (float32) payload velocities by central differences
    ux = packet[0, :, 0, 5].astype(np.float64)
    dudy = np.zeros(ny)
    dudy[1:-1] = 0.5 * (ux[2:] - ux[:-2])
    nu_t = np.minimum(p["sgs_c2"] * np.abs(dudy), 5.0 * p["nu"])
    alpha_eff = np.broadcast_to((p["alpha"] + nu_t / 0.85)[None, :, None], shape)
    # interior rows (both y neighbours fluid and inside), away from the solid cell
    sel = np.zeros(shape, bool)
    sel[:, 1:ny - 1, :] = True
    near = np.zeros(shape, bool)
    for dx, dy, dz in ((0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        near[3 + dx, 6 + dy, 5 + dz] = True
    ok = sel & ~near
    got = (t1 - t0d)[ok] / (2.0 * lap[ok])
    big = np.abs(lap[ok]) > 1e-3
    assert np.max(np.abs(got[big] - alpha_eff[ok][big]) / alpha_eff[ok][big]) < 1e-8
    assert np.max(np.abs(alpha_eff[ok] - p["alpha"])) > 0.05 * p["alpha"]  # the turbulent part is visible
    # next to the solid cell the strain is not used: alpha0 alone
    for dx, dy, dz in ((1, 0, 0), (0, 0, 1)):
        c = (3 + dx, 6 + dy, 5 + dz)
        if abs(lap[c]) > 1e-3:
            assert (t1[c] - t0d[c]) / (2.0 * lap[c]) == pytest.approx(p["alpha"], rel=1e-9)


def fan_cfg(noise=False):
    cfg = W.channel_config(0.05, 50.0, sgs=False, buoyancy=False)
    cfg.initial_velocity[0] = 0.0
    cfg.flags &= ~A.FLAG_DISABLE_FAN_FORCING
    if noise:
        cfg.flags &= ~A.FLAG_DISABLE_FAN_NOISE
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_PERIODIC
        f.thermal_kind = A.THERMAL_PERIODIC
    cfg.x_min.velocity[0] = 0.05  # reference speed: nu = 1e-3
    return cfg


def fan_packet(shape, u0, vfan, mask=1.0):
    pk = W.make_packet(np.zeros(shape, np.uint8), u0=u0, perturb=0.0).reshape(*shape, 11)
    pk[..., 1] = mask
    pk[..., 2:5] = vfan
    return pk.reshape(-1)


def mean_velocity(o, shape):
    f = o.populations()
    rho = f.sum(axis=0)
    return np.stack([(C[:, None] * f).sum(axis=0) / rho for C in (CX, CY, CZ)], axis=1).reshape(*shape, 3), rho


@pytest.mark.parametrize("speed_perp,vfan_len", [(0.03, 2.0), (0.30, 2.0)])
def test_fan_force_decays_perpendicular_velocity(oracle_mod, speed_perp, vfan_len):
    shape = (4, 4, 4)
    n = np.array([0.6, 0.8, 0.0])
    perp = np.array([-0.8, 0.6, 0.0])
    u0 = tuple((speed_perp * perp).astype(np.float32))
    o = oracle_mod.Oracle(*shape, fan_cfg(), precision=64)
    omega = 1.0 / o.params()["tau_shear"]
    kappa = 1.5 - 0.5 * omega
    o.step(fan_packet(shape, u0, tuple(vfan_len * n)))
    u1, rho = mean_velocity(o, shape)
    assert np.max(np.abs(rho - 1.0)) < 1e-12
    ub = np.array(u0, np.float64)
    sp = float(np.linalg.norm(ub))
    damp = 1.0 if sp <= 0.24 else 1.0 / (1.0 + ((sp - 0.24) / 0.04) ** 2)
    beta = 0.07 * damp
    vf = np.array(tuple(vfan_len * n), np.float32).astype(np.float64)  # the payload is float32
    target = 0.05 * float(np.linalg.norm(vf))
    n = vf / np.linalg.norm(vf)
    u_par = float(np.dot(ub, n))
    F = beta * (max(0.0, target - u_par) * n - (ub - u_par * n))
    want = ub + kappa * F
    assert np.max(np.abs(u1 - want)) < 1e-12


def test_fan_force_drives_axial_speed_to_target(oracle_mod):
    shape = (4, 4, 4)
    n = np.array([0.0, 0.0, 1.0])
    o = oracle_mod.Oracle(*shape, fan_cfg(), precision=64)
    kappa = 1.5 - 0.5 / o.params()["tau_shear"]
    target = 0.05 * 3.0
    pk = fan_packet(shape, (0.0, 0.0, 0.0), (0.0, 0.0, 3.0))
    o.step(pk)
    for _ in range(59):
        o.step(None)
    u, _ = mean_velocity(o, shape)
    want = target * (1.0 - (1.0 - 0.07 * kappa) ** 60)
    assert np.max(np.abs(u[..., 2] - want)) < 1e-12
    assert np.max(np.abs(u[..., :2])) < 1e-15


def test_fan_noise_bounds_and_mean(oracle_mod):
    shape = (24, 24, 24)
    o = oracle_mod.Oracle(*shape, fan_cfg(noise=True), precision=64)
    kappa = 1.5 - 0.5 / o.params()["tau_shear"]
    target = 0.05 * 2.0
    o.step(fan_packet(shape, (0.0, 0.0, 0.0), (2.0, 0.0, 0.0)))
    # the first step's force is uniform in space only up to the noise; streaming mixes
    # neighbouring cells, so read the post-collision momentum of each cell
    f = o.post_collision()
    jx = (CX[:, None] * f).sum(axis=0)
    ratio = jx / (kappa * 0.07 * target)
    assert ratio.min() >= 0.98 - 1e-12 and ratio.max() <= 1.02 + 1e-12
    assert ratio.max() - ratio.min() > 0.03                  # the noise is there
    sigma = 0.02 / math.sqrt(3.0) / math.sqrt(ratio.size)    # uniform xi: variance 1/3
    assert abs(ratio.mean() - 1.0) < 5.0 * sigma
