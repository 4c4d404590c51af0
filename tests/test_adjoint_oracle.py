"""Pins of the adjoint's checkers, CPU only.

* The adjoint model's FP64 forward (run through oracle/libadjoint_dual.so with a
  zero tangent) equals the FP64 oracle of the reference step on the same brick
  (velocity inlet x-, convective outflow x+, periodic y/z, adiabatic x thermal
  faces, Boussinesq, heat source, no SGS): the loss built from the oracle's
  velocities matches to 1e-9 relative.
* The FP64 dual-number directional derivative equals central finite differences
  of the same FP64 loss (1e-6 relative): the dual propagation is right.
* The hand-written reverse mode of the collision (collide_low_adjoint, the
  product's collide adjoint) equals the transposed Jacobian assembled from
  forward-mode dual numbers on the same map, in FP64, across the density floor,
  the buoyancy clamp and the speed cap.
"""
import ctypes

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

N, M, STEPS = 8, 4, 12
BETA, ALPHA = 0.1, 0.035
U_IN = 0.04


def upsample(heat, n, m):
    """Trilinear sample of the m^3 field at cell centres ((x + 0.5) m / n - 0.5,
    clamped), clamped to +-0.006: the model's heat source."""
    p = np.clip((np.arange(n) + 0.5) * m / n - 0.5, 0, m - 1)
    lo = np.minimum(np.floor(p).astype(int), max(0, m - 2))
    t = p - lo
    h = heat.reshape(m, m, m).astype(np.float64)
    out = np.zeros((n, n, n))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wx = np.where(dx, t, 1 - t).astype(np.float32)
                wy = np.where(dy, t, 1 - t).astype(np.float32)
                wz = np.where(dz, t, 1 - t).astype(np.float32)
                w = (wx[:, None, None] * wy[None, :, None]).astype(np.float32) * wz[None, None, :]
                ix = np.minimum(lo + dx, m - 1)
                out += w * h[np.ix_(ix, np.minimum(lo + dy, m - 1), np.minimum(lo + dz, m - 1))]
    return np.clip(out, -0.006, 0.006)


def case(seed=0):
    rng = np.random.default_rng(seed)
    solid = np.zeros((N, N, N), np.uint8)
    solid[4:6, 2:4, 3:5] = 1
    probes = rng.choice(np.flatnonzero(solid.reshape(-1) == 0), 20, replace=False).astype(np.int32)
    heat = rng.uniform(-0.008, 0.008, M ** 3)
    return solid, probes, heat


def test_model_forward_equals_fp64_oracle(oracle_mod):
    solid, probes, heat = case(1)
    cfg = W.channel_config(U_IN, 50.0, sgs=False, buoyancy=True)
    cfg.x_min.thermal_kind = A.THERMAL_ADIABATIC
    ov = W.reference_overrides(A.KERNEL_EXACT)
    ov.boussinesq_beta = BETA
    ov.thermal_diffusivity = ALPHA
    ov.gravity_axis = 2
    ov.gravity_sign = -1
    o = oracle_mod.Oracle(N, N, N, cfg, overrides=ov, precision=64)
    nu = o.params()["nu"]
    pk = W.make_packet(solid, u0=(U_IN, 0.0, 0.0), perturb=0.0, seed_u=1).reshape(-1, 11)
    pk[:, 9] = upsample(heat, N, M).reshape(-1)
    o.run(pk.reshape(-1), STEPS)
    u = o.velocity()
    ref_loss = float(np.sum(u[probes] ** 2))
    # the oracle's face velocity is a float32 config value
    inflow = np.tile(np.array([U_IN, 0.0, 0.0], np.float32).astype(np.float64), N * N)
    loss, _ = oracle_mod.dual_directional(N, M, STEPS, nu, BETA, ALPHA, (U_IN, 0.0, 0.0), solid, probes,
                                          np.zeros(3 * len(probes)), inflow, heat, np.zeros_like(inflow),
                                          np.zeros_like(heat))
    assert abs(loss - ref_loss) <= 1e-9 * ref_loss, (loss, ref_loss)


@pytest.mark.parametrize("heat_scale", [1.0, 0.0])
def test_dual_derivative_equals_central_differences(oracle_mod, heat_scale):
    """heat_scale 0: the temperature stays exactly 0, where the buoyancy source
    term vanishes but its derivative does not."""
    solid, probes, heat = case(2)
    heat = heat * heat_scale
    rng = np.random.default_rng(5)
    inflow = np.tile([U_IN, 0.0, 0.0], N * N) + rng.uniform(-0.005, 0.005, 3 * N * N)
    target = rng.uniform(-0.02, 0.06, 3 * len(probes))
    di = rng.standard_normal(3 * N * N)
    dh = rng.standard_normal(M ** 3) * 0.1
    args = (N, M, STEPS, 0.05, BETA, ALPHA, (U_IN, 0.0, 0.0), solid, probes, target)
    _, dl = oracle_mod.dual_directional(*args, inflow, heat, di, dh)
    eps = 1e-5
    lp, _ = oracle_mod.dual_directional(*args, inflow + eps * di, heat + eps * dh, 0 * di, 0 * dh)
    lm, _ = oracle_mod.dual_directional(*args, inflow - eps * di, heat - eps * dh, 0 * di, 0 * dh)
    fd = (lp - lm) / (2 * eps)
    assert abs(dl - fd) <= 1e-6 * abs(fd), (dl, fd)


def test_collide_reverse_mode_equals_dual_numbers():
    import os
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                                   "libadjoint_dual.so"))
    f = lib.ad_collide_vjp
    D = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [ctypes.c_double, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float, D, ctypes.c_double,
                  D, D, D, D, D, D]
    rng = np.random.default_rng(0)
    kmax = 0.57735026919 * 0.6
    branches = set()
    for trial in range(400):
        rho = 1.0 + 0.05 * rng.standard_normal()
        speed = [0.02, 0.1, 0.4][trial % 3]  # 0.4 is above the speed cap
        u = rng.standard_normal(3)
        u = u / np.linalg.norm(u) * speed
        m = np.array([rho, *(rho * u), rho * (1 / 3 + u[0] ** 2), rho * (1 / 3 + u[1] ** 2),
                      rho * (1 / 3 + u[2] ** 2), rho * u[0] * u[1], rho * u[0] * u[2], rho * u[1] * u[2]])
        m = m + 1e-3 * rng.standard_normal(10)
        T = [0.0, 0.05, 0.5, -0.5][trial % 4]  # beta * 0.5 lies outside the clamp
        beta = [0.0, 0.1][trial % 2]
        fb, ub = rng.standard_normal(27), rng.standard_normal(3)
        out = [np.zeros(10), np.zeros(1), np.zeros(10), np.zeros(1)]
        p = lambda a: a.ctypes.data_as(D)  # noqa: E731
        f(1e-3 * (1 + trial % 5), beta, trial % 3, 1 if trial % 2 else -1, kmax, p(m), T, p(fb), p(ub),
          p(out[0]), p(out[1]), p(out[2]), p(out[3]))
        rev, dual = np.concatenate(out[:2]), np.concatenate(out[2:])
        assert np.max(np.abs(rev - dual)) <= 1e-11 * max(1.0, np.max(np.abs(dual))), trial
        branches.add((speed > kmax, abs(beta * T) > 0.02))
    assert branches == {(False, False), (True, False), (False, True), (True, True)}
