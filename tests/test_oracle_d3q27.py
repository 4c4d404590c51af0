"""Pins of the D3Q27 oracle (CPU, no GPU).

The reference solver TU cannot be compiled in this image (it includes <jni.h>),
so the oracle's restatement is pinned by: (1) the reference's own hydro/thermal
core compiled into oracle/_ref (tables, equilibrium, transport, beta);
(2) closed forms of the cumulant collision (every post-collision central moment
is rho times the Isserlis moment of a Gaussian whose covariance is the relaxed
second-order tensor); (3) Guo-source moment identities; (4) conservation and
invariance laws plus the reference debug harness thresholds
(native/debug/java/.../NativeLbmDebugHarness.java:115,157,196); (5) textbook
reductions: Taylor-Green decay gives nu = (tau - 1/2)/3 and pure diffusion of a
Fourier mode decays by the FTCS factor 1 - alpha (2 - 2 cos k).
"""
import itertools
import math

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

CX = np.array([(q // 9) - 1 for q in range(27)])
CY = np.array([((q // 3) % 3) - 1 for q in range(27)])
CZ = np.array([(q % 3) - 1 for q in range(27)])


def central_moments(f, u):
    """c[a,b,c] = sum_i f_i (cx-ux)^a (cy-uy)^b (cz-uz)^c, by definition, float64."""
    f = np.asarray(f, dtype=np.float64)
    dx, dy, dz = CX - u[0], CY - u[1], CZ - u[2]
    out = np.zeros((3, 3, 3))
    for a, b, c in itertools.product(range(3), repeat=3):
        out[a, b, c] = np.sum(f * dx ** a * dy ** b * dz ** c)
    return out


def isserlis(idx, cov):
    """E[prod x_i] of a zero-mean Gaussian with covariance `cov` (sum over pairings)."""
    if len(idx) == 0:
        return 1.0
    if len(idx) % 2:
        return 0.0
    first, rest = idx[0], idx[1:]
    total = 0.0
    for k in range(len(rest)):
        total += cov[first, rest[k]] * isserlis(rest[:k] + rest[k + 1:], cov)
    return total


def gaussian_state_moments(rho, c2):
    """Central moments of rho * Gaussian(cov = c2/rho) for every (a,b,c) in {0,1,2}^3."""
    cov = c2 / rho
    out = np.zeros((3, 3, 3))
    for a, b, c in itertools.product(range(3), repeat=3):
        idx = [0] * a + [1] * b + [2] * c
        out[a, b, c] = rho * isserlis(idx, cov)
    return out


def one_cell_cfg(tau=0.8, sgs=False, buoyancy=False):
    u_ref = 0.05
    nu = (tau - 0.5) / 3.0
    cfg = W.channel_config(u_ref, reynolds=u_ref / nu, sgs=sgs, buoyancy=buoyancy)
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_PERIODIC
        f.thermal_kind = A.THERMAL_PERIODIC
    return cfg


def near_eq(rng, rho=1.01, u=(0.03, -0.02, 0.01), amp=2e-3):
    w = np.array([np.prod([(1 / 6, 2 / 3, 1 / 6)[c + 1] for c in (CX[q], CY[q], CZ[q])]) for q in range(27)])
    cu = 3 * (CX * u[0] + CY * u[1] + CZ * u[2])
    uu = sum(x * x for x in u)
    feq = w * rho * (1 + cu + 0.5 * cu * cu - 1.5 * uu)
    return feq * (1 + amp * rng.standard_normal(27))


# ---------------------------------------------------------------- (1) reference core
def test_tables_and_feq_match_reference_core(oracle_mod):
    ref = oracle_mod.load_ref()
    if ref is None:
        pytest.skip("reference tree not present (oracle/_ref not built)")
    import ctypes
    cx, cy, cz, opp = ((ctypes.c_int * 27)() for _ in range(4))
    w = (ctypes.c_float * 27)()
    ref.ref_d3q27_tables(cx, cy, cz, opp, w)
    assert list(cx) == CX.tolist() and list(cy) == CY.tolist() and list(cz) == CZ.tolist()
    assert list(opp) == [26 - q for q in range(27)]
    ow = np.zeros(27, dtype=np.float32)
    oracle_mod.load().o32_oracle_weights(ow.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert np.array_equal(np.array(list(w), dtype=np.float32), ow)
    rng = np.random.default_rng(7)
    lib = oracle_mod.load()
    for _ in range(400):
        q = int(rng.integers(27))
        rho = float(np.float32(0.97 + 0.06 * rng.random()))
        u = [float(np.float32(0.3 * (rng.random() - 0.5))) for _ in range(3)]
        a = np.float32(ref.ref_feq_d3q27(q, rho, *u))
        b = np.float32(lib.o32_oracle_feq(q, rho, *u))
        assert a.tobytes() == b.tobytes(), (q, rho, u, a, b)


def test_physics_constants_match_reference_core(oracle_mod):
    ref = oracle_mod.load_ref()
    if ref is None:
        pytest.skip("reference tree not present")
    cfg = W.channel_config(0.05, 50.0)
    o = oracle_mod.Oracle(4, 4, 4, cfg, precision=32)
    p = o.params()
    # beta = g * beta_T * T_scale * dt^2/dx with the runtime constants (jni.cpp:161-167)
    beta_ref = ref.ref_boussinesq_beta(9.81, np.float32(1.0 / 300.0), 20.0, np.float32(1.0 / 20.0), 1.0)
    assert np.float32(p["beta"]) == np.float32(beta_ref)
    assert abs(p["beta"] - 0.001635) < 1e-6
    # nu = U L / Re = 0.05/50, tau = 0.5 + 3 nu
    assert abs(p["nu"] - 1e-3) < 1e-9
    assert abs(p["tau_shear"] - 0.503) < 1e-7 and p["tau_normal"] == p["tau_shear"]
    assert abs(p["sgs_c2"] - 0.025 ** 2) < 1e-9


# ---------------------------------------------------------------- (2) cumulant closed form
@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 2e-6)])
@pytest.mark.parametrize("tau", [0.53, 0.8, 0.95])
def test_collision_is_relaxed_gaussian_state(oracle_mod, precision, tol, tau):
    cfg = one_cell_cfg(tau=tau, sgs=False)
    o = oracle_mod.Oracle(1, 1, 1, cfg, precision=precision)
    rng = np.random.default_rng(11)
    for trial in range(20):
        f = near_eq(rng, rho=1.0 + 0.02 * (rng.random() - 0.5),
                    u=tuple(0.08 * (rng.random(3) - 0.5)), amp=5e-3)
        f = f.astype(np.float32).astype(np.float64) if precision == 32 else f
        rho = f.sum()
        u = np.array([np.dot(CX, f), np.dot(CY, f), np.dot(CZ, f)]) / rho
        post = o.collide_cell(f)
        cin = central_moments(f, u)
        om = min(1.0 / o.params()["tau_shear"], 1.95)
        c2 = np.array([[cin[2, 0, 0], cin[1, 1, 0], cin[1, 0, 1]],
                       [cin[1, 1, 0], cin[0, 2, 0], cin[0, 1, 1]],
                       [cin[1, 0, 1], cin[0, 1, 1], cin[0, 0, 2]]])
        relaxed = c2 * (1.0 - om)
        relaxed[np.diag_indices(3)] = np.diag(c2) + om * (rho / 3.0 - np.diag(c2))
        expect = gaussian_state_moments(rho, relaxed)
        got = central_moments(post, u)
        assert np.max(np.abs(got - expect)) < tol, (trial, np.max(np.abs(got - expect)))


def test_sgs_relaxation_rate_follows_stress_norm(oracle_mod):
    tau0 = 0.53
    cfg = one_cell_cfg(tau=tau0, sgs=True)
    ov = W.reference_overrides()
    ov.sgs_constant = 0.1
    o = oracle_mod.Oracle(1, 1, 1, cfg, overrides=ov, precision=64)
    nu0 = o.params()["nu"]
    tau0 = o.params()["tau_shear"]
    rng = np.random.default_rng(3)
    for _ in range(10):
        f = near_eq(rng, amp=2e-2)
        rho = f.sum()
        u = np.array([np.dot(CX, f), np.dot(CY, f), np.dot(CZ, f)]) / rho
        cin = central_moments(f, u)
        c2 = np.array([[cin[2, 0, 0], cin[1, 1, 0], cin[1, 0, 1]],
                       [cin[1, 1, 0], cin[0, 2, 0], cin[0, 1, 1]],
                       [cin[1, 0, 1], cin[0, 1, 1], cin[0, 0, 2]]])
        pi_neq = c2 - rho / 3.0 * np.eye(3)
        nut = min(o.params()["sgs_c2"] * np.linalg.norm(pi_neq) / (2 * rho * nu0), 5.0 * nu0)
        ts = min(0.5 + 3 * min(nu0 + nut, 0.15), 0.95)
        tn = min(tau0 + 0.3 * (ts - tau0), 0.95)
        post = central_moments(o.collide_cell(f), u)
        # off-diagonal shear moments relax with 1 - 1/tau_s, trace part with 1/tau_n
        assert abs(post[1, 1, 0] - (1 - 1 / ts) * cin[1, 1, 0]) < 1e-13
        assert abs(post[0, 1, 1] - (1 - 1 / ts) * cin[0, 1, 1]) < 1e-13
        assert abs(post[2, 0, 0] - (cin[2, 0, 0] + (rho / 3 - cin[2, 0, 0]) / tn)) < 1e-13
        assert ts > tau0  # the eddy viscosity raised the shear relaxation time


def test_guo_force_moments(oracle_mod):
    """Buoyancy F = rho*clamp(beta*T) along +y (reference default): the output keeps
    mass and its momentum is rho*(u + F/2rho) + (1 - omega/2) F."""
    tau = 0.7
    cfg = one_cell_cfg(tau=tau, buoyancy=True)
    ov = W.reference_overrides()
    ov.boussinesq_beta = 0.1
    for axis, sign in ((1, -1), (2, -1), (0, 1)):
        ov.gravity_axis, ov.gravity_sign = axis, sign
        o = oracle_mod.Oracle(1, 1, 1, cfg, overrides=ov, precision=64)
        rng = np.random.default_rng(5)
        f = near_eq(rng)
        T = 0.05
        rho = f.sum()
        j = np.array([np.dot(CX, f), np.dot(CY, f), np.dot(CZ, f)])
        F = np.zeros(3)
        prm = o.params()
        F[axis] = -sign * rho * min(max(prm["beta"] * T, -0.02), 0.02)
        post = o.collide_cell(f, T)
        assert abs(post.sum() - rho) < 1e-14
        jp = np.array([np.dot(CX, post), np.dot(CY, post), np.dot(CZ, post)])
        expect = j + 0.5 * F + (1 - 0.5 / prm["tau_shear"]) * F
        assert np.max(np.abs(jp - expect)) < 1e-14, (axis, jp, expect)


# ---------------------------------------------------------------- (3) invariants over steps
def periodic_cfg(tau=0.6, u_ref=0.01, sgs=False):
    cfg = one_cell_cfg(tau=tau, sgs=sgs)
    cfg.initial_velocity[0] = u_ref
    cfg.x_min.velocity[0] = 0.0
    cfg.reynolds_number = u_ref / ((tau - 0.5) / 3.0)
    return cfg


def test_equilibrium_at_rest_is_invariant(oracle_mod):
    """Harness check (NativeLbmDebugHarness.java:115): maxSpeed < 1e-6, |p| < 1e-6,
    mass drift < 1e-8 for rho=1, u=0 with solids."""
    n = 8
    cfg = periodic_cfg()
    solid = np.zeros((n, n, n), np.uint8)
    solid[3:5, 3:5, 3:5] = 1
    pk = W.make_packet(solid, u0=(0, 0, 0), perturb=0.0)
    o = oracle_mod.Oracle(n, n, n, cfg, precision=32)
    o.run(pk, 50)
    fl = o.flow().reshape(-1, 4)
    assert np.abs(fl[:, :3]).max() < 1e-6 and np.abs(fl[:, 3]).max() < 1e-6
    f = o.populations().astype(np.float64)
    fluid = solid.reshape(-1) == 0
    mass = f[:, fluid].sum()
    assert abs(mass - fluid.sum()) / fluid.sum() < 1e-8 * 10


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 2e-5)])
def test_mass_and_momentum_conservation_periodic(oracle_mod, precision, tol):
    n = 8
    cfg = periodic_cfg(tau=0.56, u_ref=0.03, sgs=True)
    solid = np.zeros((n, n, n), np.uint8)
    pk = W.make_packet(solid, u0=(0.03, 0.01, -0.02), perturb=0.01, seed_u=9)
    o = oracle_mod.Oracle(n, n, n, cfg, precision=precision)
    o.run(pk, 1)
    f0 = o.populations().astype(np.float64)
    m0, j0 = f0.sum(), np.array([(f0 * c[:, None]).sum() for c in (CX, CY, CZ)])
    for _ in range(30):
        o.step(None)
    f1 = o.populations().astype(np.float64)
    m1, j1 = f1.sum(), np.array([(f1 * c[:, None]).sum() for c in (CX, CY, CZ)])
    assert abs(m1 - m0) / m0 < tol
    assert np.max(np.abs(j1 - j0)) / m0 < tol


def test_closed_box_bounce_back_conserves_mass_and_is_no_slip(oracle_mod):
    """Harness thresholds: mass drift < 6e-3 (:157, FP64 holds it to 1e-12) and
    no-slip: solid cells report zero velocity (:196)."""
    n = 10
    cfg = periodic_cfg(tau=0.6, u_ref=0.02)
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
    solid = np.zeros((n, n, n), np.uint8)
    solid[4:6, 2:8, 3:7] = 1
    pk = W.make_packet(solid, u0=(0.02, 0.0, 0.01), perturb=0.005, seed_u=3)
    o = oracle_mod.Oracle(n, n, n, cfg, precision=64)
    o.run(pk, 1)
    fluid = solid.reshape(-1) == 0
    m0 = o.populations()[:, fluid].sum()
    for _ in range(40):
        o.step(None)
    m1 = o.populations()[:, fluid].sum()
    assert abs(m1 - m0) / m0 < 1e-12
    fl = o.flow().reshape(-1, 4)
    assert np.all(fl[~fluid] == 0.0)


def test_uniform_inflow_is_a_fixed_point(oracle_mod):
    """Velocity inlet + convective outlet + periodic y/z: uniform plug flow at the
    inlet velocity is an exact steady solution."""
    n = 8
    cfg = W.channel_config(0.05, 50.0, sgs=True)
    solid = np.zeros((n, n, n), np.uint8)
    pk = W.make_packet(solid, u0=(0.05, 0, 0), perturb=0.0)
    o = oracle_mod.Oracle(n, n, n, cfg, precision=64)
    o.run(pk, 30)
    fl = o.flow().reshape(-1, 4)
    assert np.max(np.abs(fl[:, 0] - 0.05)) < 1e-12
    assert np.max(np.abs(fl[:, 1:])) < 1e-12


def test_taylor_green_decay_gives_lattice_viscosity(oracle_mod):
    n = 16
    tau = 0.6
    nu = (tau - 0.5) / 3
    U = 0.01
    cfg = periodic_cfg(tau=tau, u_ref=U)
    k = 2 * math.pi / n
    x, y, _ = np.meshgrid(np.arange(n), np.arange(n), np.arange(4), indexing="ij")
    solid = np.zeros((n, n, 4), np.uint8)
    pk = W.make_packet(solid, u0=(0, 0, 0), perturb=0.0).reshape(-1, 11)
    pk[:, 5] = (U * np.sin(k * x) * np.cos(k * y)).reshape(-1)
    pk[:, 6] = (-U * np.cos(k * x) * np.sin(k * y)).reshape(-1)
    pk[:, 8] = (-3 * U * U / 4 * (np.cos(2 * k * x) + np.cos(2 * k * y))).reshape(-1)
    o = oracle_mod.Oracle(n, n, 4, cfg, precision=64)
    o.run(pk.reshape(-1), 1)
    e0 = np.sum(o.flow().reshape(-1, 4)[:, :3] ** 2)
    steps = 200
    for _ in range(steps):
        o.step(None)
    e1 = np.sum(o.flow().reshape(-1, 4)[:, :3] ** 2)
    nu_meas = -math.log(e1 / e0) / (4 * k * k * steps)
    assert abs(nu_meas / nu - 1) < 0.03, nu_meas


def test_pure_diffusion_matches_ftcs_factor(oracle_mod):
    n = 16
    cfg = periodic_cfg(tau=0.6, u_ref=0.01)
    cfg.flags |= A.FLAG_DISABLE_BUOYANCY
    solid = np.zeros((n, 2, 2), np.uint8)
    x = np.arange(n)[:, None, None] * np.ones((1, 2, 2))
    t0 = (0.5 * np.sin(2 * math.pi * x / n)).astype(np.float32)
    pk = W.make_packet(solid, u0=(0, 0, 0), perturb=0.0, temperature=t0)
    o = oracle_mod.Oracle(n, 2, 2, cfg, precision=64)
    o.set_temperature(t0)
    steps = 40
    o.run(pk, steps)
    alpha = 0.035
    lam = 1 - alpha * (2 - 2 * math.cos(2 * math.pi / n))
    t = o.temperature_r().reshape(n, 2, 2)
    expect = t0.astype(np.float64) * lam ** steps
    assert np.max(np.abs(t - expect)) < 1e-7


def test_output_semantics_clamp_and_snap(oracle_mod):
    """write_output (jni.cpp:2535-2581): rho clamped to [0.97,1.03] before u = m/rho,
    |u| < 1e-7 snaps to 0, p = clamp(rho-1, +-0.03), solids 0."""
    cfg = periodic_cfg()
    solid = np.zeros((2, 1, 1), np.uint8)
    solid[1] = 1
    pk = W.make_packet(solid, u0=(0, 0, 0), perturb=0.0)
    o = oracle_mod.Oracle(2, 1, 1, cfg, precision=64)
    o.run(pk, 1)
    f = np.zeros((27, 2))
    f[13, 0] = 1.2          # rest population: rho = 1.2 + 0.02
    f[22, 0] = 0.02         # q=22: c = (+1, 0, 0)
    f[4, 0] = 5e-8          # q=4: c = (-1, 0, 0) -> net m_x = 0.02 - 5e-8
    o.set_populations(f)
    out = o.flow().reshape(-1, 4)
    assert out[0, 0] == np.float32((0.02 - 5e-8) / 1.03)
    assert out[0, 1] == 0.0 and out[0, 2] == 0.0
    assert out[0, 3] == np.float32(0.03)
    assert np.all(out[1] == 0)


def test_fp32_tracks_fp64_on_channel(oracle_mod):
    cfg, pk, solid = W.config1(12)
    a = oracle_mod.Oracle(12, 12, 12, cfg, precision=32)
    b = oracle_mod.Oracle(12, 12, 12, cfg, precision=64)
    a.run(pk, 40)
    b.run(pk, 40)
    fa, fb = a.flow().reshape(-1, 4), b.flow().reshape(-1, 4)
    assert np.max(np.abs(fa[:, :3] - fb[:, :3])) < 1e-5
