"""Pins of the D3Q19 mesoscale oracle (oracle/meso_oracle_impl.h), CPU only.

The reference holds no golden vectors for the mesoscale step, and its source
(native/src/aero_lbm_mesoscale.cpp) includes <jni.h>, absent here, so it is not
compiled. The oracle is pinned by:

* the transport coefficients against the reference's own hydro/thermal core
  compiled from its sources (oracle/_ref);
* an exact fixed point (quiescent isothermal atmosphere over flat ground);
* closed forms of one step at unit relaxation (tau = 1 for f and g, where the
  two-relaxation-time collision returns the equilibrium of the targets), which
  fix the streaming direction, the opposite table, the momentum relaxation
  toward the background wind (meso.cpp:483-492), the floor bounce-back and the
  wall-temperature rule (653-690), the solid-cell rule (633-641) and the
  edge/top sponge blend (671-688).
"""
import numpy as np
import pytest

# D3Q19 directions in the reference's order (meso.cpp:146-176)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
T0 = float(np.float32(288.15))  # the forcing is FP32


def opp_table():
    out = np.zeros(19, dtype=int)
    for q in range(19):
        out[q] = [k for k in range(19) if CX[k] == -CX[q] and CY[k] == -CY[q] and CZ[k] == -CZ[q]][0]
    return out


OPP = opp_table()


def hfeq(rho, ux, uy, uz):
    """D3Q19 second-order equilibrium, shape (..., 19)."""
    cu = 3.0 * (CX * ux[..., None] + CY * uy[..., None] + CZ * uz[..., None])
    uu = (ux * ux + uy * uy + uz * uz)[..., None]
    return W * rho[..., None] * (1.0 + cu + 0.5 * cu * cu - 1.5 * uu)


def tfeq(t, ux, uy, uz):
    cu = 3.0 * (CX * ux[..., None] + CY * uy[..., None] + CZ * uz[..., None])
    return W * t[..., None] * (1.0 + cu)


def forcing(nx, ny, nz, solid=None, wall_t=T0):
    F = np.zeros((nx, ny, nz, 25), np.float32)
    F[..., 2] = T0
    F[..., 3] = T0
    F[..., 4] = T0
    if solid is not None:
        F[..., 24] = solid
        F[solid > 0, 4] = wall_t
    return F


# dx = dt = 3, nu = 0.5: lattice viscosity 1/6 and tau_f = 1 exactly (also for the
# odd relaxation time, 0.5 + 0.25 / (tau - 0.5) = 1)
DX, DT, NU, PR = 3.0, 3.0, 0.5, 0.75


def make(oracle_mod, dims, dx=DX, dt=DT, nu=NU, pr=PR, precision=64, tau_min=0.58):
    return oracle_mod.MesoOracle(*dims, dx=dx, dt=dt, nu=nu, pr=pr, prt=0.85, tau_min=tau_min, precision=precision)


def test_transport_matches_reference_core(oracle_mod):
    ref = oracle_mod.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes
    for dx, dt, nu, pr in ((64.0, 5.0, 1.5e-5, 0.71), (1.0, 1.0, 1 / 6, 4 / 3), (32.0, 2.5, 0.06, 0.9), (8.0, 0.1, 2.0, 0.71)):
        out = (ctypes.c_float * 5)()
        ref.ref_derive_transport(8, 8, 8, dx, dt, nu, pr, 0.85, out)
        o = make(oracle_mod, (2, 2, 2), dx=dx, dt=dt, nu=nu, pr=pr, precision=32)
        vs, th, tt = o.transport()
        assert np.float32(vs) == np.float32(out[0])
        assert np.float32(th) == np.float32(out[3])
        assert np.float32(tt) == np.float32(out[4])


@pytest.mark.parametrize("precision", [32, 64])
def test_quiescent_isothermal_fixed_point(oracle_mod, precision):
    dims = (10, 6, 9)
    o = make(oracle_mod, dims, dx=64.0, dt=5.0, nu=1.5e-5, pr=0.71, precision=precision)
    F = forcing(*dims)
    F[..., 5] = 0.1  # roughness: drag acts on zero velocity only
    for _ in range(12):
        st = o.step(F).reshape(-1, 6)
    tol = 1e-9 if precision == 64 else 2e-4
    assert np.max(np.abs(st[:, :3] - T0)) <= tol
    assert np.max(np.abs(st[:, 3:])) <= (1e-12 if precision == 64 else 1e-6)
    f, g = o.distributions()
    assert np.max(np.abs(f - W)) <= (1e-14 if precision == 64 else 1e-6)


def unit_relaxation_step(oracle_mod, dims, solid=None, seed=0):
    """Seed + one step, then a random (rho, u) state with a non-equilibrium part in f
    and g at the thermal equilibrium of the relaxed velocity, then one more step at
    tau_f = 1. Returns (input f, input g, output f, output g)."""
    o = make(oracle_mod, dims)
    F = forcing(*dims, solid=solid)
    o.step(F)
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    rho = rng.uniform(0.98, 1.02, n)
    u = rng.uniform(-0.03, 0.03, (3, n))
    f0 = hfeq(rho, *u) + rng.uniform(-1e-3, 1e-3, (n, 19))
    _, rx, ry, rz = relaxed_targets(f0)
    g0 = tfeq(np.full(n, T0), rx, ry, rz)
    o.set_distributions(f0, g0)
    o.step(F)
    f1, g1 = o.distributions()
    sh = lambda a: a.reshape(*dims, 19)
    return sh(f0), sh(g0), sh(f1), sh(g1)


def relaxed_targets(f0, dt=DT):
    """Post-collision state at unit relaxation: equilibrium of the cell density and of
    the velocity relaxed toward the (zero) background wind by dt/60 (meso.cpp:483-492)."""
    rho = f0.sum(-1)
    ux = (f0 * CX).sum(-1) / rho
    uy = (f0 * CY).sum(-1) / rho
    uz = (f0 * CZ).sum(-1) / rho
    wr = dt / 60.0
    return rho, (1 - wr) * ux, (1 - wr) * uy, (1 - wr) * uz


def test_unit_relaxation_stream_closed_form(oracle_mod):
    dims = (12, 8, 11)
    nx, ny, nz = dims
    f0, g0, f1, g1 = unit_relaxation_step(oracle_mod, dims)
    rho, ux, uy, uz = relaxed_targets(f0)
    post = hfeq(rho, ux, uy, uz)
    gpost = g0  # already the thermal equilibrium of the targets
    checked = 0
    for x in range(2, nx - 2):
        for y in range(0, ny - 2):
            for z in range(2, nz - 2):
                for q in range(19):
                    sx, sy, sz = x - CX[q], y - CY[q], z - CZ[q]
                    if sy < 0:  # floor: halfway bounce-back of the cell's own post-collision value
                        ef, eg = post[x, y, z, OPP[q]], W[q] * T0
                    else:
                        ef, eg = post[sx, sy, sz, q], gpost[sx, sy, sz, q]
                    assert abs(f1[x, y, z, q] - ef) <= 1e-13, (x, y, z, q)
                    assert abs(g1[x, y, z, q] - eg) <= 1e-9, (x, y, z, q)
                    checked += 1
    assert checked > 6000


def test_edge_and_top_sponge(oracle_mod):
    dims = (10, 7, 10)
    nx, ny, nz = dims
    f0, g0, f1, g1 = unit_relaxation_step(oracle_mod, dims, seed=3)
    rho, ux, uy, uz = relaxed_targets(f0)
    post = hfeq(rho, ux, uy, uz)
    # distance 0 from an edge (and off the floor): background equilibrium, zero wind
    for (x, y, z) in ((0, 3, 4), (nx - 1, 2, 5), (4, ny - 1, 4), (3, 2, 0)):
        assert np.max(np.abs(f1[x, y, z] - W)) <= 1e-15
        assert np.max(np.abs(g1[x, y, z] - W * T0)) <= 1e-12
    # distance 1: half blend of the streamed value and the background
    x, y, z = 1, 3, 5
    for q in range(19):
        streamed = post[x - CX[q], y - CY[q], z - CZ[q], q]
        assert abs(f1[x, y, z, q] - (0.5 * streamed + 0.5 * W[q])) <= 1e-13


def test_solid_cells_and_wall_links(oracle_mod):
    dims = (12, 8, 12)
    solid = np.zeros(dims, np.float32)
    solid[5:8, 0:3, 5:8] = 1.0
    f0, g0, f1, g1 = unit_relaxation_step(oracle_mod, dims, solid=solid, seed=7)
    rho, ux, uy, uz = relaxed_targets(f0)
    post = hfeq(rho, ux, uy, uz)
    s = solid > 0
    assert np.max(np.abs(f1[s] - W)) <= 1e-15
    assert np.max(np.abs(g1[s] - W * T0)) <= 1e-12
    # fluid cell beside the block: links whose source is solid bounce back
    x, y, z = 4, 1, 6
    assert not s[x, y, z]
    n_bounce = 0
    for q in range(19):
        sx, sy, sz = x - CX[q], y - CY[q], z - CZ[q]
        if sy >= 0 and s[sx, sy, sz]:
            assert abs(f1[x, y, z, q] - post[x, y, z, OPP[q]]) <= 1e-13
            assert abs(g1[x, y, z, q] - W[q] * T0) <= 1e-12
            n_bounce += 1
    assert n_bounce >= 3


def test_fp32_tracks_fp64_on_terrain(oracle_mod):
    import workloads as Wl
    dims = (24, 12, 24)
    F = Wl.meso_forcing(*dims, seed=3, max_height=6.0)
    o32 = oracle_mod.MesoOracle(*dims, dx=64.0, dt=5.0, nu=5e-4 * 64 * 64 / 5.0, pr=0.71, prt=0.85, precision=32)
    o64 = oracle_mod.MesoOracle(*dims, dx=64.0, dt=5.0, nu=5e-4 * 64 * 64 / 5.0, pr=0.71, prt=0.85, precision=64)
    for _ in range(40):
        a = o32.step(F).reshape(-1, 6)
        b = o64.step(F).reshape(-1, 6)
    assert np.max(np.abs(a[:, :3] - b[:, :3])) < 2e-3      # kelvin
    assert np.max(np.abs(a[:, 3:] - b[:, 3:])) < 1e-4       # m/s (wind ~0.5 m/s)
    assert np.max(np.abs(b[:, 3:])) > 0.05                  # the wind field is not trivial


@pytest.mark.parametrize("tau", [0.65, 1.3])
def test_trt_shear_wave_decays_with_even_viscosity(oracle_mod, tau):
    """Two-relaxation-time collision away from unit relaxation (meso.cpp:548-738):
    a shear wave u_x = A sin(k z) in a quiescent atmosphere decays as exp(-nu k^2 t)
    with nu = (tau_even - 1/2)/3, whatever the odd rate (magic parameter 1/4 gives
    tau_odd = 0.5 + 0.25/(tau_even - 0.5), which would give a very different decay).
    The background wind of the forcing equals the initial wave (so the momentum
    relaxation toward it, dt/60 per step, acts only on the decay itself, a 1e-7
    effect at dt = 1 ms), the seeding puts the equilibrium of that wind in every
    cell, and the amplitude is read at the domain centre after steps 10 and 30:
    populations move at most one cell per step, so a centre more than 30 cells from
    every face and the floor sees no boundary."""
    nx, ny, nz, lam = 64, 64, 128, 64
    dx, dt = 1.0, 1.0e-3
    nu_l = (tau - 0.5) / 3.0
    o = oracle_mod.MesoOracle(nx, ny, nz, dx=dx, dt=dt, nu=nu_l * dx * dx / dt, pr=0.71, prt=0.85, tau_min=0.51,
                              precision=64)
    vs = dx / dt
    k = 2.0 * np.pi / lam
    z = np.arange(nz)
    F = np.zeros((nx, ny, nz, 25), np.float32)
    F[..., 2] = T0
    F[..., 3] = T0
    F[..., 4] = T0
    F[..., 6] = (0.01 * np.sin(k * (z - 32)) * vs)[None, None, :]
    zs = np.arange(32, 32 + lam)
    basis = np.sin(k * (zs - 32))
    amp = {}
    for s in range(1, 31):
        out = o.step(F).reshape(nx, ny, nz, 6)
        if s in (10, 30):
            ux = out[nx // 2, ny // 2, zs, 3].astype(np.float64) / vs
            amp[s] = float(np.dot(ux, basis) / np.dot(basis, basis))
    nu_eff = -np.log(amp[30] / amp[10]) / (20.0 * k * k)
    assert nu_eff == pytest.approx(nu_l, rel=5e-3)
    tau_odd = 0.5 + 0.25 / (tau - 0.5)
    assert abs(nu_eff - (tau_odd - 0.5) / 3.0) > 0.1 * nu_l
