"""Oracle pins for the heated-cavity presets: their resolved parameters and the BFECC
thermal advection (ref: native/src/aero_lbm_jni.cpp:823-828, 933-993 parameters,
2050-2139 BFECC branch of update_temperature_field).

The checks come from the mathematics, not from the reference's formulas retyped:
the resolved (nu, alpha, beta) must give back the preset's Rayleigh, Prandtl and Mach
numbers; at rest BFECC is the identity, leaving explicit FTCS diffusion with the
face rules; on one periodic Fourier mode with uniform velocity the update is the
published BFECC amplification factor for linear interpolation (Dupont and Liu 2003,
Kim et al. 2005); on random data the limiter keeps the result inside the bounds of
the old field around each cell.
"""
import math

import numpy as np
import pytest

import aero_lbm as A
import workloads as W


def reference_ov(alpha=None):
    ov = W.reference_overrides(A.KERNEL_EXACT)
    if alpha is not None:
        ov.thermal_diffusivity = alpha
    return ov


def periodic_bfecc_cfg():
    cfg = W.heated_cavity_config(True)
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_PERIODIC
        f.thermal_kind = A.THERMAL_PERIODIC
    return cfg


def first_temperature_step(oracle_mod, cfg, ov, t0, u0=(0.0, 0.0, 0.0), perturb=0.0, solid=None, seed_u=3):
    """One oracle step from temperature t0 with reference velocity u0 (+- perturb):
    the first temperature update advects with that velocity."""
    nx, ny, nz = t0.shape
    solid = np.zeros(t0.shape, np.uint8) if solid is None else solid
    packet = W.make_packet(solid, u0=u0, perturb=perturb, seed_u=seed_u, temperature=t0)
    o = oracle_mod.Oracle(nx, ny, nz, cfg, overrides=ov, precision=64)
    o.set_temperature(t0.reshape(-1).astype(np.float32))
    o.step(packet)
    u = packet.reshape(-1, 11)[:, 5:8].astype(np.float64)
    return o, o.temperature_r().reshape(t0.shape), u.reshape(nx, ny, nz, 3)


@pytest.mark.parametrize("length", [1.0, 24.0])
@pytest.mark.parametrize("three_d", [True, False])
def test_heated_cavity_parameters_give_back_ra_pr_ma(oracle_mod, length, three_d):
    cfg = W.heated_cavity_config(three_d, reference_length=length)
    o = oracle_mod.Oracle(6, 6, 6, cfg, overrides=reference_ov(), precision=64)
    p = o.params()
    nu, alpha, beta = p["nu"], p["alpha"], p["beta"]
    dT = 0.5 - (-0.5)
    # Pr = nu / alpha, Ra = g beta dT L^3 / (nu alpha) (beta carries g), U = sqrt(g beta dT L)
    assert nu / alpha == pytest.approx(0.71, rel=1e-6)
    assert beta * dT * length ** 3 / (nu * alpha) == pytest.approx(1.0e5, rel=1e-5)
    assert math.sqrt(beta * dT * length) / math.sqrt(1.0 / 3.0) == pytest.approx(0.15, rel=1e-6)
    assert p["tau_shear"] == pytest.approx(0.5 + 3.0 * nu, rel=1e-6)


def test_bfecc_at_rest_is_ftcs_diffusion_with_face_rules(oracle_mod):
    rng = np.random.default_rng(11)
    shape = (10, 8, 6)
    t0 = rng.uniform(-0.4, 0.4, shape).astype(np.float32)
    solid = (rng.uniform(size=shape) < 0.12).astype(np.uint8)
    cfg = W.heated_cavity_config(True)
    o, t1, _ = first_temperature_step(oracle_mod, cfg, reference_ov(), t0, solid=solid)
    alpha = o.params()["alpha"]
    T = t0.astype(np.float64)
    sol = solid.astype(bool)
    # ghost values: x- wall +0.5, x+ wall -0.5 (Dirichlet), y / z walls adiabatic (the cell)
    lap = -6.0 * T
    for axis in range(3):
        for step in (1, -1):
            nb = np.roll(T, -step, axis=axis)
            nb_solid = np.roll(sol, -step, axis=axis)
            edge = [slice(None)] * 3
            edge[axis] = -1 if step == 1 else 0
            edge = tuple(edge)
            nb = np.where(nb_solid, T, nb)
            if axis == 0:
                nb[edge] = -0.5 if step == 1 else 0.5
            else:
                nb[edge] = T[edge]
            lap += nb
    want = np.clip(T + alpha * lap, -1.0, 1.0)
    want[sol] = 0.0
    assert np.max(np.abs(t1 - want)) < 1e-12


@pytest.mark.parametrize("axis,c,mode", [(0, 0.5, 1), (1, -0.3, 2), (2, 0.7, 1), (0, 1.0, 3)])
def test_bfecc_single_mode_matches_amplification_factor(oracle_mod, axis, c, mode):
    shape = [8, 8, 8]
    shape[axis] = 32
    n = shape[axis]
    amp = 0.4
    k = 2.0 * math.pi * mode / n
    coord = np.arange(n, dtype=np.float64)
    prof = (amp * np.cos(k * coord)).astype(np.float32)
    bshape = [1, 1, 1]
    bshape[axis] = n
    t0 = np.broadcast_to(prof.reshape(bshape), shape).astype(np.float32).copy()
    u0 = [0.0, 0.0, 0.0]
    u0[axis] = c
    _, t1, u = first_temperature_step(oracle_mod, periodic_bfecc_cfg(), reference_ov(alpha=0.0), t0, u0=tuple(u0))
    cc = float(u[0, 0, 0, axis])  # the float32 velocity the oracle advects with
    # one step of linear-interpolation semi-Lagrangian advection multiplies mode k by
    # a = 1 - |c| + |c| e^{-ik sgn c} (0 <= |c| <= 1); BFECC gives G = a (3 - |a|^2) / 2
    a = 1.0 - abs(cc) + abs(cc) * np.exp(-1j * k * math.copysign(1.0, cc))
    G = a * (3.0 - abs(a) ** 2) / 2.0
    want_prof = np.real(G * np.fft.fft(prof.astype(np.float64))[mode] * 2.0 / n * np.exp(1j * k * coord))
    got_prof = np.moveaxis(t1, axis, 0)[:, 0, 0]
    assert np.max(np.abs(got_prof - want_prof)) < 1e-6 * amp
    assert np.max(np.abs(t1 - np.broadcast_to(got_prof.reshape(bshape), shape))) < 1e-12


def test_bfecc_limiter_keeps_local_bounds_and_differs_from_plain_advection(oracle_mod):
    rng = np.random.default_rng(5)
    shape = (12, 10, 8)
    t0 = rng.uniform(-0.5, 0.5, shape).astype(np.float32)
    _, t1, _ = first_temperature_step(oracle_mod, periodic_bfecc_cfg(), reference_ov(alpha=0.0), t0,
                                      perturb=0.8, seed_u=9)
    T = t0.astype(np.float64)
    lo, hi = T.copy(), T.copy()
    for axis in range(3):
        for step in (1, -1):
            nb = np.roll(T, -step, axis=axis)
            lo, hi = np.minimum(lo, nb), np.maximum(hi, nb)
    assert np.all(t1 >= lo - 1e-15) and np.all(t1 <= hi + 1e-15)
    at_bound = np.count_nonzero((t1 == lo) | (t1 == hi))
    assert at_bound > 0.02 * T.size  # the limiter is engaged on rough data
    plain = periodic_bfecc_cfg()
    plain.preset = A.PRESET_NONE
    _, t1_sl, _ = first_temperature_step(oracle_mod, plain, reference_ov(alpha=0.0), t0, perturb=0.8, seed_u=9)
    assert np.max(np.abs(t1 - t1_sl)) > 1e-2
