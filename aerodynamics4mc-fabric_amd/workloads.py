"""Synthetic workloads for the local-airflow solver (seeded, deterministic).

Every generator uses SplitMix64 (the mixing constants the reference uses for
its fan noise hash, native/src/aero_lbm_jni.cpp:1318-1327) so inputs are
reproducible across machines. Payload packets follow the reference's 11-channel
per-cell layout (native/src/aero_lbm_jni.cpp:64-74): [0] obstacle (>0.5 solid),
[1] fan mask, [2-4] fan velocity, [5-7] reference velocity (lattice),
[8] reference pressure, [9] thermal source, [10] reference temperature. Cell
order is x-major, z fastest: cell = (x*ny + y)*nz + z.
"""
from __future__ import annotations

import numpy as np

import aero_lbm as A  # the package directory is on sys.path (see __graft_entry__.PKG_DIR)

_M64 = (1 << 64) - 1


class SplitMix64:
    """Sequential SplitMix64 generator."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform in [0, 1) with 53-bit resolution."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] inclusive."""
        return lo + int(self.uniform() * (hi - lo + 1))


def splitmix_array(seed: int, n: int) -> np.ndarray:
    """n uniforms in [0,1) from SplitMix64 stream `seed` (vectorised, counter-based)."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def solid_boxes(nx: int, ny: int, nz: int, fill: float, seed: int, edge=(2, 8),
                keep_clear_x: int = 2) -> np.ndarray:
    """Random axis-aligned solid boxes (edges uniform in `edge`) placed until the
    solid fraction reaches `fill`. Planes x < keep_clear_x and x >= nx-keep_clear_x
    stay fluid so inflow/outflow faces are open. Returns uint8 [nx,ny,nz]."""
    rng = SplitMix64(seed)
    solid = np.zeros((nx, ny, nz), dtype=np.uint8)
    target = fill * nx * ny * nz
    guard = 0
    while solid.sum() < target and guard < 100000:
        guard += 1
        ex, ey, ez = (rng.randint(edge[0], edge[1]) for _ in range(3))
        x0 = rng.randint(keep_clear_x, max(keep_clear_x, nx - keep_clear_x - ex))
        y0 = rng.randint(0, max(0, ny - ey))
        z0 = rng.randint(0, max(0, nz - ez))
        solid[x0:min(x0 + ex, nx - keep_clear_x), y0:y0 + ey, z0:z0 + ez] = 1
    return solid


def make_packet(solid: np.ndarray, u0=(0.05, 0.0, 0.0), perturb: float = 1e-3, seed_u: int = 1,
                temperature: np.ndarray | None = None, source: np.ndarray | None = None,
                channels: int = 11) -> np.ndarray:
    """Payload packet [cells*channels] float32 with reference velocity u0 + U(-perturb, perturb)."""
    nx, ny, nz = solid.shape
    cells = nx * ny * nz
    p = np.zeros((cells, channels), dtype=np.float32)
    p[:, 0] = solid.reshape(-1).astype(np.float32)
    r = splitmix_array(seed_u, 3 * cells).reshape(cells, 3)
    for k in range(3):
        p[:, 5 + k] = (u0[k] + perturb * (2.0 * r[:, k] - 1.0)).astype(np.float32)
    if temperature is not None and channels > 10:
        p[:, 10] = temperature.reshape(-1).astype(np.float32)
    if source is not None and channels > 9:
        p[:, 9] = source.reshape(-1).astype(np.float32)
    return p.reshape(-1)


def hot_blob(nx: int, ny: int, nz: int, amplitude: float = 0.02, sigma: float = 8.0, seed: int = 4,
             axis: int = 2) -> np.ndarray:
    """Gaussian temperature blob centred on the floor (coordinate 0 of `axis`), with a
    seeded jitter of the horizontal centre by up to +-2 cells."""
    rng = SplitMix64(seed)
    c = [nx / 2.0, ny / 2.0, nz / 2.0]
    for k in range(3):
        if k != axis:
            c[k] += 4.0 * rng.uniform() - 2.0
    c[axis] = 0.0
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    return (amplitude * np.exp(-r2 / (2.0 * sigma * sigma))).astype(np.float32)


def channel_config(u_in: float = 0.05, reynolds: float = 50.0, sgs: bool = True, buoyancy: bool = False,
                   periodic_yz: bool = True, thermal_periodic: bool = True) -> "A.BenchmarkConfig":
    """Benchmark-mode wind-tunnel configuration: velocity inlet at x-, convective
    outflow at x+, periodic y/z (reference face kinds, aero_lbm_capi.h:45-54).
    Reynolds 50 with reference length 1 gives nu = u_in / 50 (1e-3 at 0.05)."""
    cfg = A.BenchmarkConfig()
    cfg.abi_version = 1
    cfg.struct_size = 320
    cfg.flags = (A.FLAG_DISABLE_FAN_FORCING | A.FLAG_DISABLE_FAN_NOISE | A.FLAG_DISABLE_SPONGE
                 | A.FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND)
    if not sgs:
        cfg.flags |= A.FLAG_DISABLE_SGS
    if not buoyancy:
        cfg.flags |= A.FLAG_DISABLE_BUOYANCY
    cfg.enabled = 1
    cfg.preset = A.PRESET_NONE
    cfg.reynolds_number = reynolds
    cfg.rayleigh_number = 1.0e5
    cfg.prandtl_number = 0.71
    cfg.mach_number = 0.05
    cfg.reference_density = 1.0
    cfg.reference_length = 1.0
    cfg.gravity[1] = -1.0
    cfg.initial_velocity[0] = u_in
    tk = A.THERMAL_PERIODIC if thermal_periodic else A.THERMAL_ADIABATIC
    cfg.x_min.hydrodynamic_kind = A.HYDRO_VELOCITY
    cfg.x_min.velocity[0] = u_in
    cfg.x_min.thermal_kind = A.THERMAL_DIRICHLET
    cfg.x_max.hydrodynamic_kind = A.HYDRO_CONVECTIVE
    cfg.x_max.thermal_kind = A.THERMAL_ADIABATIC
    for f in (cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_PERIODIC if periodic_yz else A.HYDRO_SYMMETRY
        f.thermal_kind = tk if periodic_yz else A.THERMAL_ADIABATIC
    return cfg


def config1(n: int = 32):
    """Plain D3Q27 cumulant LBM, one n^3 brick: 15% solid boxes (seed 2), u=(0.05,0,0)
    +-1e-3 (seed 1), nu 1e-3, SGS on (reference C), buoyancy off."""
    solid = solid_boxes(n, n, n, 0.15, seed=2)
    packet = make_packet(solid, seed_u=1)
    cfg = channel_config(0.05, 50.0, sgs=True, buoyancy=False)
    return cfg, packet, solid


def config3(n: int = 128, beta: float = 0.1, cs: float = 0.1):
    """D3Q27 cumulant + SGS + Boussinesq, one n^3 brick: ~10% solid boxes (seed 5),
    inflow 0.05, Gaussian hot blob dT=0.02 sigma=8 at the z=0 floor (seed 4),
    gravity -z with beta; returns (cfg, overrides, packet, temperature, solid)."""
    solid = solid_boxes(n, n, n, 0.10, seed=5)
    t = hot_blob(n, n, n, 0.02, 8.0, seed=4, axis=2)
    t[solid.astype(bool)] = 0.0
    packet = make_packet(solid, seed_u=1, temperature=t)
    cfg = channel_config(0.05, 50.0, sgs=True, buoyancy=True)
    # z is the gravity axis: keep the floor/ceiling as walls for the thermal field
    cfg.z_min.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
    cfg.z_max.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
    cfg.z_min.thermal_kind = A.THERMAL_ADIABATIC
    cfg.z_max.thermal_kind = A.THERMAL_ADIABATIC
    ov = A.SolverOverrides()
    ov.struct_size = ctypes_sizeof_overrides()
    ov.sgs_constant = cs
    ov.sgs_nut_cap_ratio = float("nan")
    ov.boussinesq_beta = beta
    ov.thermal_diffusivity = float("nan")
    ov.gravity_axis = 2
    ov.gravity_sign = -1
    ov.kernel_variant = A.KERNEL_FAST
    return cfg, ov, packet, t, solid


def heated_cavity_config(three_d: bool = True, reference_length: float = 1.0) -> "A.BenchmarkConfig":
    """Differentially heated cavity preset (ref: native/src/aero_lbm_jni.cpp:307-338
    defaults, 384-400 preset): Ra 1e5, Pr 0.71, Ma 0.15, internal source off, no-slip
    walls, hot x- wall at +0.5 and cold x+ wall at -0.5, other walls adiabatic; the
    2-D variant is periodic in z. The solver runs it with BFECC thermal advection."""
    cfg = A.BenchmarkConfig()
    cfg.abi_version = 1
    cfg.struct_size = 320
    cfg.flags = (A.FLAG_DISABLE_FAN_FORCING | A.FLAG_DISABLE_FAN_NOISE | A.FLAG_DISABLE_SPONGE
                 | A.FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND | A.FLAG_DISABLE_SGS
                 | A.FLAG_DISABLE_INTERNAL_THERMAL_SOURCE)
    cfg.enabled = 1
    cfg.preset = A.PRESET_HEATED_CAVITY_3D if three_d else A.PRESET_HEATED_CAVITY_2D
    cfg.reynolds_number = 100.0
    cfg.rayleigh_number = 1.0e5
    cfg.prandtl_number = 0.71
    cfg.mach_number = 0.15
    cfg.reference_density = 1.0
    cfg.reference_length = reference_length
    cfg.gravity[1] = -1.0
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
        f.thermal_kind = A.THERMAL_ADIABATIC
    cfg.x_min.thermal_kind = A.THERMAL_DIRICHLET
    cfg.x_min.temperature = 0.5
    cfg.x_max.thermal_kind = A.THERMAL_DIRICHLET
    cfg.x_max.temperature = -0.5
    if not three_d:
        for f in (cfg.z_min, cfg.z_max):
            f.hydrodynamic_kind = A.HYDRO_PERIODIC
            f.thermal_kind = A.THERMAL_PERIODIC
    return cfg


def cylinder_config() -> "A.BenchmarkConfig":
    """Cylinder crossflow preset (ref: native/src/aero_lbm_jni.cpp:307-338 defaults,
    366-380 preset): Re 100, buoyancy and internal source off, no-slip walls, x- velocity
    inlet with u_max 0.12 (mean 0.08, parabolic in y), x+ convective outflow, periodic
    z. The solid cylinder comes with the payload; the solver adds the drag integral."""
    cfg = heated_cavity_config(True)
    cfg.preset = A.PRESET_CYLINDER_CROSSFLOW_2D
    cfg.flags = (A.FLAG_DISABLE_FAN_FORCING | A.FLAG_DISABLE_FAN_NOISE | A.FLAG_DISABLE_SPONGE
                 | A.FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND | A.FLAG_DISABLE_SGS
                 | A.FLAG_DISABLE_INTERNAL_THERMAL_SOURCE | A.FLAG_DISABLE_BUOYANCY)
    cfg.reynolds_number = 100.0
    cfg.mach_number = 0.05
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_BOUNCE_BACK
        f.thermal_kind = A.THERMAL_DISABLED
        f.temperature = 0.0
    cfg.x_min.hydrodynamic_kind = A.HYDRO_VELOCITY
    cfg.initial_velocity[0] = 0.08
    cfg.x_min.velocity[0] = 1.5 * 0.08
    cfg.x_max.hydrodynamic_kind = A.HYDRO_CONVECTIVE
    for f in (cfg.z_min, cfg.z_max):
        f.hydrodynamic_kind = A.HYDRO_PERIODIC
        f.thermal_kind = A.THERMAL_PERIODIC
    return cfg


def cylinder_mask(nx: int, ny: int, nz: int) -> np.ndarray:
    """Solid mask of the preset's cylinder (ref: native/src/aero_lbm_jni.cpp:200-228),
    binary32 geometry, extruded along z."""
    f32 = np.float32
    length, height = f32(max(1, nx - 1)), f32(max(1, ny - 1))
    d = max(f32(2.5), f32(0.1) / f32(2.2) * length)
    d = min(d, max(f32(2.5), height - f32(4.0)))
    r = f32(0.5) * d
    cx = f32(0.2) / f32(2.2) * length
    cy = f32(0.0) + f32(0.2) / f32(0.41) * height
    X, Y = np.meshgrid(np.arange(nx, dtype=f32), np.arange(ny, dtype=f32), indexing="ij")
    dx, dy = X - cx, Y - cy
    disk = dx * dx + dy * dy <= r * r
    return np.repeat(disk[:, :, None], nz, axis=2).astype(np.uint8)


def ctypes_sizeof_overrides() -> int:
    import ctypes
    return ctypes.sizeof(A.SolverOverrides)


def reference_overrides(kernel_variant: int = 0):
    """Overrides that keep every reference constant (NaN / -1 = keep)."""
    ov = A.SolverOverrides()
    ov.struct_size = ctypes_sizeof_overrides()
    nan = float("nan")
    ov.sgs_constant = nan
    ov.sgs_nut_cap_ratio = nan
    ov.boussinesq_beta = nan
    ov.thermal_diffusivity = nan
    ov.gravity_axis = -1
    ov.gravity_sign = 0
    ov.kernel_variant = kernel_variant
    return ov


# ---------------------------------------------------------------- mesoscale (D3Q19) workloads
# 25 forcing channels per cell (ref: native/src/aero_lbm_mesoscale.cpp:110-139); y is vertical.
MESO_CHANNELS = 25


def value_noise_2d(nx: int, nz: int, seed: int, octaves: int = 4, base_cells: int = 64) -> np.ndarray:
    """Seeded multi-octave value noise in [0, 1) on an nx x nz grid (bilinear lattice noise)."""
    out = np.zeros((nx, nz))
    amp_total = 0.0
    amp = 1.0
    period = base_cells
    for o in range(octaves):
        gx, gz = nx // period + 2, nz // period + 2
        lattice = splitmix_array(seed * 1000 + o, gx * gz).reshape(gx, gz)
        xs = np.arange(nx) / period
        zs = np.arange(nz) / period
        x0 = np.floor(xs).astype(int)
        z0 = np.floor(zs).astype(int)
        fx = (xs - x0)[:, None]
        fz = (zs - z0)[None, :]
        a = lattice[x0][:, z0]
        b = lattice[x0 + 1][:, z0]
        c = lattice[x0][:, z0 + 1]
        d = lattice[x0 + 1][:, z0 + 1]
        out += amp * ((a * (1 - fx) + b * fx) * (1 - fz) + (c * (1 - fx) + d * fx) * fz)
        amp_total += amp
        amp *= 0.5
        period = max(1, period // 2)
    return out / amp_total


def meso_forcing(nx: int, ny: int, nz: int, wind_lattice: float = 0.04, angle_deg: float = 30.0,
                 velocity_scale: float = 12.8, max_height: float = 24.0, seed: int = 3, z0: float = 0.1,
                 ambient: float = 288.15, deep: float = 289.65, roughness: float = 0.1) -> np.ndarray:
    """Terrain-following mesoscale forcing: heightmap from 4-octave value noise (heights
    0..max_height cells) as solid cells below the surface, geostrophic wind of
    `wind_lattice` lattice units rotated `angle_deg` from +x with a log-law profile in
    the vertical (y), ambient/surface targets at `ambient`, deep ground at `deep`."""
    h = value_noise_2d(nx, nz, seed) * max_height
    y = np.arange(ny)[None, :, None].astype(np.float64)
    hh = h[:, None, :]
    solid = (y < np.floor(hh)).astype(np.float32)
    above = np.maximum(y - np.floor(hh) + 0.5, 0.0)
    profile = np.log1p(above / z0) / np.log1p((ny - 0.5) / z0)
    u_ms = wind_lattice * velocity_scale
    ang = np.deg2rad(angle_deg)
    F = np.zeros((nx, ny, nz, MESO_CHANNELS), dtype=np.float32)
    F[..., 0] = np.broadcast_to(hh, (nx, ny, nz))
    F[..., 2] = ambient
    F[..., 3] = deep
    F[..., 4] = ambient
    F[..., 5] = roughness
    F[..., 6] = (u_ms * np.cos(ang) * profile).astype(np.float32)
    F[..., 7] = (u_ms * np.sin(ang) * profile).astype(np.float32)
    F[..., 24] = solid
    return F.reshape(-1)


def meso_forcing_varying(nx: int, ny: int, nz: int, seed: int = 11, base: np.ndarray | None = None) -> np.ndarray:
    """meso_forcing (or `base`) with every channel the coefficient planes derive from
    varied from cell to cell by a seeded uniform amount: temperatures and nested deltas
    (up to +1 K / +0.5 K), roughness, humidity, heating, updrafts. No coefficient plane
    is then uniform over the domain."""
    F = (meso_forcing(nx, ny, nz) if base is None else base).reshape(nx, ny, nz, MESO_CHANNELS).astype(np.float64)
    r = np.random.default_rng(seed).random(F.shape)
    for ch, amp in ((2, 1.0), (3, 1.0), (4, 1.0), (5, 0.1), (9, 0.2), (10, 0.3), (11, 0.1), (16, 0.2),
                    (17, 0.1), (18, 0.05), (19, 0.05), (22, 0.5), (23, 0.5)):
        F[..., ch] += amp * r[..., ch]
    return F.astype(np.float32).reshape(-1)


def meso_config(nx: int, ny: int, nz: int, nu_lattice: float = 5e-4, dx: float = 64.0, dt: float = 5.0):
    """Mesoscale config whose molecular viscosity gives lattice viscosity `nu_lattice`."""
    cfg = A.MesoscaleConfig()
    cfg.abi_version = 3
    cfg.struct_size = 96
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    cfg.dx_m, cfg.dt_s = dx, dt
    cfg.molecular_nu_m2_s = nu_lattice * dx * dx / dt
    cfg.prandtl_air = 0.71
    cfg.turbulent_prandtl = 0.85
    cfg.reference_density_kg_m3 = 1.225
    cfg.ambient_air_temperature_k = 288.15
    cfg.deep_ground_temperature_k = 289.65
    return cfg


def adjoint_case(n: int = 64, heat_grid: int = 16, probes: int = 512, fill: float = 0.05, u_in: float = 0.04,
                 seed: int = 21) -> dict:
    """Inflow/heat fit of one n^3 brick: solid boxes, seeded probe cells, a hidden
    ground-truth parameter set and the starting guess.

    inflow_* are [n*n*3] float32 (y-major, z fastest, then u, v, w) on the x = 0
    face; heat_* are [heat_grid^3] float32 lattice source rates. The truth is a
    log-law shear profile in z with a lateral modulation and small cross
    components; the start is the uniform inflow u_in and no heat."""
    solid = solid_boxes(n, n, n, fill, seed, edge=(2, max(3, n // 8)), keep_clear_x=3)
    r = splitmix_array(seed + 1, probes + heat_grid ** 3)
    fluid = np.flatnonzero(solid.reshape(-1) == 0)
    pick = np.unique((r[:probes] * len(fluid)).astype(np.int64))
    extra = np.setdiff1d(fluid, fluid[pick])[: probes - len(pick)]
    probe_cells = np.sort(np.concatenate([fluid[pick], extra])).astype(np.int32)
    y, z = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    prof = u_in * (0.6 + 0.4 * np.log1p(z) / np.log1p(n - 1)) * (1.0 + 0.1 * np.sin(2.0 * np.pi * y / n))
    truth = np.stack([prof, 0.004 * np.cos(2.0 * np.pi * z / n), 0.003 * np.sin(2.0 * np.pi * y / n)], axis=-1)
    heat_truth = 0.004 * (2.0 * r[probes:] - 1.0)
    start = np.zeros((n, n, 3))
    start[..., 0] = u_in
    return {
        "n": n, "heat_grid": heat_grid, "solid": solid, "probes": probe_cells,
        "inflow_truth": truth.reshape(-1).astype(np.float32),
        "heat_truth": heat_truth.astype(np.float32),
        "inflow_start": start.reshape(-1).astype(np.float32),
        "heat_start": np.zeros(heat_grid ** 3, np.float32),
    }


# the batched workload: (edge, count) of its bricks
BATCH_BRICKS = ((32, 40), (64, 12), (128, 4))


def batch_brick_specs(seed: int = 2024) -> list:
    """The 56 bricks of the batched workload (40 x 32^3 + 12 x 64^3 + 4 x 128^3), drawn
    from one seeded generator: per brick a solid-box seed, a fill in [0.05, 0.20], an
    inflow speed in [0.02, 0.08] and a direction uniform in the x-y plane."""
    import math
    rng = np.random.default_rng(seed)
    specs = []
    for edge, count in BATCH_BRICKS:
        for _ in range(count):
            s = int(rng.integers(1, 1 << 30))
            fill = float(rng.uniform(0.05, 0.20))
            speed = float(rng.uniform(0.02, 0.08))
            ang = float(rng.uniform(0.0, 2.0 * math.pi))
            specs.append({"edge": edge, "seed": s, "fill": fill, "speed": speed, "angle": ang})
    return specs


def batch_brick_case(spec: dict):
    """(config, packet, temperature, solid) of one batched brick: velocity inlet with
    the brick's inflow, SGS and Boussinesq buoyancy on, seeded solid boxes and a seeded
    hot blob (amplitude 0.02, sigma edge / 16) on the floor."""
    import math
    edge, seed, speed, ang = spec["edge"], spec["seed"], spec["speed"], spec["angle"]
    v = (speed * math.cos(ang), speed * math.sin(ang), 0.0)
    cfg = channel_config(speed, 50.0, sgs=True, buoyancy=True)
    cfg.initial_velocity[0], cfg.initial_velocity[1] = v[0], v[1]
    cfg.x_min.velocity[0], cfg.x_min.velocity[1] = v[0], v[1]
    solid = solid_boxes(edge, edge, edge, spec["fill"], seed=seed, edge=(2, max(3, edge // 8)))
    t = hot_blob(edge, edge, edge, 0.02, edge / 16.0, seed=seed + 1, axis=2)
    t[solid.astype(bool)] = 0.0
    pk = make_packet(solid, u0=v, perturb=1e-3, seed_u=seed + 2, temperature=t)
    return cfg, pk, t, solid
