"""Wind-tunnel cases shared by the aero_solver_* tests: the benchmark
configuration and payload packet that the reference's wind-tunnel layer builds
(native/src/aero_solver_capi.cpp:130-160, 209-258), restated here as the
expected values."""
import ctypes
import math

import numpy as np

import aero_lbm as A

DIMS = (24, 12, 10)
DX, DT = 0.5, 0.02


def tunnel_config(bc):
    """The benchmark configuration the reference's layer builds for boundary bc."""
    lib = A.load()
    cfg = A.default_config()
    cfg.enabled = 1
    cfg.preset = A.PRESET_NONE
    cfg.flags |= (A.FLAG_DISABLE_FAN_FORCING | A.FLAG_DISABLE_FAN_NOISE | A.FLAG_DISABLE_SPONGE
                  | A.FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND | A.FLAG_DISABLE_SGS | A.FLAG_DISABLE_INTERNAL_THERMAL_SOURCE
                  | A.FLAG_DISABLE_BUOYANCY)
    cfg.flags &= ~A.FLAG_DISABLE_CONVECTIVE_OUTFLOW
    f32 = np.float32
    scale = f32(DX) / f32(DT)
    u = [f32(v) / scale for v in (bc.inlet_vx, bc.inlet_vy, bc.inlet_vz)]
    cfg.reference_density = max(1e-6, bc.density)
    cfg.reference_temperature = 0.0
    cfg.reference_length = 1.0
    for k in range(3):
        cfg.initial_velocity[k] = u[k]
    cfg.initial_pressure = bc.outlet_pressure
    speed = max(f32(math.sqrt(f32(u[0] * u[0] + u[1] * u[1]) + u[2] * u[2])), f32(1e-6))
    nu = f32(bc.viscosity) * f32(DT) / (f32(DX) * f32(DX))
    cfg.reynolds_number = max(1e-6, f32(speed / nu))
    cfg.mach_number = min(max(f32(speed / f32(math.sqrt(f32(1.0) / f32(3.0)))), 1e-4), 0.30)
    faces = [cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max]
    kind = {A.SOLVER_CLOSED: A.HYDRO_BOUNCE_BACK, A.SOLVER_PERIODIC: A.HYDRO_PERIODIC}.get(bc.mode, A.HYDRO_SYMMETRY)
    for f in faces:
        ctypes.memset(ctypes.addressof(f), 0, ctypes.sizeof(f))
        f.hydrodynamic_kind = kind
        f.thermal_kind = A.THERMAL_DISABLED
    if bc.mode in (A.SOLVER_WIND_TUNNEL, A.SOLVER_WIND_TUNNEL_PRESSURE_OUTLET):
        cfg.x_min.hydrodynamic_kind = A.HYDRO_VELOCITY
        if bc.mode == A.SOLVER_WIND_TUNNEL:
            cfg.x_max.hydrodynamic_kind = A.HYDRO_CONVECTIVE
        else:
            cfg.x_max.hydrodynamic_kind = A.HYDRO_PRESSURE
            cfg.x_max.pressure = bc.outlet_pressure
    for k in range(3):
        cfg.x_min.velocity[k] = u[k]
    return cfg


def tunnel_packet(solid, bc, flow=None, obstacle=True):
    n = solid.size
    pk = np.zeros((n, 11), np.float32)
    if obstacle:
        pk[:, 0] = solid.reshape(-1) != 0
    scale = np.float32(DX) / np.float32(DT)
    if flow is None:
        vals = np.array([bc.inlet_vx, bc.inlet_vy, bc.inlet_vz], np.float32) / scale
        pk[:, 5:8] = vals
        pk[:, 8] = bc.outlet_pressure
    else:
        pk[:, 5:8] = flow.reshape(-1, 4)[:, :3].astype(np.float32) / scale
        pk[:, 8] = flow.reshape(-1, 4)[:, 3]
    pk[solid.reshape(-1) != 0, 5:9] = 0.0
    return pk.reshape(-1)


def boundary(mode, vx=1.25, vy=0.1, p=0.004, visc=0.3):
    b = A.BoundaryDesc()
    A.load().aero_solver_default_boundary(ctypes.byref(b))
    b.mode, b.inlet_vx, b.inlet_vy, b.outlet_pressure, b.viscosity = mode, vx, vy, p, visc
    return b


