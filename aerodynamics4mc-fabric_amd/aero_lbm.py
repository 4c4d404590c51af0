"""ctypes binding of libaero_lbm.so (the MI355X build of the reference's C ABI).

This is the Python host mirror of the reference's native interface, the same
way the reference's own tools bind it (native/tools/benchmark_solver_dll.py,
native/tools/check_d3q27_fp16.py): the struct layouts follow
native/include/aero_lbm_capi.h:76-164 and every function keeps the reference's
name, argument meaning and 1/0 return convention.

Loading fails loudly: there is no CPU fallback. If the shared library has not
been built (``python -c "import __graft_entry__ as g; g.build()"``) or cannot
find a GPU, calls raise ``AeroLbmError``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_float, c_int, c_longlong,
                    c_uint8, c_uint16, c_int16, c_uint32, c_uint64, c_double, c_void_p)

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AERO_LBM_LIBRARY selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("AERO_LBM_LIBRARY") or os.path.join(_HERE, "lib", "libaero_lbm.so")

# ---- enums (native/include/aero_lbm_capi.h:16-74) ----
SOLVER_AUTO = 0
SOLVER_CLASSIC_D3Q27 = 1
SOLVER_COMPACT_EXPERIMENTAL = 2
SOLVER_D3Q27_FP16_INPLACE = 3
SOLVER_D3Q27_BF16_INPLACE = 4  # extension
SOLVER_D3Q27_FP32_INPLACE = 5  # extension
SOLVER_D3Q27_S16_INPLACE = 6  # extension: 16-bit fixed point of f - w_q

PRESET_NONE = 0
PRESET_TAYLOR_GREEN_3D = 1
PRESET_LID_DRIVEN_CAVITY_2D = 2
PRESET_LID_DRIVEN_CAVITY_3D = 3
PRESET_CYLINDER_CROSSFLOW_2D = 4
PRESET_HEATED_CAVITY_2D = 5
PRESET_HEATED_CAVITY_3D = 6

HYDRO_INHERIT, HYDRO_PERIODIC, HYDRO_BOUNCE_BACK, HYDRO_MOVING_WALL = 0, 1, 2, 3
HYDRO_VELOCITY, HYDRO_PRESSURE, HYDRO_CONVECTIVE, HYDRO_SYMMETRY = 4, 5, 6, 7
THERMAL_INHERIT, THERMAL_DISABLED, THERMAL_PERIODIC, THERMAL_ADIABATIC = 0, 1, 2, 3
THERMAL_DIRICHLET, THERMAL_NEUMANN = 4, 5

FLAG_DISABLE_FAN_FORCING = 1 << 0
FLAG_DISABLE_FAN_NOISE = 1 << 1
FLAG_DISABLE_SPONGE = 1 << 2
FLAG_DISABLE_CONVECTIVE_OUTFLOW = 1 << 3
FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND = 1 << 4
FLAG_DISABLE_SGS = 1 << 5
FLAG_DISABLE_INTERNAL_THERMAL_SOURCE = 1 << 6
FLAG_DISABLE_BUOYANCY = 1 << 7

KERNEL_FAST = 0
KERNEL_EXACT = 1


class BoundaryFace(Structure):
    _fields_ = [("hydrodynamic_kind", c_int), ("thermal_kind", c_int),
                ("velocity", c_float * 3), ("pressure", c_float),
                ("temperature", c_float), ("heat_flux", c_float)]


class BenchmarkConfig(Structure):
    _fields_ = [("abi_version", c_uint32), ("struct_size", c_uint32), ("flags", c_uint32),
                ("enabled", c_int), ("preset", c_int), ("reserved0", c_int),
                ("reynolds_number", c_float), ("rayleigh_number", c_float),
                ("prandtl_number", c_float), ("mach_number", c_float),
                ("reference_density", c_float), ("reference_temperature", c_float),
                ("reference_length", c_float), ("body_force", c_float * 3),
                ("gravity", c_float * 3), ("initial_velocity", c_float * 3),
                ("initial_pressure", c_float), ("initial_temperature", c_float),
                ("x_min", BoundaryFace), ("x_max", BoundaryFace),
                ("y_min", BoundaryFace), ("y_max", BoundaryFace),
                ("z_min", BoundaryFace), ("z_max", BoundaryFace),
                ("reserved", c_uint32 * 8)]


class SolverOverrides(Structure):
    _fields_ = [("struct_size", c_uint32), ("sgs_constant", c_float),
                ("sgs_nut_cap_ratio", c_float), ("boussinesq_beta", c_float),
                ("thermal_diffusivity", c_float), ("gravity_axis", c_int),
                ("gravity_sign", c_int), ("kernel_variant", c_int), ("reserved", c_int * 8)]


class TimingSnapshot(Structure):
    _fields_ = [("ticks", c_uint64)] + [(n, c_double) for n in (
        "last_payload_copy_ms", "last_solver_ms", "last_readback_ms", "last_total_ms",
        "avg_payload_copy_ms", "avg_solver_ms", "avg_readback_ms", "avg_total_ms")]


class MesoscaleConfig(Structure):
    _fields_ = [("abi_version", c_uint32), ("struct_size", c_uint32),
                ("nx", c_int), ("ny", c_int), ("nz", c_int),
                ("dx_m", c_float), ("dt_s", c_float), ("molecular_nu_m2_s", c_float),
                ("prandtl_air", c_float), ("turbulent_prandtl", c_float),
                ("reference_density_kg_m3", c_float), ("ambient_air_temperature_k", c_float),
                ("deep_ground_temperature_k", c_float), ("background_wind_m_s", c_float * 3),
                ("reserved", c_uint32 * 8)]


class MesoscaleTransport(Structure):
    _fields_ = [("velocity_scale_m_s_per_lattice", c_float), ("nu_molecular_lattice", c_float),
                ("alpha_molecular_lattice", c_float), ("tau_shear_molecular", c_float),
                ("tau_thermal_molecular", c_float)]


assert ctypes.sizeof(BoundaryFace) == 32
assert ctypes.sizeof(BenchmarkConfig) == 320
assert ctypes.sizeof(TimingSnapshot) == 72
assert ctypes.sizeof(MesoscaleConfig) == 96
assert ctypes.sizeof(MesoscaleTransport) == 20


class WorldDelta(Structure):
    _fields_ = [("type", c_int), ("x", c_int), ("y", c_int), ("z", c_int), ("data0", c_int), ("data1", c_int),
                ("data2", c_int), ("data3", c_int), ("value0", c_float), ("value1", c_float), ("value2", c_float),
                ("value3", c_float)]


class AdjointConfig(Structure):
    _fields_ = [("n", c_int), ("heat_grid", c_int), ("steps", c_int), ("checkpoint_every", c_int), ("nu", c_float),
                ("beta", c_float), ("alpha", c_float), ("u0", c_float * 3), ("probe_count", c_int)]


class GridDesc(Structure):
    """AeroGridDesc (include/aero_solver_capi.h)."""
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("dx", c_float), ("dt", c_float)]


class BoundaryDesc(Structure):
    """AeroBoundaryDesc (include/aero_solver_capi.h)."""
    _fields_ = [("mode", c_int), ("inlet_vx", c_float), ("inlet_vy", c_float), ("inlet_vz", c_float),
                ("outlet_pressure", c_float), ("density", c_float), ("viscosity", c_float)]


class StepInput(Structure):
    _fields_ = [("solid_mask", POINTER(ctypes.c_uint8)), ("prev_flow", POINTER(c_float)), ("grid", GridDesc),
                ("boundary", BoundaryDesc), ("steps", c_int)]


class StepOutput(Structure):
    _fields_ = [("flow_out", POINTER(c_float)), ("max_velocity", c_float), ("status", c_int)]


SOLVER_WIND_TUNNEL, SOLVER_WIND_TUNNEL_PRESSURE_OUTLET, SOLVER_CLOSED, SOLVER_PERIODIC = 1, 2, 3, 4


class AeroLbmError(RuntimeError):
    pass


_F = POINTER(c_float)
_I = POINTER(c_int)
_U8 = POINTER(ctypes.c_uint8)
_U16 = POINTER(ctypes.c_uint16)
_I16 = POINTER(ctypes.c_int16)

# name -> (restype, argtypes); mirrors include/aero_lbm_capi.h
PROTOTYPES = {
    "aero_solver_default_grid": (None, [POINTER(GridDesc)]),
    "aero_solver_default_boundary": (None, [POINTER(BoundaryDesc)]),
    "aero_solver_create": (c_int, [c_int, c_int, c_int, c_float, c_float, POINTER(c_longlong)]),
    "aero_solver_create_with_grid": (c_int, [POINTER(GridDesc), POINTER(c_longlong)]),
    "aero_solver_set_solid_mask": (c_int, [c_longlong, POINTER(ctypes.c_uint8), c_int]),
    "aero_solver_set_flow_state": (c_int, [c_longlong, POINTER(c_float), c_int]),
    "aero_solver_step_wind_tunnel": (c_int, [c_longlong, POINTER(BoundaryDesc), c_int, POINTER(c_float), c_int]),
    "aero_solver_advance_wind_tunnel": (c_int, [c_longlong, POINTER(BoundaryDesc), c_int]),
    "aero_solver_run_wind_tunnel": (c_int, [POINTER(StepInput), POINTER(StepOutput)]),
    "aero_solver_destroy": (None, [c_longlong]),
    "aero_solver_last_error": (c_char_p, []),
    "aero_solver_runtime_info": (c_char_p, []),
    "aero_solver_finish": (c_int, []),
    "aero_lbm_init": (c_int, [c_int, c_int, c_int]),
    "aero_lbm_step": (c_int, [_F, c_int, c_longlong, _F]),
    "aero_lbm_init_rect": (c_int, [c_int, c_int, c_int, c_int, c_int]),
    "aero_lbm_step_rect": (c_int, [_F, c_int, c_int, c_int, c_longlong, _F]),
    "aero_lbm_step_rect_scaled": (c_int, [_F, c_int, c_int, c_int, c_longlong, c_float, _F]),
    "aero_lbm_step_rect_cached": (c_int, [c_int, c_int, c_int, c_longlong, _F]),
    "aero_lbm_step_rect_cached_scaled": (c_int, [c_int, c_int, c_int, c_longlong, c_float, _F]),
    "aero_lbm_patch_d3q27_f16_static_cells_rect": (c_int, [c_int, c_int, c_int, c_longlong, c_int, _I, _U8, _U8]),
    "aero_lbm_step_rect_with_sparse_overlays": (c_int, [_F, c_int, c_int, c_int, c_longlong, _I, _F, c_int, _F]),
    "aero_lbm_shift_context": (c_int, [c_int, c_longlong, c_int, c_int, c_int]),
    "aero_lbm_context_compact_initialized": (c_int, [c_longlong]),
    "aero_lbm_exchange_halo": (c_int, [c_int, c_longlong, c_longlong, c_int, c_int, c_int]),
    "aero_lbm_exchange_halo_layers_rect": (c_int, [c_int, c_int, c_int, c_int, c_longlong, c_longlong, c_int, c_int,
                                                   c_int]),
    "aero_lbm_has_context": (c_int, [c_longlong]),
    "aero_lbm_context_realtime_cached_initialized": (c_int, [c_longlong]),
    "aero_lbm_get_temperature_state_rect": (c_int, [c_int, c_int, c_int, c_longlong, _F]),
    "aero_lbm_sample_temperature_point_rect": (c_int, [c_int, c_int, c_int, c_longlong, c_int, c_int, c_int, _F]),
    "aero_lbm_sample_flow_point_rect": (c_int, [c_int, c_int, c_int, c_longlong, c_int, c_int, c_int, _F]),
    "aero_lbm_extract_flow_atlas_rect": (c_int, [c_int, c_int, c_int, c_longlong, c_int, _F, c_int]),
    "aero_lbm_copy_flow_temperature_subrect": (c_int, [c_int, c_int, c_int, c_longlong, c_int, c_int, c_int,
                                                       c_int, c_int, c_int, _F, _F]),
    "aero_lbm_get_flow_state_rect": (c_int, [c_int, c_int, c_int, c_longlong, _F]),
    "aero_lbm_set_temperature_state_rect": (c_int, [c_int, c_int, c_int, c_longlong, _F]),
    "aero_lbm_get_last_force": (c_int, [c_longlong, _F, _F, _F]),
    "aero_lbm_release_context": (None, [c_longlong]),
    "aero_lbm_shutdown": (None, []),
    "aero_lbm_runtime_info": (c_char_p, []),
    "aero_lbm_last_error": (c_char_p, []),
    "aero_lbm_timing_info": (c_char_p, []),
    "aero_lbm_memory_info": (c_char_p, []),
    "aero_lbm_reset_timing": (None, []),
    "aero_lbm_set_realtime_solver_mode": (None, [c_int]),
    "aero_lbm_finish": (c_int, []),
    "aero_lbm_get_timing_snapshot": (c_int, [POINTER(TimingSnapshot)]),
    "aero_lbm_benchmark_default_config": (None, [POINTER(BenchmarkConfig)]),
    "aero_lbm_benchmark_default_preset_config": (c_int, [c_int, POINTER(BenchmarkConfig)]),
    "aero_lbm_benchmark_set_config": (c_int, [POINTER(BenchmarkConfig)]),
    "aero_lbm_benchmark_get_config": (c_int, [POINTER(BenchmarkConfig)]),
    "aero_lbm_benchmark_reset_config": (None, []),
    "aero_lbm_benchmark_info": (c_char_p, []),
    "aero_lbm_default_solver_overrides": (None, [POINTER(SolverOverrides)]),
    "aero_lbm_set_solver_overrides": (c_int, [POINTER(SolverOverrides)]),
    "aero_lbm_get_solver_overrides": (c_int, [POINTER(SolverOverrides)]),
    "aero_lbm_run_cached_steps_timed": (c_int, [c_int, c_int, c_int, c_longlong, c_int, _F]),
    "aero_lbm_get_conservation_rect": (c_int, [c_int, c_int, c_int, c_longlong, POINTER(c_double)]),
    "aero_lbm_analysis_max_compressed_size_3d": (c_longlong, [c_int, c_int, c_int]),
    "aero_lbm_analysis_compress_float_grid_3d": (c_longlong, [_F, c_int, c_int, c_int, c_double, _U8, c_longlong]),
    "aero_lbm_analysis_decompress_float_grid_3d": (c_int, [_U8, c_longlong, c_int, c_int, c_int, _F]),
    "aero_lbm_analysis_compress_float_grids_3d_device": (c_int, [c_int, POINTER(c_void_p), _I, POINTER(c_double), c_void_p,
                                                                 POINTER(c_longlong), POINTER(c_longlong),
                                                                 POINTER(c_longlong)]),
    "aero_lbm_analysis_compress_grids_timed": (c_int, [c_int, POINTER(c_void_p), _I, POINTER(c_double), c_void_p,
                                                       POINTER(c_longlong), POINTER(c_longlong), c_int, _F]),
    "aero_lbm_mesoscale_default_config": (None, [POINTER(MesoscaleConfig)]),
    "aero_lbm_mesoscale_derive_transport": (c_int, [POINTER(MesoscaleConfig), POINTER(MesoscaleTransport)]),
    "aero_lbm_mesoscale_create_context": (c_int, [POINTER(MesoscaleConfig), POINTER(c_longlong)]),
    "aero_lbm_mesoscale_step_context": (c_int, [c_longlong, POINTER(MesoscaleConfig), _F, _F]),
    "aero_lbm_mesoscale_release_context": (None, [c_longlong]),
    "aero_lbm_mesoscale_set_tau_floor": (c_int, [c_float]),
    "aero_lbm_mesoscale_run_steps_timed": (c_int, [c_longlong, POINTER(MesoscaleConfig), _F, c_int, _F]),
    "aero_lbm_mesoscale_get_state": (c_int, [c_longlong, _F]),
    "aero_lbm_set_d3q27_f16_boundary_shell_rect": (c_int, [c_int, c_int, c_int, c_longlong, _F, c_int, c_int, c_int,
                                                           c_float]),
    "aero_lbm_clear_d3q27_f16_boundary_shell_rect": (c_int, [c_longlong]),
    "aero_lbm_simulation_create_service": (c_int, [POINTER(c_longlong)]),
    "aero_lbm_simulation_release_service": (None, [c_longlong]),
    "aero_lbm_simulation_set_focus": (c_int, [c_longlong, c_int, c_int, c_int, c_int]),
    "aero_lbm_simulation_submit_world_deltas": (c_int, [c_longlong, POINTER(WorldDelta), c_int]),
    "aero_lbm_simulation_ensure_brick_world_runtime": (c_int, [c_longlong, c_longlong, c_int, c_float, c_float]),
    "aero_lbm_simulation_set_brick_world_solver_mode": (c_int, [c_longlong, c_longlong, c_int]),
    "aero_lbm_simulation_set_brick_world_batching": (c_int, [c_longlong, c_longlong, c_int]),
    "aero_lbm_simulation_set_brick_world_active_hints": (c_int, [c_longlong, c_longlong, c_int, _I, c_int]),
    "aero_lbm_simulation_set_brick_world_exact_active_hints": (c_int, [c_longlong, c_longlong, c_int, _I, c_int]),
    "aero_lbm_simulation_get_brick_world_runtime_status": (c_int, [c_longlong, c_longlong, _I, c_int]),
    "aero_lbm_simulation_step_brick_world_runtime": (c_int, [c_longlong, c_longlong, c_int]),
    "aero_lbm_simulation_get_brick_world_resident_brick_count": (c_int, [c_longlong, c_longlong]),
    "aero_lbm_simulation_copy_brick_world_resident_bricks": (c_int, [c_longlong, c_longlong, _I, c_int]),
    "aero_lbm_simulation_get_brick_world_active_brick_count": (c_int, [c_longlong, c_longlong]),
    "aero_lbm_simulation_copy_brick_world_active_bricks": (c_int, [c_longlong, c_longlong, _I, c_int]),
    "aero_lbm_simulation_upload_brick_world_static_brick": (c_int, [c_longlong, c_longlong, c_int, c_int, c_int, c_int,
                                                                    _U8, _U8, _U16, _F, _U8, _U8]),
    "aero_lbm_simulation_upload_brick_world_static_brick_with_sources": (
        c_int, [c_longlong, c_longlong, c_int, c_int, c_int, c_int, _U8, _U8, _U16, _F, _U8, _F, _U8, _U8]),
    "aero_lbm_simulation_queue_brick_world_static_brick_upload": (c_int, [c_longlong, c_longlong, c_int, c_int, c_int,
                                                                          c_int, _U8, _U8, _U16, _F, _U8, _U8]),
    "aero_lbm_simulation_upload_brick_world_dynamic_brick": (c_int, [c_longlong, c_longlong, c_int, c_int, c_int,
                                                                     c_int, _F, _F, _F]),
    "aero_lbm_simulation_upload_brick_world_boundary_reference_brick": (c_int, [c_longlong, c_longlong, c_int, c_int,
                                                                                c_int, c_int, _F, _F, _F]),
    "aero_lbm_simulation_copy_brick_world_dynamic_brick": (c_int, [c_longlong, c_longlong, c_int, c_int, c_int, c_int,
                                                                   _F, _F, _F]),
    "aero_lbm_simulation_copy_brick_world_packed_flow_atlas": (c_int, [c_longlong, c_longlong, c_int, c_int, c_int,
                                                                       c_int, c_int, _I16, c_int]),
    "aero_lbm_simulation_set_packed_flow_atlas": (c_int, [c_longlong, c_longlong, _I16, c_int]),
    "aero_lbm_simulation_poll_packed_flow_atlas": (c_int, [c_longlong, c_longlong, _I16, c_int]),
    "aero_lbm_simulation_runtime_info": (c_char_p, []),
    "aero_lbm_simulation_last_error": (c_char_p, []),
    "aero_lbm_batch_create": (c_int, [POINTER(c_longlong), c_int, c_int, c_int, _F, c_int, POINTER(c_longlong)]),
    "aero_lbm_batch_run": (c_int, [c_longlong, c_int, c_int, _F]),
    "aero_lbm_batch_run_per_brick": (c_int, [c_longlong, c_int, _F]),
    "aero_lbm_batch_copy_atlas": (c_int, [c_longlong, c_int, _I16, c_int]),
    "aero_lbm_batch_info": (c_char_p, [c_longlong]),
    "aero_lbm_batch_release": (None, [c_longlong]),
    "aero_lbm_adjoint_create": (c_int, [POINTER(AdjointConfig), _U8, _I, _F, POINTER(c_longlong)]),
    "aero_lbm_adjoint_forward": (c_int, [c_longlong, _F, _F, POINTER(c_double), _F]),
    "aero_lbm_adjoint_backward": (c_int, [c_longlong, _F, _F]),
    "aero_lbm_adjoint_forward_fp64": (c_int, [c_longlong, POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
    "aero_lbm_adjoint_release": (None, [c_longlong]),
}


def header_symbols(header_path: str) -> list[str]:
    """Names of every function declared with AERO_LBM_CAPI_EXPORT / JNIEXPORT in a header."""
    import re
    text = open(header_path).read()
    return sorted(set(re.findall(r"AERO_LBM_CAPI_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text)))


_lib = None


def load(path: str = LIB_PATH):
    """Load libaero_lbm.so and attach prototypes. Raises AeroLbmError when absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AeroLbmError(f"libaero_lbm.so not built at {path}; run __graft_entry__.build()")
    # One HIP runtime per process: when torch is installed, load the one it ships
    # before the library binds libamdhip64, so torch tensors and the library can share
    # the device (torch fails to initialise under a runtime loaded before it).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_F)


def last_error() -> str:
    msg = load().aero_lbm_last_error()
    return msg.decode() if msg else ""


def check(ok: int, what: str):
    if not ok:
        raise AeroLbmError(f"{what} failed: {last_error()}")


def default_config() -> BenchmarkConfig:
    cfg = BenchmarkConfig()
    load().aero_lbm_benchmark_default_config(ctypes.byref(cfg))
    return cfg


def default_overrides() -> SolverOverrides:
    ov = SolverOverrides()
    load().aero_lbm_default_solver_overrides(ctypes.byref(ov))
    return ov


class Context:
    """One solver context (a brick) of shape (nx, ny, nz), keyed like the reference."""

    def __init__(self, nx: int, ny: int, nz: int, key: int = 1, in_ch: int = 11, out_ch: int = 4):
        self.lib = load()
        self.nx, self.ny, self.nz, self.key = nx, ny, nz, key
        self.cells = nx * ny * nz
        self.in_ch, self.out_ch = in_ch, out_ch

    def step(self, packet: np.ndarray | None, out: np.ndarray | None = None, scale: float | None = None) -> None:
        outp = fptr(out) if out is not None else None
        if packet is None:
            if scale is None:
                ok = self.lib.aero_lbm_step_rect_cached(self.nx, self.ny, self.nz, self.key, outp)
            else:
                ok = self.lib.aero_lbm_step_rect_cached_scaled(self.nx, self.ny, self.nz, self.key, scale, outp)
        else:
            packet = np.ascontiguousarray(packet, dtype=np.float32)
            if scale is None:
                ok = self.lib.aero_lbm_step_rect(fptr(packet), self.nx, self.ny, self.nz, self.key, outp)
            else:
                ok = self.lib.aero_lbm_step_rect_scaled(fptr(packet), self.nx, self.ny, self.nz, self.key,
                                                        scale, outp)
        check(ok, "step")

    def flow(self) -> np.ndarray:
        out = np.zeros(self.cells * 4, dtype=np.float32)
        check(self.lib.aero_lbm_get_flow_state_rect(self.nx, self.ny, self.nz, self.key, fptr(out)),
              "get_flow_state_rect")
        return out

    def temperature(self) -> np.ndarray:
        out = np.zeros(self.cells, dtype=np.float32)
        check(self.lib.aero_lbm_get_temperature_state_rect(self.nx, self.ny, self.nz, self.key, fptr(out)),
              "get_temperature_state_rect")
        return out

    def set_temperature(self, t: np.ndarray) -> None:
        t = np.ascontiguousarray(t, dtype=np.float32)
        check(self.lib.aero_lbm_set_temperature_state_rect(self.nx, self.ny, self.nz, self.key, fptr(t)),
              "set_temperature_state_rect")

    def atlas(self, stride: int) -> np.ndarray:
        sx = -(-self.nx // stride)
        sy = -(-self.ny // stride)
        sz = -(-self.nz // stride)
        out = np.zeros(sx * sy * sz * 4, dtype=np.float32)
        check(self.lib.aero_lbm_extract_flow_atlas_rect(self.nx, self.ny, self.nz, self.key, stride,
                                                        fptr(out), out.size), "extract_flow_atlas_rect")
        return out

    def sample(self, x: int, y: int, z: int) -> np.ndarray:
        out = np.zeros(4, dtype=np.float32)
        check(self.lib.aero_lbm_sample_flow_point_rect(self.nx, self.ny, self.nz, self.key, x, y, z, fptr(out)),
              "sample_flow_point_rect")
        return out

    def run_timed(self, steps: int) -> float:
        ms = c_float(0.0)
        check(self.lib.aero_lbm_run_cached_steps_timed(self.nx, self.ny, self.nz, self.key, steps,
                                                       ctypes.byref(ms)), "run_cached_steps_timed")
        return float(ms.value)

    def release(self) -> None:
        self.lib.aero_lbm_release_context(self.key)


class Mesoscale:
    """One D3Q19 mesoscale context of the library (aero_lbm_mesoscale_*)."""

    def __init__(self, cfg: MesoscaleConfig):
        self.lib = load()
        self.cfg = cfg
        key = c_longlong(0)
        check(self.lib.aero_lbm_mesoscale_create_context(ctypes.byref(cfg), ctypes.byref(key)), "mesoscale create")
        self.key = key.value
        self.cells = cfg.nx * cfg.ny * cfg.nz

    def step(self, forcing):
        forcing = np.ascontiguousarray(forcing, dtype=np.float32).reshape(-1)
        if forcing.size != self.cells * 25:
            raise ValueError("forcing must hold cells x 25 floats")
        out = np.zeros(self.cells * 6, dtype=np.float32)
        check(self.lib.aero_lbm_mesoscale_step_context(self.key, ctypes.byref(self.cfg), forcing.ctypes.data_as(_F),
                                                        out.ctypes.data_as(_F)), "mesoscale step")
        return out

    def run_timed(self, forcing, steps):
        forcing = np.ascontiguousarray(forcing, dtype=np.float32).reshape(-1)
        ms = c_float(0.0)
        check(self.lib.aero_lbm_mesoscale_run_steps_timed(self.key, ctypes.byref(self.cfg), forcing.ctypes.data_as(_F),
                                                           steps, ctypes.byref(ms)), "mesoscale timed run")
        return ms.value

    def state(self):
        out = np.zeros(self.cells * 6, dtype=np.float32)
        check(self.lib.aero_lbm_mesoscale_get_state(self.key, out.ctypes.data_as(_F)), "mesoscale state")
        return out

    def release(self):
        if self.key:
            self.lib.aero_lbm_mesoscale_release_context(self.key)
            self.key = 0


def _ptr(a: np.ndarray, kind):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(kind)


class BrickWorld:
    """One brick world of the native brick runtime (aero_lbm_simulation_*)."""

    def __init__(self, world: int = 1, size: int = 16, dx: float = 1.0, dt: float = 0.05,
                 mode: int = SOLVER_D3Q27_FP16_INPLACE):
        self.lib = load()
        key = c_longlong(0)
        self._check(self.lib.aero_lbm_simulation_create_service(ctypes.byref(key)), "create service")
        self.service, self.world, self.size = key.value, world, size
        self._check(self.lib.aero_lbm_simulation_ensure_brick_world_runtime(self.service, world, size, dx, dt),
                    "ensure world")
        self._check(self.lib.aero_lbm_simulation_set_brick_world_solver_mode(self.service, world, mode), "solver mode")

    def _check(self, ok, what):
        if not ok:
            msg = self.lib.aero_lbm_simulation_last_error()
            raise AeroLbmError(f"{what} failed: {msg.decode() if msg else ''}")

    def upload_static(self, coord, obstacle, surface_kind=None, open_face=None, emitter=None, fan_src=None,
                      emitter_src=None):
        n = self.size
        cells = n ** 3
        ob = np.ascontiguousarray(obstacle, dtype=np.uint8).reshape(-1)
        kind = np.zeros(cells, np.uint8) if surface_kind is None else np.ascontiguousarray(surface_kind, np.uint8)
        of = np.zeros(cells, np.uint16) if open_face is None else np.ascontiguousarray(open_face, np.uint16)
        em = np.zeros(cells, np.float32) if emitter is None else np.ascontiguousarray(emitter, np.float32)
        sky = np.zeros(cells * 6, np.uint8)
        fs = None if fan_src is None else np.ascontiguousarray(fan_src, np.uint8)
        es = None if emitter_src is None else np.ascontiguousarray(emitter_src, np.float32)
        self._check(self.lib.aero_lbm_simulation_upload_brick_world_static_brick_with_sources(
            self.service, self.world, n, *coord, _ptr(ob, _U8), _ptr(kind, _U8), _ptr(of, _U16), _ptr(em, _F),
            None if fs is None else _ptr(fs, _U8), None if es is None else _ptr(es, _F), _ptr(sky, _U8),
            _ptr(sky, _U8)), "upload static")

    def upload_dynamic(self, coord, flow, air=None, surface=None, reference=False):
        n = self.size
        fl = np.ascontiguousarray(flow, np.float32).reshape(-1)
        a = np.zeros(n ** 3, np.float32) if air is None else np.ascontiguousarray(air, np.float32)
        s = np.zeros(n ** 3, np.float32) if surface is None else np.ascontiguousarray(surface, np.float32)
        fn = (self.lib.aero_lbm_simulation_upload_brick_world_boundary_reference_brick if reference
              else self.lib.aero_lbm_simulation_upload_brick_world_dynamic_brick)
        self._check(fn(self.service, self.world, n, *coord, _ptr(fl, _F), _ptr(a, _F), _ptr(s, _F)), "upload dynamic")

    def hints(self, coords, exact=True):
        c = np.ascontiguousarray(np.array(coords, np.int32).reshape(-1))
        fn = (self.lib.aero_lbm_simulation_set_brick_world_exact_active_hints if exact
              else self.lib.aero_lbm_simulation_set_brick_world_active_hints)
        self._check(fn(self.service, self.world, self.size, _ptr(c, _I), len(c) // 3), "hints")

    def step(self, steps: int = 1):
        self._check(self.lib.aero_lbm_simulation_step_brick_world_runtime(self.service, self.world, steps), "step")

    def batching(self, enabled: bool):
        """Cached in-place bricks step in shared launches (the default) or one by one."""
        self._check(self.lib.aero_lbm_simulation_set_brick_world_batching(self.service, self.world, int(enabled)),
                    "batching")

    def status(self):
        out = np.zeros(8, np.int32)
        self._check(self.lib.aero_lbm_simulation_get_brick_world_runtime_status(self.service, self.world,
                                                                                 _ptr(out, _I), 8), "status")
        return out

    def dynamic(self, coord):
        n = self.size
        fl = np.zeros(n ** 3 * 4, np.float32)
        a = np.zeros(n ** 3, np.float32)
        s = np.zeros(n ** 3, np.float32)
        self._check(self.lib.aero_lbm_simulation_copy_brick_world_dynamic_brick(
            self.service, self.world, n, *coord, _ptr(fl, _F), _ptr(a, _F), _ptr(s, _F)), "copy dynamic")
        return fl, a, s

    def atlas(self, coord, stride: int = 4):
        sn = (self.size + stride - 1) // stride
        out = np.zeros(sn ** 3 * 4, np.int16)
        self._check(self.lib.aero_lbm_simulation_copy_brick_world_packed_flow_atlas(
            self.service, self.world, self.size, *coord, stride, _ptr(out, _I16), out.size), "atlas")
        return out

    def bricks(self, active=True):
        count_fn = (self.lib.aero_lbm_simulation_get_brick_world_active_brick_count if active
                    else self.lib.aero_lbm_simulation_get_brick_world_resident_brick_count)
        copy_fn = (self.lib.aero_lbm_simulation_copy_brick_world_active_bricks if active
                   else self.lib.aero_lbm_simulation_copy_brick_world_resident_bricks)
        n = count_fn(self.service, self.world)
        out = np.zeros(max(n, 1) * 3, np.int32)
        got = copy_fn(self.service, self.world, _ptr(out, _I), n)
        assert got == n
        return sorted(tuple(int(v) for v in out[3 * i:3 * i + 3]) for i in range(n))

    def release(self):
        self.lib.aero_lbm_simulation_release_service(self.service)


class Batch:
    """Contexts stepped together (aero_lbm_batch_*)."""

    def __init__(self, keys, steps_per_publish=4, stride=4, scales=None, autotune=False):
        self.lib = load()
        self.keys = list(keys)
        k = (c_longlong * len(self.keys))(*self.keys)
        sc = None if scales is None else np.ascontiguousarray(scales, np.float32)
        out = c_longlong(0)
        check(self.lib.aero_lbm_batch_create(k, len(self.keys), steps_per_publish, stride,
                                             None if sc is None else fptr(sc), 1 if autotune else 0,
                                             ctypes.byref(out)), "batch_create")
        self.id, self.steps, self.stride = out.value, steps_per_publish, stride

    def run(self, cycles=1, graph=True):
        ms = c_float(0)
        check(self.lib.aero_lbm_batch_run(self.id, cycles, 1 if graph else 0, ctypes.byref(ms)), "batch_run")
        return ms.value

    def run_per_brick(self, cycles=1):
        ms = c_float(0)
        check(self.lib.aero_lbm_batch_run_per_brick(self.id, cycles, ctypes.byref(ms)), "batch_run_per_brick")
        return ms.value

    def atlas(self, index, n):
        sn = (n + self.stride - 1) // self.stride
        out = np.zeros(sn ** 3 * 4, np.int16)
        check(self.lib.aero_lbm_batch_copy_atlas(self.id, index, _ptr(out, _I16), out.size), "batch_copy_atlas")
        return out

    def info(self):
        return self.lib.aero_lbm_batch_info(self.id).decode()

    def release(self):
        self.lib.aero_lbm_batch_release(self.id)


class Adjoint:
    """Forward / backward pair of the brick-step adjoint (aero_lbm_adjoint_*)."""

    def __init__(self, n, heat_grid, steps, every, nu, beta, alpha, u0, solid, probes, target):
        self.lib = load()
        cfg = AdjointConfig(n, heat_grid, steps, every, nu, beta, alpha, (c_float * 3)(*u0), len(probes))
        sol = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(-1))
        pr = np.ascontiguousarray(np.asarray(probes, np.int32))
        tg = np.ascontiguousarray(np.asarray(target, np.float32).reshape(-1))
        h = c_longlong(0)
        check(self.lib.aero_lbm_adjoint_create(ctypes.byref(cfg), _ptr(sol, _U8), _ptr(pr, _I), fptr(tg),
                                               ctypes.byref(h)), "adjoint_create")
        self.h, self.n, self.m, self.nprobe = h.value, n, heat_grid, len(probes)

    def forward(self, inflow, heat):
        fi = np.ascontiguousarray(inflow, np.float32).reshape(-1)
        hh = np.ascontiguousarray(heat, np.float32).reshape(-1)
        loss = c_double(0)
        u = np.zeros(3 * self.nprobe, np.float32)
        check(self.lib.aero_lbm_adjoint_forward(self.h, fptr(fi), fptr(hh), ctypes.byref(loss), fptr(u)),
              "adjoint_forward")
        return loss.value, u.reshape(-1, 3)

    def backward(self):
        gi = np.zeros(3 * self.n * self.n, np.float32)
        gh = np.zeros(self.m ** 3, np.float32)
        check(self.lib.aero_lbm_adjoint_backward(self.h, fptr(gi), fptr(gh)), "adjoint_backward")
        return gi, gh

    def forward_fp64(self, inflow, heat):
        fi = np.ascontiguousarray(inflow, np.float64).reshape(-1)
        hh = np.ascontiguousarray(heat, np.float64).reshape(-1)
        loss = c_double(0)
        D = POINTER(c_double)
        check(self.lib.aero_lbm_adjoint_forward_fp64(self.h, fi.ctypes.data_as(D), hh.ctypes.data_as(D),
                                                     ctypes.byref(loss)), "adjoint_forward_fp64")
        return loss.value

    def release(self):
        self.lib.aero_lbm_adjoint_release(self.h)


# ---- analysis codec (zfp streams of 3D float grids; native/src/aero_lbm_analysis_codec.cpp)

def compress_float_grid_3d(values: np.ndarray, tolerance: float, lib=None) -> bytes | None:
    """values: float32 grid of shape (nz, ny, nx) (x fastest). The stream's bytes, or
    None on error (reason in last_error()), as the reference's compressFloatGrid3d."""
    lib = lib or load()
    v = np.ascontiguousarray(values, np.float32)
    nz, ny, nx = v.shape
    cap = lib.aero_lbm_analysis_max_compressed_size_3d(nx, ny, nz)
    if cap <= 0:
        return None
    out = np.zeros(cap, np.uint8)
    n = lib.aero_lbm_analysis_compress_float_grid_3d(fptr(v), nx, ny, nz, float(tolerance), _ptr(out, _U8), cap)
    return out[:n].tobytes() if n > 0 else None


def decompress_float_grid_3d(data: bytes, nx: int, ny: int, nz: int, lib=None) -> np.ndarray | None:
    """The grid of shape (nz, ny, nx) a stream decodes to, or None on error."""
    lib = lib or load()
    buf = np.frombuffer(data, np.uint8).copy()
    out = np.zeros((nz, ny, nx), np.float32)
    ok = lib.aero_lbm_analysis_decompress_float_grid_3d(_ptr(buf, _U8), len(buf), nx, ny, nz, fptr(out))
    return out if ok else None


def _device_grid_args(grids, tolerances):
    n = len(grids)
    vals = (c_void_p * n)(*[g.data_ptr() for g in grids])
    dims = np.array([[g.shape[2], g.shape[1], g.shape[0]] for g in grids], np.int32).reshape(-1)
    tols = (c_double * n)(*[float(t) for t in tolerances])
    return n, vals, dims, tols


def compress_device_grids(grids, tolerances, lib=None, repeats: int = 0):
    """Device float32 tensors of shape (nz, ny, nx), coded in one set of launches.
    Returns the list of streams (bytes); with repeats > 0 also the device time per
    repeat in ms (the launches run that many times back to back)."""
    import torch
    lib = lib or load()
    n, vals, dims, tols = _device_grid_args(grids, tolerances)
    caps = [lib.aero_lbm_analysis_max_compressed_size_3d(int(dims[3 * i]), int(dims[3 * i + 1]), int(dims[3 * i + 2]))
            for i in range(n)]
    offs = np.cumsum([0] + caps[:-1]).astype(np.int64)
    out = torch.zeros(int(sum(caps)), dtype=torch.uint8, device=grids[0].device)
    c_offs = (c_longlong * n)(*offs.tolist())
    c_caps = (c_longlong * n)(*caps)
    sizes = (c_longlong * n)()
    torch.cuda.synchronize()
    check(lib.aero_lbm_analysis_compress_float_grids_3d_device(n, vals, _ptr(dims, _I), tols, c_void_p(out.data_ptr()),
                                                               c_offs, c_caps, sizes), "compress_float_grids_3d_device")
    host = out.cpu().numpy()
    streams = [host[offs[i]:offs[i] + sizes[i]].tobytes() for i in range(n)]
    if repeats <= 0:
        return streams
    ms = c_float(0)
    check(lib.aero_lbm_analysis_compress_grids_timed(n, vals, _ptr(dims, _I), tols, c_void_p(out.data_ptr()), c_offs,
                                                     c_caps, repeats, ctypes.byref(ms)), "compress_grids_timed")
    return streams, ms.value
