"""C-ABI surface checks that need no GPU: the library loads, exports every
symbol declared in include/*.h, and keeps the reference's struct layouts."""
import ctypes
import glob
import os
import re

import pytest

import aero_lbm as A
from conftest import ROOT, gpu_available

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        names.update(re.findall(r"AERO_LBM_CAPI_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text))
        names.update(re.findall(r"JNIEXPORT\s+\w+\s+JNICALL\s+(\w+)\s*\(", text))
    return sorted(names)


def test_headers_declare_symbols():
    syms = declared_symbols()
    assert "aero_lbm_init_rect" in syms and "aero_lbm_step_rect" in syms
    assert len(syms) > 30


def test_library_exports_every_declared_symbol():
    lib = A.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_layouts_match_reference():
    assert ctypes.sizeof(A.BoundaryFace) == 32
    assert ctypes.sizeof(A.BenchmarkConfig) == 320
    assert ctypes.sizeof(A.TimingSnapshot) == 72
    assert ctypes.sizeof(A.MesoscaleConfig) == 96
    assert ctypes.sizeof(A.MesoscaleTransport) == 20
    assert A.BenchmarkConfig.x_min.offset == 96


def test_benchmark_config_roundtrip_without_gpu():
    lib = A.load()
    cfg = A.default_config()
    assert cfg.abi_version == 1 and cfg.struct_size == 320 and cfg.enabled == 0
    assert cfg.flags == (A.FLAG_DISABLE_FAN_FORCING | A.FLAG_DISABLE_FAN_NOISE | A.FLAG_DISABLE_SPONGE
                         | A.FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND | A.FLAG_DISABLE_SGS)
    tg = A.BenchmarkConfig()
    assert lib.aero_lbm_benchmark_default_preset_config(A.PRESET_TAYLOR_GREEN_3D, ctypes.byref(tg)) == 1
    assert tg.enabled == 1 and tg.x_min.hydrodynamic_kind == A.HYDRO_PERIODIC
    bad = A.BenchmarkConfig()
    bad.abi_version = 2
    assert lib.aero_lbm_benchmark_set_config(ctypes.byref(bad)) == 0
    tg.reynolds_number = float("nan")
    assert lib.aero_lbm_benchmark_set_config(ctypes.byref(tg)) == 1
    back = A.BenchmarkConfig()
    lib.aero_lbm_benchmark_get_config(ctypes.byref(back))
    assert back.reynolds_number == pytest.approx(100.0)  # sanitised like the reference
    lib.aero_lbm_benchmark_reset_config()


@pytest.mark.parametrize("three_d", [True, False])
def test_heated_cavity_preset_matches_reference_preset(three_d):
    """The library's heated-cavity preset is the reference's (jni.cpp:307-338, 384-400),
    as restated by the workload helper, byte for byte."""
    import workloads as W
    lib = A.load()
    got = A.BenchmarkConfig()
    preset = A.PRESET_HEATED_CAVITY_3D if three_d else A.PRESET_HEATED_CAVITY_2D
    assert lib.aero_lbm_benchmark_default_preset_config(preset, ctypes.byref(got)) == 1
    want = W.heated_cavity_config(three_d)
    assert bytes(got) == bytes(want)


def test_cylinder_preset_matches_reference_preset():
    import workloads as W
    lib = A.load()
    got = A.BenchmarkConfig()
    assert lib.aero_lbm_benchmark_default_preset_config(A.PRESET_CYLINDER_CROSSFLOW_2D, ctypes.byref(got)) == 1
    assert bytes(got) == bytes(W.cylinder_config())


def test_overrides_validation():
    lib = A.load()
    ov = A.default_overrides()
    assert ov.struct_size == ctypes.sizeof(A.SolverOverrides)
    assert ov.kernel_variant == A.KERNEL_FAST
    ov.kernel_variant = 7
    assert lib.aero_lbm_set_solver_overrides(ctypes.byref(ov)) == 0
    assert "kernel_variant" in A.last_error()


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    lib = A.load()
    assert lib.aero_lbm_init_rect(8, 8, 8, 11, 4) == 0
    assert "no HIP device" in A.last_error()
    assert lib.aero_lbm_runtime_info().decode().startswith(("hip|unavailable", "uninitialized"))


def test_mesoscale_transport_and_config_validation(oracle_mod):
    """Host-side mesoscale entry points: transport identical to the reference's own
    hydro/thermal core (oracle/_ref), reference defaults, ABI-version check."""
    lib = A.load()
    cfg = A.MesoscaleConfig()
    lib.aero_lbm_mesoscale_default_config(ctypes.byref(cfg))
    assert (cfg.abi_version, cfg.struct_size, cfg.nx, cfg.ny, cfg.nz) == (3, 96, 32, 4, 32)
    assert cfg.dx_m == 64.0 and cfg.dt_s == 5.0
    ref = oracle_mod.load_ref()
    for dx, dt, nu, pr in ((64.0, 5.0, 1.5e-5, 0.71), (32.0, 2.5, 0.06, 0.9), (8.0, 0.1, 2.0, 0.71)):
        cfg.dx_m, cfg.dt_s, cfg.molecular_nu_m2_s, cfg.prandtl_air = dx, dt, nu, pr
        tr = A.MesoscaleTransport()
        assert lib.aero_lbm_mesoscale_derive_transport(ctypes.byref(cfg), ctypes.byref(tr)) == 1
        if ref is not None:
            out = (ctypes.c_float * 5)()
            ref.ref_derive_transport(cfg.nx, cfg.ny, cfg.nz, dx, dt, nu, pr, cfg.turbulent_prandtl, out)
            got = (tr.velocity_scale_m_s_per_lattice, tr.nu_molecular_lattice, tr.alpha_molecular_lattice,
                   tr.tau_shear_molecular, tr.tau_thermal_molecular)
            assert got == tuple(out)
    cfg.abi_version = 2
    assert lib.aero_lbm_mesoscale_derive_transport(ctypes.byref(cfg), ctypes.byref(tr)) == 0
    key = ctypes.c_longlong(0)
    assert lib.aero_lbm_mesoscale_create_context(ctypes.byref(cfg), ctypes.byref(key)) == 0
    assert lib.aero_lbm_mesoscale_set_tau_floor(0.4) == 0
