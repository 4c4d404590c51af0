"""Wind-tunnel solver C ABI (aero_solver_*) without a GPU: struct layouts, defaults
and the argument checks that run before any device work."""
import ctypes

import numpy as np

import aero_lbm as A


def test_struct_layouts():
    # x86-64 layouts of native/include/aero_solver_capi.h
    assert ctypes.sizeof(A.GridDesc) == 20
    assert ctypes.sizeof(A.BoundaryDesc) == 28
    assert ctypes.sizeof(A.StepInput) == 72 and A.StepInput.grid.offset == 16 and A.StepInput.steps.offset == 64
    assert ctypes.sizeof(A.StepOutput) == 16


def test_defaults():
    lib = A.load()
    g, b = A.GridDesc(), A.BoundaryDesc()
    lib.aero_solver_default_grid(ctypes.byref(g))
    lib.aero_solver_default_boundary(ctypes.byref(b))
    assert (g.nx, g.ny, g.nz) == (64, 64, 64) and g.dx == 1.0 and np.float32(g.dt) == np.float32(0.05)
    assert b.mode == A.SOLVER_WIND_TUNNEL and b.inlet_vx == 5.0 and b.inlet_vy == 0.0 and b.outlet_pressure == 0.0
    assert np.float32(b.density) == np.float32(1.225) and np.float32(b.viscosity) == np.float32(1.5e-5)


def test_argument_errors():
    lib = A.load()
    h = ctypes.c_longlong(0)
    assert lib.aero_solver_create(0, 4, 4, 1.0, 0.05, ctypes.byref(h)) == 0
    assert b"invalid grid dimensions" in lib.aero_solver_last_error()
    assert lib.aero_solver_create(4, 4, 4, float("nan"), 0.05, ctypes.byref(h)) == 0
    assert b"invalid grid dx/dt" in lib.aero_solver_last_error()
    assert lib.aero_solver_create(4, 4, 4, 1.0, -1.0, ctypes.byref(h)) == 0
    assert b"invalid grid dx/dt" in lib.aero_solver_last_error()
    assert lib.aero_solver_create(2000, 2000, 1000, 1.0, 0.05, ctypes.byref(h)) == 0
    assert b"overflows int" in lib.aero_solver_last_error()
    assert lib.aero_solver_create_with_grid(None, ctypes.byref(h)) == 0
    assert b"missing grid" in lib.aero_solver_last_error()
    assert lib.aero_solver_run_wind_tunnel(None, None) == 0
    assert b"missing step input or output" in lib.aero_solver_last_error()
    assert lib.aero_solver_advance_wind_tunnel(987654, None, 1) == 0
    assert b"invalid solver handle" in lib.aero_solver_last_error()
    lib.aero_solver_destroy(987654)  # unknown handle: no effect

