"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker. The product (libaero_lbm.so and its Python binding) never
imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_core.so")

_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    for p, real in (("o32_", c_float), ("o64_", c_double)):
        R = POINTER(real)
        getattr(lib, p + "oracle_create").restype = c_void_p
        getattr(lib, p + "oracle_create").argtypes = [c_int] * 5 + [c_void_p, c_void_p]
        getattr(lib, p + "oracle_destroy").argtypes = [c_void_p]
        getattr(lib, p + "oracle_set_temperature").argtypes = [c_void_p, POINTER(c_float)]
        getattr(lib, p + "oracle_get_temperature").argtypes = [c_void_p, POINTER(c_float)]
        getattr(lib, p + "oracle_step").argtypes = [c_void_p, POINTER(c_float), POINTER(c_float), c_float]
        getattr(lib, p + "oracle_step").restype = c_int
        getattr(lib, p + "oracle_run").argtypes = [c_void_p, POINTER(c_float), c_int]
        getattr(lib, p + "oracle_run").restype = c_int
        getattr(lib, p + "oracle_extract_atlas").argtypes = [c_void_p, c_int, POINTER(c_float)]
        getattr(lib, p + "oracle_get_f").argtypes = [c_void_p, R]
        getattr(lib, p + "oracle_set_f").argtypes = [c_void_p, R]
        getattr(lib, p + "oracle_get_f_post").argtypes = [c_void_p, R]
        getattr(lib, p + "oracle_get_velocity").argtypes = [c_void_p, R]
        getattr(lib, p + "oracle_get_temperature_r").argtypes = [c_void_p, R]
        getattr(lib, p + "oracle_get_params").argtypes = [c_void_p, R]
        getattr(lib, p + "oracle_get_force").argtypes = [c_void_p, POINTER(c_float)]
        getattr(lib, p + "oracle_collide_cell").argtypes = [c_void_p, R, real, R]
        getattr(lib, p + "oracle_feq").argtypes = [c_int, real, real, real, real]
        getattr(lib, p + "oracle_feq").restype = real
        getattr(lib, p + "oracle_weights").argtypes = [R]
        getattr(lib, p + "oracle_set_shell").argtypes = [c_void_p, c_void_p, POINTER(c_float), c_int, c_float]
        getattr(lib, p + "oracle_patch_cells").argtypes = [c_void_p, c_int, POINTER(c_int), POINTER(ctypes.c_uint8)]
        getattr(lib, p + "oracle_shift").argtypes = [c_void_p, c_int, c_int, c_int]
        getattr(lib, p + "oracle_shift").restype = c_int
        getattr(lib, p + "oracle_exchange_halo").argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_int]
        getattr(lib, p + "oracle_exchange_halo").restype = c_int
    lib.oracle_sanitize_config.argtypes = [c_void_p]
    lib.oracle_set_threads.argtypes = [c_int]
    lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    _lib = lib
    return lib


def set_threads(n: int) -> None:
    """Threads for the oracle's per-cell loops; results are identical for any count."""
    load().oracle_set_threads(n)


def _fp(a):
    return a.ctypes.data_as(POINTER(c_float))


class Oracle:
    """CPU restatement of the reference step. precision: 32 or 64."""

    def __init__(self, nx, ny, nz, cfg, overrides=None, precision=32, in_ch=11, out_ch=4):
        self.lib = load()
        self.p = "o32_" if precision == 32 else "o64_"
        self.real = np.float32 if precision == 32 else np.float64
        self.nx, self.ny, self.nz = nx, ny, nz
        self.cells = nx * ny * nz
        self.out_ch = out_ch
        ovp = ctypes.cast(ctypes.byref(overrides), c_void_p) if overrides is not None else None
        self.h = getattr(self.lib, self.p + "oracle_create")(nx, ny, nz, in_ch, out_ch,
                                                             ctypes.cast(ctypes.byref(cfg), c_void_p), ovp)
        if not self.h:
            raise ValueError("oracle does not restate this configuration")

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.p + "oracle_destroy")(self.h)
            self.h = None

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    def set_temperature(self, t):
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        self._f("oracle_set_temperature")(self.h, _fp(t))

    def temperature(self):
        out = np.zeros(self.cells, dtype=np.float32)
        self._f("oracle_get_temperature")(self.h, _fp(out))
        return out

    def shift(self, dx, dy, dz):
        """The reference's context shift (every field: new(x) = old(x + d))."""
        if not self._f("oracle_shift")(self.h, dx, dy, dz):
            raise ValueError("shift exceeds the grid")

    def exchange_halo(self, other, offset, halo_layers):
        """The reference's halo exchange with `other` lying at `offset` (-1..1 per axis)."""
        if not self._f("oracle_exchange_halo")(self.h, other.h, *offset, halo_layers):
            raise ValueError("invalid halo exchange")

    def patch_cells(self, cells, solid):
        """The reference's static-cell patch on the oracle state (obstacle flags set in
        order; every changed cell restarts from zero if solid, the rest equilibrium if fluid)."""
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        solid = np.ascontiguousarray(solid, dtype=np.uint8)
        self._f("oracle_patch_cells")(self.h, len(cells), cells.ctypes.data_as(POINTER(c_int)),
                                      solid.ctypes.data_as(POINTER(ctypes.c_uint8)))

    def temperature_r(self):
        out = np.zeros(self.cells, dtype=self.real)
        self._f("oracle_get_temperature_r")(self.h, out.ctypes.data_as(POINTER(np.ctypeslib.as_ctypes_type(self.real))))
        return out

    def step(self, packet=None, out=None, scale=1.0):
        pk = None if packet is None else _fp(np.ascontiguousarray(packet, dtype=np.float32))
        ok = self._f("oracle_step")(self.h, pk, None if out is None else _fp(out), scale)
        if not ok:
            raise RuntimeError("oracle step failed")

    def run(self, packet, steps):
        pk = np.ascontiguousarray(packet, dtype=np.float32)
        if not self._f("oracle_run")(self.h, _fp(pk), steps):
            raise RuntimeError("oracle run failed")

    def atlas(self, stride=1):
        sx, sy, sz = (-(-n // stride) for n in (self.nx, self.ny, self.nz))
        out = np.zeros(sx * sy * sz * 4, dtype=np.float32)
        self._f("oracle_extract_atlas")(self.h, stride, _fp(out))
        return out

    def flow(self):
        return self.atlas(1)

    def populations(self):
        out = np.zeros(27 * self.cells, dtype=self.real)
        self._f("oracle_get_f")(self.h, out.ctypes.data_as(POINTER(np.ctypeslib.as_ctypes_type(self.real))))
        return out.reshape(27, self.cells)

    def post_collision(self):
        """Post-collision populations of the last step, before streaming, [27, cells]."""
        out = np.zeros(27 * self.cells, dtype=self.real)
        self._f("oracle_get_f_post")(self.h, out.ctypes.data_as(POINTER(np.ctypeslib.as_ctypes_type(self.real))))
        return out.reshape(27, self.cells)

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=self.real).reshape(-1)
        self._f("oracle_set_f")(self.h, f.ctypes.data_as(POINTER(np.ctypeslib.as_ctypes_type(self.real))))

    def velocity(self):
        out = np.zeros(3 * self.cells, dtype=self.real)
        self._f("oracle_get_velocity")(self.h, out.ctypes.data_as(POINTER(np.ctypeslib.as_ctypes_type(self.real))))
        return out.reshape(self.cells, 3)

    def force(self):
        """Last step's drag integral (fx, fy, fz); zero unless the cylinder preset runs."""
        out = np.zeros(3, dtype=np.float32)
        self._f("oracle_get_force")(self.h, _fp(out))
        return out

    def params(self):
        out = np.zeros(6, dtype=self.real)
        self._f("oracle_get_params")(self.h, out.ctypes.data_as(POINTER(np.ctypeslib.as_ctypes_type(self.real))))
        return dict(zip(("tau_shear", "tau_normal", "nu", "alpha", "beta", "sgs_c2"), out.tolist()))

    def set_shell(self, layer, targets, layers, relax):
        """Boundary-reference shell: layer per cell (uint8, 0 = none), decoded targets
        (cells x 4: u_x, u_y, u_z, p), depth and outer relaxation; layers <= 0 removes it."""
        self._layer = np.ascontiguousarray(layer, dtype=np.uint8).reshape(-1)
        self._targets = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        self._f("oracle_set_shell")(self.h, self._layer.ctypes.data_as(c_void_p), _fp(self._targets), int(layers),
                                    float(relax))

    def collide_cell(self, f, temperature=0.0):
        """Post-collision populations of one fluid cell (context must be 1x1x1)."""
        ct = np.ctypeslib.as_ctypes_type(self.real)
        fin = np.ascontiguousarray(f, dtype=self.real)
        out = np.zeros(27, dtype=self.real)
        self._f("oracle_collide_cell")(self.h, fin.ctypes.data_as(POINTER(ct)), temperature,
                                       out.ctypes.data_as(POINTER(ct)))
        return out


def load_ref():
    """The reference's own hydro/thermal core (oracle/_ref), or None when absent."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB)
    lib.ref_d3q27_tables.argtypes = [POINTER(c_int)] * 4 + [POINTER(c_float)]
    lib.ref_feq_d3q27.restype = c_float
    lib.ref_feq_d3q27.argtypes = [c_int] + [c_float] * 4
    lib.ref_feq_d3q7.restype = c_float
    lib.ref_feq_d3q7.argtypes = [c_int] + [c_float] * 4
    lib.ref_derive_transport.argtypes = [c_int] * 3 + [c_float] * 5 + [POINTER(c_float)]
    lib.ref_boussinesq_beta.restype = c_float
    lib.ref_boussinesq_beta.argtypes = [c_float] * 5
    lib.ref_lattice_diffusivity.restype = c_float
    lib.ref_lattice_diffusivity.argtypes = [c_float] * 3
    return lib


class MesoOracle:
    """CPU restatement of the reference's D3Q19 mesoscale step (precision 32 or 64)."""

    def __init__(self, nx, ny, nz, dx=64.0, dt=5.0, nu=1.5e-5, pr=0.71, prt=0.85, tau_min=0.58, precision=32):
        self.lib = load()
        self.p = "m32_" if precision == 32 else "m64_"
        self.real = np.float32 if precision == 32 else np.float64
        ct = np.ctypeslib.as_ctypes_type(self.real)
        f = getattr(self.lib, self.p + "meso_create")
        f.restype = c_void_p
        f.argtypes = [c_int] * 3 + [c_float] * 6
        getattr(self.lib, self.p + "meso_destroy").argtypes = [c_void_p]
        getattr(self.lib, self.p + "meso_step").argtypes = [c_void_p, POINTER(c_float), POINTER(c_float)]
        getattr(self.lib, self.p + "meso_step").restype = c_int
        getattr(self.lib, self.p + "meso_get_f").argtypes = [c_void_p, POINTER(ct), POINTER(ct)]
        getattr(self.lib, self.p + "meso_set_f").argtypes = [c_void_p, POINTER(ct), POINTER(ct)]
        getattr(self.lib, self.p + "meso_get_fields").argtypes = [c_void_p] + [POINTER(ct)] * 3
        getattr(self.lib, self.p + "meso_transport").argtypes = [c_void_p, POINTER(ct)]
        self.cells = nx * ny * nz
        self.h = f(nx, ny, nz, dx, dt, nu, pr, prt, tau_min)
        if not self.h:
            raise ValueError("invalid mesoscale domain")

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.p + "meso_destroy")(self.h)
            self.h = None

    def step(self, forcing):
        forcing = np.ascontiguousarray(forcing, dtype=np.float32).reshape(-1)
        out = np.zeros(self.cells * 6, dtype=np.float32)
        if not getattr(self.lib, self.p + "meso_step")(self.h, _fp(forcing), _fp(out)):
            raise RuntimeError("mesoscale oracle step failed")
        return out

    def distributions(self):
        ct = np.ctypeslib.as_ctypes_type(self.real)
        f = np.zeros(self.cells * 19, dtype=self.real)
        g = np.zeros(self.cells * 19, dtype=self.real)
        getattr(self.lib, self.p + "meso_get_f")(self.h, f.ctypes.data_as(POINTER(ct)), g.ctypes.data_as(POINTER(ct)))
        return f.reshape(self.cells, 19), g.reshape(self.cells, 19)

    def set_distributions(self, f, g):
        ct = np.ctypeslib.as_ctypes_type(self.real)
        f = np.ascontiguousarray(f, dtype=self.real).reshape(-1)
        g = np.ascontiguousarray(g, dtype=self.real).reshape(-1)
        getattr(self.lib, self.p + "meso_set_f")(self.h, f.ctypes.data_as(POINTER(ct)), g.ctypes.data_as(POINTER(ct)))

    def fields(self):
        """(ambient, surface, deep) per cell."""
        ct = np.ctypeslib.as_ctypes_type(self.real)
        a, s, d = (np.zeros(self.cells, dtype=self.real) for _ in range(3))
        getattr(self.lib, self.p + "meso_get_fields")(self.h, *(v.ctypes.data_as(POINTER(ct)) for v in (a, s, d)))
        return a, s, d

    def transport(self):
        """(velocity scale m/s per lattice unit, tau_shear, tau_thermal) before clamping."""
        ct = np.ctypeslib.as_ctypes_type(self.real)
        out = np.zeros(3, dtype=self.real)
        getattr(self.lib, self.p + "meso_transport")(self.h, out.ctypes.data_as(POINTER(ct)))
        return tuple(out.tolist())


_dual = None


def dual_directional(n, m, steps, nu, beta, alpha, u0, solid, probes, target, inflow, heat, d_inflow, d_heat):
    """FP64 dual-number loss and directional derivative dL/dp . d of the adjoint
    model (oracle/libadjoint_dual.so)."""
    global _dual
    if _dual is None:
        path = os.path.join(HERE, "libadjoint_dual.so")
        if not os.path.exists(path):
            build()
        _dual = ctypes.CDLL(path)
        D = POINTER(c_double)
        _dual.ad_dual_directional.argtypes = [c_int, c_int, c_int, c_double, c_float, c_float, POINTER(c_float),
                                              c_void_p, c_int, POINTER(c_int), D, D, D, D, D, D, D]
        _dual.ad_dual_directional.restype = c_int
    f64 = lambda a: np.ascontiguousarray(a, np.float64).reshape(-1)
    D = POINTER(c_double)
    sol = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(-1))
    pr = np.ascontiguousarray(np.asarray(probes, np.int32))
    u0a = np.ascontiguousarray(u0, np.float32)
    arrs = [f64(target), f64(inflow), f64(heat), f64(d_inflow), f64(d_heat)]
    loss, dl = c_double(0), c_double(0)
    ok = _dual.ad_dual_directional(n, m, steps, nu, beta, alpha, u0a.ctypes.data_as(POINTER(c_float)),
                                   sol.ctypes.data_as(c_void_p), len(pr), pr.ctypes.data_as(POINTER(c_int)),
                                   *[a.ctypes.data_as(D) for a in arrs], ctypes.byref(loss), ctypes.byref(dl))
    if not ok:
        raise RuntimeError("dual directional derivative failed")
    return loss.value, dl.value
