"""Inputs and checker bindings for the analysis codec tests (test infrastructure).

Grids are float32 arrays of shape (nz, ny, nx), x fastest, as the reference's
compressFloatGrid3d takes them (src/main/java/com/aerodynamics4mc/runtime/
NativeSimulationBridge.java:1128-1147). The checkers are the plain-C restatement in
oracle/liboracle.so and, when it was built from the reference's sources,
oracle/_ref/libcodec_ref.so."""
import ctypes
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libcodec_ref.so")
_C = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_longlong]
_D = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]

# the reference's analysis tolerances, Java floats widened to double
# (src/main/java/com/aerodynamics4mc/runtime/AeroServerRuntime.java:139-140, 160-165)
VELOCITY_TOL = float(np.float32(0.02))
PRESSURE_TOL = float(np.float32(0.0005))
PATCH_TOLS = [float(np.float32(t)) for t in (0.0005, 0.01, 0.1, 0.001, 0.00025)]
SPECIALS = np.array([np.inf, -np.inf, np.nan, 1e-40, -1e-45, 3.4e38, -3.4e38, 0.0, 1.0, -0.0], np.float32)


class Codec:
    """compress(v, tol) -> bytes | None; decompress(b, nx, ny, nz) -> array | None"""

    def __init__(self, path, compress, decompress):
        self.lib = ctypes.CDLL(path)
        self.c = getattr(self.lib, compress)
        self.c.restype, self.c.argtypes = ctypes.c_longlong, _C
        self.d = getattr(self.lib, decompress)
        self.d.restype, self.d.argtypes = ctypes.c_int, _D

    def compress(self, v, tol):
        v = np.ascontiguousarray(v, np.float32)
        nz, ny, nx = v.shape
        cap = max_bytes(nx, ny, nz)
        out = np.zeros(cap, np.uint8)
        n = self.c(v.ctypes.data, nx, ny, nz, float(tol), out.ctypes.data, cap)
        return out[:n].tobytes() if n > 0 else None

    def decompress(self, data, nx, ny, nz):
        buf = np.frombuffer(data, np.uint8).copy()
        out = np.zeros((nz, ny, nx), np.float32)
        return out if self.d(buf.ctypes.data, len(buf), nx, ny, nz, out.ctypes.data) else None


def oracle():
    return Codec(ORACLE, "oracle_codec_compress_3f", "oracle_codec_decompress_3f")


def reference():
    """The reference codec built from its sources, or None where it was not built."""
    if not os.path.exists(REF):
        return None
    r = Codec(REF, "ref_codec_compress_3f", "ref_codec_decompress_3f")
    r.mode = r.lib.ref_zfp_compress_3f_mode
    r.mode.restype = ctypes.c_longlong
    r.mode.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                       ctypes.c_void_p, ctypes.c_longlong]
    return r


def max_bytes(nx, ny, nz):
    blocks = ((nx + 3) // 4) * ((ny + 3) // 4) * ((nz + 3) // 4)
    return ((148 + blocks * 2120 + 63) // 64) * 8


def header_bits(tol):
    e = math.frexp(tol)[1] - 1
    return 96 if (e != -1074 and e <= 843) else 148


def grid(kind, shape, seed=0):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    if kind == "smooth":
        v = np.sin(0.3 * x) * np.cos(0.2 * y) + 0.1 * z
    elif kind == "noise":
        v = rng.standard_normal(shape) * 10.0 ** rng.integers(-8, 8)
    elif kind == "zero":
        v = np.zeros(shape)
    elif kind == "mixed":
        v = rng.standard_normal(shape)
        v[rng.random(shape) < 0.5] = 0
    elif kind == "special":
        v = rng.choice(SPECIALS, size=shape)
    elif kind == "subnormal":
        v = rng.standard_normal(shape) * 1e-39
    elif kind == "velocity":  # a residual velocity field: smooth swirl plus small noise
        v = 0.05 * np.sin(2 * np.pi * x / 17) * np.cos(2 * np.pi * z / 23) + 0.003 * rng.standard_normal(shape)
    elif kind == "pressure":
        v = 1e-3 * np.cos(2 * np.pi * (x + y) / 29) + 1e-4 * rng.standard_normal(shape)
    else:
        raise ValueError(kind)
    return np.ascontiguousarray(v, np.float32)


SHAPES = [(4, 4, 4), (1, 1, 1), (5, 6, 7), (13, 9, 7), (16, 16, 16), (3, 17, 2), (33, 32, 31)]
KINDS = ["smooth", "noise", "zero", "mixed", "special", "subnormal"]
TOLS = [1e-1, VELOCITY_TOL, PRESSURE_TOL, 1e-7, 1e-30, 1e-40, 5e-324, 1e300]


def sweep():
    """(shape, kind, tolerance, seed) cases covering partial blocks, all-zero and
    special-value blocks, and both header forms"""
    out = []
    for si, shape in enumerate(SHAPES):
        for ki, kind in enumerate(KINDS):
            for ti, tol in enumerate(TOLS):
                out.append((shape, kind, tol, 1000 * si + 10 * ki + ti))
    return out
