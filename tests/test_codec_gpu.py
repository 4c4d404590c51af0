"""Analysis codec on the GPU: the device encoder's streams byte for byte against the
oracle (oracle/codec_oracle.c), the golden streams of the reference codec
(native/src/aero_lbm_analysis_codec.cpp:19-99) and, where it was built, the
reference codec itself; the batched device entry point against single calls; the
JNI compression native (native/src/aero_lbm_simulation_bridge.cpp:8880-8927)."""
import base64
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import aero_lbm as A
import codec_cases as CC
from jni_fake import JNI_ABORT, FakeJVM, Natives

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "codec_streams.json")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    return A.load()


@pytest.fixture(scope="module")
def oracle():
    return CC.oracle()


def test_encoder_matches_golden_streams(lib):
    with open(GOLDEN) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        shape = tuple(c["shape"])
        log(f"golden {c['kind']} {shape} tol={c['tolerance']}")
        v = np.frombuffer(base64.b64decode(c["input"]), np.float32).reshape(shape)
        got = A.compress_float_grid_3d(v, c["tolerance"], lib)
        assert got == base64.b64decode(c["stream"]), (c["kind"], shape)


def test_encoder_matches_oracle_on_sweep(lib, oracle):
    ref = CC.reference()
    for i, (shape, kind, tol, seed) in enumerate(CC.sweep()):
        if i % 48 == 0:
            log(f"sweep case {i}: {shape} {kind} tol={tol}")
        v = CC.grid(kind, shape, seed)
        got = A.compress_float_grid_3d(v, tol, lib)
        assert got is not None, A.last_error()
        assert got == oracle.compress(v, tol), (shape, kind, tol)
        if ref is not None and i % 7 == 0:
            assert got == ref.compress(v, tol), (shape, kind, tol)


def test_reference_payload_grids_batched(lib, oracle):
    # four 64^3 grids as the reference's analysis payload codes them (vx, vy, vz at
    # the velocity tolerance, pressure at the pressure tolerance), plus a 62^3 grid
    # with partial blocks and an unaligned row length, in one set of launches
    import torch
    shapes = [(64, 64, 64)] * 4 + [(62, 62, 62)]
    kinds = ["velocity", "velocity", "velocity", "pressure", "smooth"]
    tols = [CC.VELOCITY_TOL] * 3 + [CC.PRESSURE_TOL, 1e-4]
    grids = [CC.grid(k, s, 40 + i) for i, (k, s) in enumerate(zip(kinds, shapes))]
    dev = [torch.from_numpy(g).cuda() for g in grids]
    log("batched 4 x 64^3 + 62^3")
    streams = A.compress_device_grids(dev, tols, lib)
    for g, t, s in zip(grids, tols, streams):
        assert s == oracle.compress(g, t)
        assert s == A.compress_float_grid_3d(g, t, lib)
        d = A.decompress_float_grid_3d(s, g.shape[2], g.shape[1], g.shape[0], lib)
        assert float(np.max(np.abs(d.astype(np.float64) - g))) <= t


def test_large_grids_against_oracle(lib, oracle):
    for shape, kind, tol in (((128, 128, 128), "velocity", CC.VELOCITY_TOL), ((130, 97, 131), "mixed", 1e-3),
                             ((256, 256, 256), "pressure", CC.PRESSURE_TOL)):
        log(f"large {shape} {kind}")
        v = CC.grid(kind, shape, 11)
        got = A.compress_float_grid_3d(v, tol, lib)
        assert got == oracle.compress(v, tol), shape


def test_capacity_shortfall_is_an_error_and_writes_nothing_past_it(lib):
    import torch
    v = torch.from_numpy(CC.grid("mixed", (32, 32, 32), 3)).cuda()
    need = len(A.compress_device_grids([v], [1e-6], lib)[0])
    cap = (need // 2) & ~7
    out = torch.full((cap + 4096,), 0xA5, dtype=torch.uint8, device="cuda")
    vals = (ctypes.c_void_p * 1)(v.data_ptr())
    dims = np.array([32, 32, 32], np.int32)
    tols = (ctypes.c_double * 1)(1e-6)
    offs = (ctypes.c_longlong * 1)(0)
    caps = (ctypes.c_longlong * 1)(cap)
    sizes = (ctypes.c_longlong * 1)()
    torch.cuda.synchronize()
    ok = lib.aero_lbm_analysis_compress_float_grids_3d_device(1, vals, dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                              tols, ctypes.c_void_p(out.data_ptr()), offs, caps, sizes)
    assert ok == 0 and "capacity" in A.last_error()
    assert bool((out[cap:] == 0xA5).all())


def test_errors_match_reference_strings(lib):
    v = CC.grid("smooth", (4, 4, 4))
    assert A.compress_float_grid_3d(v, 0.0, lib) is None and A.last_error() == "invalid_tolerance"
    assert A.compress_float_grid_3d(v, float("nan"), lib) is None and A.last_error() == "invalid_tolerance"
    big = np.zeros((1, 1, 65537), np.float32)
    assert A.compress_float_grid_3d(big, 1e-3, lib) is None and A.last_error() == "zfp_write_header_failed"


def test_jni_compress_native_roundtrip(lib, oracle):
    jvm = FakeJVM()
    nat = Natives(lib, jvm)
    shape = (20, 16, 18)
    v = CC.grid("velocity", shape, 9)
    nz, ny, nx = shape
    arr = jvm.array("float", v.reshape(-1))
    h = nat("nativeCompressFloatGrid3d", arr, nx, ny, nz, CC.VELOCITY_TOL)
    assert h
    stream = jvm.value(h).tobytes()
    assert stream == oracle.compress(v, CC.VELOCITY_TOL)
    assert jvm.modes(arr) == [JNI_ABORT]  # the input goes back unwritten
    out = jvm.array("float", np.zeros(v.size, np.float32))
    assert nat("nativeDecompressFloatGrid3d", h, nx, ny, nz, out) == 1
    assert jvm.value(out).tobytes() == oracle.decompress(stream, nx, ny, nz).reshape(-1).tobytes()
    # wrong length or a bad tolerance: null, as the reference returns
    assert not nat("nativeCompressFloatGrid3d", arr, nx + 1, ny, nz, CC.VELOCITY_TOL)
    assert not nat("nativeCompressFloatGrid3d", arr, nx, ny, nz, -1.0)
