"""Analysis codec on the CPU (native/src/aero_lbm_analysis_codec.cpp, the zfp 1.0.1
coder it drives, and the JNI decompression native at
native/src/aero_lbm_simulation_bridge.cpp:8929-8974).

The oracle (oracle/codec_oracle.c) is pinned three ways: byte for byte against the
golden streams the reference codec wrote (tests/golden/codec_streams.json), against
the reference codec compiled from its sources where that build exists, and by
properties the format fixes without the reference (the tolerance bound of
fixed-accuracy mode, the header's fields, one bit per all-zero block). The library's
host decoder and the decompression native need no GPU and are checked here too."""
import base64
import json
import math
import os

import numpy as np
import pytest

import aero_lbm as A
import codec_cases as CC
from jni_fake import JNI_ABORT, FakeJVM, Natives

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "codec_streams.json")


def golden_cases():
    with open(GOLDEN) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        shape = tuple(c["shape"])
        v = np.frombuffer(base64.b64decode(c["input"]), np.float32).reshape(shape)
        d = np.frombuffer(base64.b64decode(c["decoded"]), np.float32).reshape(shape)
        yield c["kind"], shape, c["tolerance"], v, base64.b64decode(c["stream"]), d


@pytest.fixture(scope="module")
def oracle():
    return CC.oracle()


@pytest.fixture(scope="module")
def lib():
    return A.load()


def test_oracle_reproduces_golden_streams(oracle):
    for kind, shape, tol, v, stream, decoded in golden_cases():
        assert oracle.compress(v, tol) == stream, (kind, shape, tol)
        got = oracle.decompress(stream, shape[2], shape[1], shape[0])
        assert got.tobytes() == decoded.tobytes(), (kind, shape, tol)


def test_golden_inputs_are_the_seeded_grids():
    for kind, shape, tol, v, _, _ in golden_cases():
        seed = [c for c in json.load(open(GOLDEN))["cases"] if c["kind"] == kind and tuple(c["shape"]) == shape][0]["seed"]
        assert CC.grid(kind, shape, seed).tobytes() == v.tobytes()


def test_oracle_matches_reference_build_on_sweep(oracle):
    ref = CC.reference()
    if ref is None:
        pytest.skip("oracle/_ref/libcodec_ref.so not built (reference tree absent)")
    for shape, kind, tol, seed in CC.sweep():
        v = CC.grid(kind, shape, seed)
        a, b = oracle.compress(v, tol), ref.compress(v, tol)
        assert a == b, (shape, kind, tol)
        nz, ny, nx = shape
        da, db = oracle.decompress(a, nx, ny, nz), ref.decompress(a, nx, ny, nz)
        assert da.tobytes() == db.tobytes(), (shape, kind, tol)


def test_fixed_accuracy_error_bound(oracle):
    # fixed-accuracy mode keeps every finite value within the tolerance, up to the
    # rounding of the decoded value to float, while the tolerance stays well above
    # the 30-bit block-floating-point step (magnitudes here are below 8)
    for kind in ("smooth", "mixed", "velocity", "pressure"):
        for tol in (1e-1, CC.VELOCITY_TOL, CC.PRESSURE_TOL, 1e-6):
            for shape in ((13, 9, 7), (16, 16, 16)):
                v = CC.grid(kind, shape, 5)
                s = oracle.compress(v, tol)
                d = oracle.decompress(s, shape[2], shape[1], shape[0])
                err = np.abs(d.astype(np.float64) - v.astype(np.float64))
                bound = tol + np.spacing(np.maximum(np.abs(v), np.abs(d))).astype(np.float64)
                assert np.all(err <= bound), (kind, tol, shape, float(np.max(err - bound)))


def test_header_fields_and_zero_blocks(oracle):
    for shape, tol in (((7, 6, 5), 1e-3), ((1, 1, 1), 0.5), ((40, 3, 9), 1e-9)):
        nz, ny, nx = shape
        s = oracle.compress(np.zeros(shape, np.float32), tol)
        bits = np.unpackbits(np.frombuffer(s, np.uint8), bitorder="little")
        field = lambda a, n: int(sum(int(bits[a + i]) << i for i in range(n)))  # noqa: E731
        assert bytes(field(8 * i, 8) for i in range(3)) == b"zfp" and field(24, 8) == 5
        meta = field(32, 52)
        assert meta & 3 == 2 and (meta >> 2) & 3 == 2  # float, three dimensions
        assert ((meta >> 4) & 0xFFFF) + 1 == nx and ((meta >> 20) & 0xFFFF) + 1 == ny and ((meta >> 36) & 0xFFFF) + 1 == nz
        minexp = math.frexp(tol)[1] - 1
        assert field(84, 12) == minexp + 1074 + 2177
        blocks = ((nx + 3) // 4) * ((ny + 3) // 4) * ((nz + 3) // 4)
        assert len(s) == ((96 + blocks + 63) // 64) * 8  # one zero bit per block, word aligned
        assert not bits[96:].any()


def test_long_header_for_extreme_tolerances(oracle):
    v = CC.grid("smooth", (4, 4, 4))
    for tol in (5e-324, 1e300):
        s = oracle.compress(v, tol)
        bits = np.unpackbits(np.frombuffer(s, np.uint8), bitorder="little")
        assert bits[84:96].all()  # 0xfff in the short field announces the 64-bit mode
        d = oracle.decompress(s, 4, 4, 4)
        assert d is not None


def test_library_max_size_matches_zfp_bound(lib):
    for shape in ((1, 1, 1), (4, 4, 4), (64, 64, 64), (7, 33, 2)):
        nz, ny, nx = shape
        assert lib.aero_lbm_analysis_max_compressed_size_3d(nx, ny, nz) == CC.max_bytes(nx, ny, nz)
    assert lib.aero_lbm_analysis_max_compressed_size_3d(0, 4, 4) == 0


def test_host_decoder_matches_golden_and_oracle(lib, oracle):
    for kind, shape, tol, v, stream, decoded in golden_cases():
        nz, ny, nx = shape
        got = A.decompress_float_grid_3d(stream, nx, ny, nz, lib)
        assert got is not None and got.tobytes() == decoded.tobytes(), (kind, shape, tol)
    for shape, kind, tol, seed in CC.sweep()[::3]:
        v = CC.grid(kind, shape, seed)
        s = oracle.compress(v, tol)
        nz, ny, nx = shape
        assert A.decompress_float_grid_3d(s, nx, ny, nz, lib).tobytes() == oracle.decompress(s, nx, ny, nz).tobytes()


def test_host_decoder_other_zfp_modes(lib):
    # fixed-rate and fixed-precision streams written by zfp itself (reference build)
    ref = CC.reference()
    if ref is None:
        pytest.skip("oracle/_ref/libcodec_ref.so not built (reference tree absent)")
    for shape in ((13, 9, 7), (16, 16, 16)):
        nz, ny, nx = shape
        v = CC.grid("smooth", shape, 3)
        for mode, param in ((0, 2.0), (0, 8.0), (0, 23.0), (1, 6), (1, 20), (2, 1e-3)):
            cap = CC.max_bytes(nx, ny, nz) * 2
            out = np.zeros(cap, np.uint8)
            n = ref.mode(v.ctypes.data, nx, ny, nz, mode, float(param), out.ctypes.data, cap)
            assert n > 0
            s = out[:n].tobytes()
            want = ref.decompress(s, nx, ny, nz)
            got = A.decompress_float_grid_3d(s, nx, ny, nz, lib)
            assert got is not None and got.tobytes() == want.tobytes(), (shape, mode, param)


def test_host_decoder_errors(lib):
    kind, shape, tol, v, stream, decoded = next(golden_cases())
    nz, ny, nx = shape
    assert A.decompress_float_grid_3d(stream, nx + 1, ny, nz, lib) is None
    assert A.last_error() == "zfp_header_mismatch"
    assert A.decompress_float_grid_3d(b"zfq" + stream[3:], nx, ny, nz, lib) is None
    assert A.last_error() == "zfp_read_header_failed"
    assert A.decompress_float_grid_3d(stream[:40], nx, ny, nz, lib) is None
    assert A.last_error() == "zfp_decompress_failed"
    assert A.decompress_float_grid_3d(b"", nx, ny, nz, lib) is None
    assert A.last_error() == "invalid_input"


def test_jni_decompress_native(lib):
    jvm = FakeJVM()
    nat = Natives(lib, jvm)
    kind, shape, tol, v, stream, decoded = next(golden_cases())
    nz, ny, nx = shape
    comp = jvm.array("byte", np.frombuffer(stream, np.int8))
    out = jvm.array("float", np.zeros(v.size, np.float32))
    assert nat("nativeDecompressFloatGrid3d", comp, nx, ny, nz, out) == 1
    assert jvm.value(out).tobytes() == decoded.reshape(-1).tobytes()
    assert jvm.modes(out) == [0]
    # a wrong length is refused before any array is pinned; a bad stream releases
    # the output with JNI_ABORT
    short = jvm.array("float", np.zeros(v.size - 1, np.float32))
    assert nat("nativeDecompressFloatGrid3d", comp, nx, ny, nz, short) == 0
    assert jvm.modes(short) == []
    bad = jvm.array("byte", np.frombuffer(stream[:40], np.int8))
    out2 = jvm.array("float", np.full(v.size, 7.0, np.float32))
    assert nat("nativeDecompressFloatGrid3d", bad, nx, ny, nz, out2) == 0
    assert jvm.modes(out2) == [JNI_ABORT] and np.all(jvm.value(out2) == 7.0)
    assert nat("nativeDecompressFloatGrid3d", None, nx, ny, nz, out2) == 0
