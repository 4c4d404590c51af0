"""The JNI bridge linked into libaero_lbm.so, called the way a JVM calls it: through a
JNIEnv whose interface function table is a stand-in JVM (tests/jni_fake.py). Covers
the natives the mod's ClientL2Solver calls (NativeSimulationBridge.java:1239-1656)
and the four mesoscale natives (MesoscaleNativeBridge.java:129-165), their results
against the C ABI and bricks_oracle, and the release mode of every Java array: input
arrays go back with JNI_ABORT, output arrays with mode 0 when the call succeeded and
JNI_ABORT when it failed (native/src/aero_lbm_simulation_bridge.cpp:7717-7727).
No GPU: stepping returns false with an error here (tests/test_jni_bridge_gpu.py
steps through the same natives on the GPU)."""
import subprocess

import numpy as np
import pytest

import aero_lbm as A
import bricks_oracle as B
from jni_fake import JNI_ABORT, MNB, NSB, FakeJVM, Natives, SIGNATURES

N_BRICK = 8


@pytest.fixture()
def jni():
    lib = A.load()
    jvm = FakeJVM()
    return jvm, Natives(lib, jvm)


def test_library_exports_every_native_method():
    out = subprocess.check_output(["nm", "-D", "--defined-only", A.LIB_PATH], text=True)
    for name in SIGNATURES:
        assert f" T {name}\n" in out, name
    for extra in ("nativeSetFocus", "nativeSetBrickWorldActiveHints", "nativeGetBrickWorldResidentBrickCoords",
                  "nativeGetBrickWorldActiveBrickCoords", "nativeUploadBrickWorldStaticBrick",
                  "nativeQueueBrickWorldStaticBrickUpload"):
        assert f" T {NSB}{extra}\n" in out, extra


def static_arrays(jvm, n, seed=3):
    rng = np.random.default_rng(seed)
    cells = n ** 3
    ob = (rng.random(cells) < 0.15).astype(np.int8)
    kind = np.zeros(cells, np.int8)
    kind[7] = 3
    arrays = dict(obstacle=jvm.array("byte", ob), kind=jvm.array("byte", kind),
                  open=jvm.array("short", np.zeros(cells)), emitter=jvm.array("float", np.zeros(cells)),
                  fan_src=jvm.array("byte", np.zeros(cells)), emitter_src=jvm.array("float", np.zeros(cells)),
                  sky=jvm.array("byte", np.zeros(cells * 6)), direct=jvm.array("byte", np.zeros(cells * 6)))
    return ob.astype(np.uint8), kind.astype(np.uint8), arrays


def open_world(jvm, N, world=21):
    svc = N("nativeCreateService")
    assert svc != 0
    assert N("nativeEnsureBrickWorldRuntime", svc, world, N_BRICK, 0.5, 0.05) == 1
    assert N("nativeSetBrickWorldSolverMode", svc, world, A.SOLVER_D3Q27_FP16_INPLACE) == 1
    coords = jvm.array("int", [0, 0, 0])
    assert N("nativeSetBrickWorldExactActiveHints", svc, world, N_BRICK, coords, 1) == 1
    assert jvm.modes(coords) == [JNI_ABORT]
    return svc


def test_world_uploads_copies_status_atlas_and_release_modes(jni):
    jvm, N = jni
    n = N_BRICK
    svc = open_world(jvm, N)
    ob, kind, st = static_arrays(jvm, n)
    order = ("obstacle", "kind", "open", "emitter", "fan_src", "emitter_src", "sky", "direct")
    assert N("nativeUploadBrickWorldStaticBrickWithSources", svc, 21, n, 0, 0, 0, *[st[k] for k in order]) == 1
    for k in order:
        assert jvm.modes(st[k]) == [JNI_ABORT], k        # inputs are never written back
    rng = np.random.default_rng(9)
    flow = rng.uniform(-0.05, 0.05, n ** 3 * 4).astype(np.float32)
    air = rng.uniform(-0.5, 0.5, n ** 3).astype(np.float32)
    surf = rng.uniform(0.0, 1.0, n ** 3).astype(np.float32)
    hf, ha, hs = jvm.array("float", flow), jvm.array("float", air), jvm.array("float", surf)
    assert N("nativeUploadBrickWorldDynamicBrick", svc, 21, n, 0, 0, 0, hf, ha, hs) == 1
    assert jvm.modes(hf) == jvm.modes(ha) == jvm.modes(hs) == [JNI_ABORT]
    bref = [jvm.array("float", flow * 0.5), jvm.array("float", air), jvm.array("float", surf)]
    assert N("nativeUploadBrickWorldBoundaryReferenceBrick", svc, 21, n, 0, 0, 0, *bref) == 1

    # copy back: output arrays released with mode 0 and holding the constrained state
    of, oa, osf = jvm.array("float", np.zeros(n ** 3 * 4)), jvm.array("float", np.zeros(n ** 3)), \
        jvm.array("float", np.zeros(n ** 3))
    assert N("nativeCopyBrickWorldDynamicBrick", svc, 21, n, 0, 0, 0, of, oa, osf) == 1
    assert jvm.modes(of) == jvm.modes(oa) == jvm.modes(osf) == [0]
    # with source arrays the upload derives surface kinds, open faces and emitters from
    # them (simulation_bridge.cpp:1987-2097)
    kd, opd, emd = B.derive_from_sources(n, ob, np.zeros(n ** 3, np.uint8), np.zeros(n ** 3, np.float32))
    ef, ea, es = flow.copy(), air.copy(), surf.copy()
    B.apply_constraints(ob, opd, kd, emd, ef, ea, es)
    assert np.array_equal(jvm.value(of), ef) and np.array_equal(jvm.value(oa), ea)
    assert np.array_equal(jvm.value(osf), es)

    # an output array of the wrong length: refused before any array is pinned
    bad = jvm.array("float", np.full(10, -7.0))
    oa2, os2 = jvm.array("float", np.zeros(n ** 3)), jvm.array("float", np.zeros(n ** 3))
    assert N("nativeCopyBrickWorldDynamicBrick", svc, 21, n, 0, 0, 0, bad, oa2, os2) == 0
    assert jvm.modes(bad) == jvm.modes(oa2) == []
    assert np.all(jvm.value(bad) == -7.0)
    # a copy the C ABI refuses (a brick without state): pinned outputs go back with JNI_ABORT
    of3, oa3, os3 = jvm.array("float", np.full(n ** 3 * 4, 5.0)), jvm.array("float", np.zeros(n ** 3)), \
        jvm.array("float", np.zeros(n ** 3))
    assert N("nativeCopyBrickWorldDynamicBrick", svc, 21, n, 4, 4, 4, of3, oa3, os3) == 0
    assert jvm.modes(of3) == jvm.modes(oa3) == jvm.modes(os3) == [JNI_ABORT]
    assert np.all(jvm.value(of3) == 5.0)

    status = jvm.array("int", np.full(8, -1))
    assert N("nativeGetBrickWorldRuntimeStatus", svc, 21, status) == 1
    assert jvm.modes(status) == [0]
    s = jvm.value(status)
    assert s[0] == n and s[1] == 1 and s[3] == 1

    for stride in (1, 4):
        sn = -(-n // stride)
        atlas = jvm.array("short", np.zeros(sn ** 3 * 4))
        assert N("nativeCopyBrickWorldPackedFlowAtlas", svc, 21, n, 0, 0, 0, stride, atlas) == 1
        assert jvm.modes(atlas) == [0]
        assert np.array_equal(jvm.value(atlas), B.packed_atlas(n, stride, ef, velocity_scale=0.5 / 0.05))

    # world deltas: 8 ints + 4 floats each (NativeSimulationBridge.java:30-31)
    ints = jvm.array("int", [2, 3, 3, 3, 5, 0, 0, 0])
    floats = jvm.array("float", [0.0, 0.0, 0.0, 0.0])
    assert N("nativeSubmitWorldDeltas", svc, ints, floats) == 1
    assert jvm.modes(ints) == jvm.modes(floats) == [JNI_ABORT]
    ragged = jvm.array("int", [2, 3, 3, 3, 5, 0, 0])
    assert N("nativeSubmitWorldDeltas", svc, ragged, floats) == 0

    info = N("nativeRuntimeInfo")
    text = jvm.value(info)
    assert text.startswith("service=1 worlds=1") and "bricks=1" in text, text
    N("nativeReleaseService", svc)


def test_packed_flow_atlas_store(jni):
    """nativeSetPackedFlowAtlas / nativePollPackedFlowAtlas (simulation_bridge.cpp:
    6711-6762, 8789-8840): the atlas comes back whole under its key; the input array
    goes back with JNI_ABORT, the polled array with mode 0 whatever the result."""
    jvm, N = jni
    svc = N("nativeCreateService")
    vals = np.arange(-40, 40, dtype=np.int16) * 97
    src = jvm.array("short", vals)
    assert N("nativeSetPackedFlowAtlas", svc, 5, src) == 1
    assert jvm.modes(src) == [JNI_ABORT]
    out = jvm.array("short", np.zeros(vals.size))
    assert N("nativePollPackedFlowAtlas", svc, 5, out) == 1
    assert jvm.modes(out) == [0] and np.array_equal(jvm.value(out), vals)
    short = jvm.array("short", np.full(vals.size - 1, 3))
    assert N("nativePollPackedFlowAtlas", svc, 5, short) == 0
    assert jvm.modes(short) == [0] and np.all(jvm.value(short) == 3)
    assert b"size mismatch" in A.load().aero_lbm_simulation_last_error()
    assert N("nativePollPackedFlowAtlas", svc, 6, jvm.array("short", np.zeros(vals.size))) == 0
    assert b"missing atlas" in A.load().aero_lbm_simulation_last_error()
    assert N("nativeSetPackedFlowAtlas", svc + 1, 5, src) == 0
    assert b"missing service" in A.load().aero_lbm_simulation_last_error()
    # the C ABI view of the same store
    lib = A.load()
    got = np.zeros(vals.size, np.int16)
    assert lib.aero_lbm_simulation_poll_packed_flow_atlas(svc, 5, got.ctypes.data_as(A._I16), got.size) == 1
    assert np.array_equal(got, vals)
    assert lib.aero_lbm_simulation_set_packed_flow_atlas(svc, 5, None, 4) == 0
    assert b"invalid atlas buffer" in lib.aero_lbm_simulation_last_error()
    assert "atlases=1" in lib.aero_lbm_simulation_runtime_info().decode()
    assert lib.aero_lbm_context_compact_initialized(12345) == 0
    N("nativeReleaseService", svc)


def test_step_without_gpu_fails_with_a_message(jni):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: stepping succeeds (tests/test_jni_bridge_gpu.py)")
    jvm, N = jni
    svc = open_world(jvm, N, world=22)
    _, _, st = static_arrays(jvm, N_BRICK)
    order = ("obstacle", "kind", "open", "emitter", "fan_src", "emitter_src", "sky", "direct")
    assert N("nativeUploadBrickWorldStaticBrickWithSources", svc, 22, N_BRICK, 0, 0, 0, *[st[k] for k in order]) == 1
    assert N("nativeStepBrickWorldRuntime", svc, 22, 1) == 0
    msg = jvm.value(N("nativeLastError"))
    assert isinstance(msg, str) and msg
    N("nativeReleaseService", svc)


def test_mesoscale_natives(jni):
    jvm, N = jni
    dims, dx, dt, nu, pr, prt = (8, 4, 6), 64.0, 5.0, 1.5e-5, 0.71, 0.85
    out = N("nativeDeriveTransport", *dims, dx, dt, nu, pr, prt)
    v = jvm.value(out).astype(np.float64)
    nu_l = nu * dt / dx ** 2
    alpha_l = nu / pr * dt / dx ** 2
    want = [dx / dt, nu_l, alpha_l, 0.5 + 3.0 * nu_l, 0.5 + alpha_l / 0.25]
    assert np.allclose(v, want, rtol=1e-6, atol=0.0)
    cells = int(np.prod(dims))
    forcing = jvm.array("float", np.zeros(cells * 25))
    state = jvm.array("float", np.full(cells * 6, 3.0))
    bad_state = jvm.array("float", np.zeros(cells * 6 - 1))
    key = N("nativeCreateContext", *dims, dx, dt, nu, pr, prt)
    # a state array of the wrong length is refused before any work (mesoscale.cpp:1566-1570)
    assert N("nativeStepContext", key, *dims, dx, dt, nu, pr, prt, forcing, bad_state) == 0
    assert jvm.modes(bad_state) == []
    import torch
    if not torch.cuda.is_available():
        # without a device the step fails and the output array is not written back
        assert N("nativeStepContext", key, *dims, dx, dt, nu, pr, prt, forcing, state) == 0
        assert jvm.modes(state) == [JNI_ABORT] and np.all(jvm.value(state) == 3.0)
    N("nativeReleaseContext", key)
