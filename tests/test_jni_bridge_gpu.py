"""The JNI bridge on the GPU: a brick world stepped through the Java natives
(NativeSimulationBridge.java:1239-1656) equals the same world driven through the C
ABI, bit for bit, in the in-place and the classic solver modes; the mesoscale natives
(MesoscaleNativeBridge.java:129-165) equal the C ABI's mesoscale context. Output
arrays come back with mode 0 on success."""
import numpy as np
import pytest

import aero_lbm as A
import workloads as W
from jni_fake import FakeJVM, Natives

pytestmark = pytest.mark.gpu


def static_brick(n, seed):
    ob = W.solid_boxes(n, n, n, 0.08, seed=seed, edge=(2, 4), keep_clear_x=2).reshape(-1).astype(np.uint8)
    kind = np.zeros(n ** 3, np.uint8)
    kind[(n // 2) * n * n + 5 * n + 5] = 33 if not ob[(n // 2) * n * n + 5 * n + 5] else 0
    return ob, kind


@pytest.mark.parametrize("mode", [A.SOLVER_D3Q27_FP16_INPLACE, A.SOLVER_CLASSIC_D3Q27])
def test_world_through_jni_equals_c_abi(mode):
    lib = A.load()
    lib.aero_lbm_benchmark_reset_config()
    n = 16
    ob, kind = static_brick(n, 17)
    rng = np.random.default_rng(2)
    flow = np.tile(np.array([0.04, 0.0, 0.0, 0.0], np.float32), n ** 3)
    flow[0::4] += rng.uniform(-1e-3, 1e-3, n ** 3).astype(np.float32)
    air = (0.02 * rng.uniform(-1, 1, n ** 3)).astype(np.float32)
    surf = np.zeros(n ** 3, np.float32)
    cells = n ** 3

    jvm = FakeJVM()
    N = Natives(lib, jvm)
    svc = N("nativeCreateService")
    assert N("nativeEnsureBrickWorldRuntime", svc, 31, n, 1.0, 0.05) == 1
    assert N("nativeSetBrickWorldSolverMode", svc, 31, mode) == 1
    assert N("nativeSetBrickWorldExactActiveHints", svc, 31, n, jvm.array("int", [0, 0, 0]), 1) == 1
    statics = [jvm.array("byte", ob), jvm.array("byte", kind), jvm.array("short", np.zeros(cells)),
               jvm.array("float", np.zeros(cells)), jvm.array("byte", np.zeros(cells * 6)),
               jvm.array("byte", np.zeros(cells * 6))]
    assert N("nativeUploadBrickWorldStaticBrick", svc, 31, n, 0, 0, 0, *statics) == 1
    assert N("nativeUploadBrickWorldDynamicBrick", svc, 31, n, 0, 0, 0, jvm.array("float", flow),
             jvm.array("float", air), jvm.array("float", surf)) == 1

    w = A.BrickWorld(world=32, size=n, dx=1.0, dt=0.05, mode=mode)
    w.hints([(0, 0, 0)], exact=True)
    w.upload_static((0, 0, 0), ob, surface_kind=kind)
    w.upload_dynamic((0, 0, 0), flow, air, surf)
    try:
        for step in range(5):
            assert N("nativeStepBrickWorldRuntime", svc, 31, 1) == 1, jvm.value(N("nativeLastError"))
            w.step(1)
            of, oa, osf = jvm.array("float", np.zeros(cells * 4)), jvm.array("float", np.zeros(cells)), \
                jvm.array("float", np.zeros(cells))
            assert N("nativeCopyBrickWorldDynamicBrick", svc, 31, n, 0, 0, 0, of, oa, osf) == 1
            assert jvm.modes(of) == [0]
            gf, ga, _ = w.dynamic((0, 0, 0))
            assert np.array_equal(jvm.value(of), gf), step
            assert np.array_equal(jvm.value(oa), ga), step
        atlas = jvm.array("short", np.zeros(4 ** 3 * 4))
        assert N("nativeCopyBrickWorldPackedFlowAtlas", svc, 31, n, 0, 0, 0, 4, atlas) == 1
        assert np.array_equal(jvm.value(atlas), w.atlas((0, 0, 0), 4))
        status = jvm.array("int", np.zeros(8))
        assert N("nativeGetBrickWorldRuntimeStatus", svc, 31, status) == 1
        assert jvm.value(status)[3] == 1
    finally:
        w.release()
        N("nativeReleaseService", svc)


def test_mesoscale_natives_equal_c_abi():
    lib = A.load()
    jvm = FakeJVM()
    N = Natives(lib, jvm)
    nx, ny, nz = 48, 12, 40
    cfg = W.meso_config(nx, ny, nz)
    F = W.meso_forcing(nx, ny, nz).reshape(-1).astype(np.float32)
    args = (nx, ny, nz, cfg.dx_m, cfg.dt_s, cfg.molecular_nu_m2_s, cfg.prandtl_air, cfg.turbulent_prandtl)
    key = N("nativeCreateContext", *args)
    assert key != 0
    ref = A.Mesoscale(cfg)
    try:
        for step in range(4):
            state = jvm.array("float", np.zeros(nx * ny * nz * 6))
            assert N("nativeStepContext", key, *args, jvm.array("float", F), state) == 1
            assert jvm.modes(state) == [0]
            want = ref.step(F)
            assert np.array_equal(jvm.value(state), np.asarray(want).reshape(-1)), step
    finally:
        N("nativeReleaseContext", key)
        ref.release()
