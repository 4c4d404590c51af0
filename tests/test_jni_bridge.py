"""The JNI bridge (aerodynamics4mc-fabric_amd/jni/aero_lbm_jni_bridge.cpp) compiles
and defines every native method the mod's Java classes declare
(src/main/java/com/aerodynamics4mc/runtime/NativeSimulationBridge.java:1239-1656,
MesoscaleNativeBridge.java:129-165). No JDK is installed here, so the check
compiles against tests/jni_stub/jni.h, a declaration-only stand-in, and reads the
exported names from the object file."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "aerodynamics4mc-fabric_amd", "jni", "aero_lbm_jni_bridge.cpp")

NSB = ["nativeCreateService", "nativeReleaseService", "nativeSetFocus", "nativeSubmitWorldDeltas",
       "nativeEnsureBrickWorldRuntime", "nativeSetBrickWorldSolverMode", "nativeSetBrickWorldActiveHints",
       "nativeSetBrickWorldExactActiveHints", "nativeGetBrickWorldRuntimeStatus", "nativeStepBrickWorldRuntime",
       "nativeGetBrickWorldResidentBrickCoords", "nativeGetBrickWorldActiveBrickCoords",
       "nativeUploadBrickWorldStaticBrick", "nativeUploadBrickWorldStaticBrickWithSources",
       "nativeQueueBrickWorldStaticBrickUpload", "nativeCopyBrickWorldDynamicBrick",
       "nativeCopyBrickWorldPackedFlowAtlas", "nativeUploadBrickWorldDynamicBrick",
       "nativeUploadBrickWorldBoundaryReferenceBrick", "nativeRuntimeInfo", "nativeLastError"]
MNB = ["nativeDeriveTransport", "nativeCreateContext", "nativeStepContext", "nativeReleaseContext"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("nm") is None, reason="needs g++ and nm")
def test_bridge_compiles_and_exports_every_native_method(tmp_path):
    obj = tmp_path / "bridge.o"
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-c", SRC, "-o", str(obj),
                           "-I", os.path.join(ROOT, "tests", "jni_stub")])
    syms = subprocess.check_output(["nm", "--defined-only", str(obj)], text=True)
    for m in NSB:
        assert f" T Java_com_aerodynamics4mc_runtime_NativeSimulationBridge_{m}\n" in syms, m
    for m in MNB:
        assert f" T Java_com_aerodynamics4mc_runtime_MesoscaleNativeBridge_{m}\n" in syms, m
