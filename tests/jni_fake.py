"""A stand-in for a JVM's side of JNI, for calling the library's Java_* exports from
Python (test infrastructure).

It builds a JNIEnv: a pointer to a 235-slot interface function table whose slots at
the specification's indices (NewStringUTF 167, GetArrayLength 171, NewByteArray 176,
NewIntArray 179, NewFloatArray 181, Get<Byte|Short|Int|Float>ArrayElements
184/186/187/189, Release<...>ArrayElements 192/194/195/197, GetByteArrayRegion 200,
Set<Byte|Int|Float>ArrayRegion 208/211/213, ExceptionCheck 228) are Python
callbacks. Java arrays are numpy arrays behind opaque handles; Get*Elements
hands out a copy, and Release*Elements copies it back only for modes 0 and
JNI_COMMIT, as a copying JVM does, so a test sees exactly which arrays a native
method wrote back. Every release is recorded with its mode."""
import ctypes
from ctypes import CFUNCTYPE, POINTER, c_char_p, c_int32, c_int64, c_uint8, c_void_p

import numpy as np

JNI_COMMIT, JNI_ABORT = 1, 2
TABLE_SLOTS = 235
KINDS = {"byte": (np.int8, 184, 192), "short": (np.int16, 186, 194), "int": (np.int32, 187, 195),
         "float": (np.float32, 189, 197)}


class FakeJVM:
    def __init__(self):
        self.table = (c_void_p * TABLE_SLOTS)()
        self.objects = {}          # handle -> numpy array or str
        self.copies = {}           # element pointer -> (handle, numpy copy)
        self.releases = []         # (handle, mode)
        self._next = 0x1000
        self._keep = []
        self._install()
        self.env_struct = c_void_p(ctypes.addressof(self.table))
        self.env = ctypes.cast(ctypes.pointer(self.env_struct), c_void_p)

    # ---- objects
    def _new(self, obj):
        h = self._next
        self._next += 0x10
        self.objects[h] = obj
        return h

    def array(self, kind, values):
        dt = KINDS[kind][0]
        return self._new(np.array(values, dtype=dt).reshape(-1).copy())

    def value(self, handle):
        return self.objects[handle]

    def modes(self, handle):
        return [m for h, m in self.releases if h == handle]

    # ---- table
    def _slot(self, index, restype, *argtypes):
        def deco(fn):
            f = CFUNCTYPE(restype, *argtypes)(fn)
            self._keep.append(f)
            self.table[index] = ctypes.cast(f, c_void_p)
            return fn
        return deco

    def _install(self):
        @self._slot(167, c_void_p, c_void_p, c_char_p)
        def new_string_utf(env, s):
            return self._new(s.decode("utf-8"))

        @self._slot(171, c_int32, c_void_p, c_void_p)
        def get_array_length(env, a):
            return int(self.objects[a].size)

        @self._slot(176, c_void_p, c_void_p, c_int32)
        def new_byte_array(env, n):
            return self._new(np.zeros(n, np.int8))

        @self._slot(200, None, c_void_p, c_void_p, c_int32, c_int32, c_void_p)
        def get_byte_region(env, a, start, n, buf):
            ctypes.memmove(buf, self.objects[a][start:start + n].ctypes.data, n)

        @self._slot(228, c_uint8, c_void_p)
        def exception_check(env):
            return 0

        @self._slot(179, c_void_p, c_void_p, c_int32)
        def new_int_array(env, n):
            return self._new(np.zeros(n, np.int32))

        @self._slot(181, c_void_p, c_void_p, c_int32)
        def new_float_array(env, n):
            return self._new(np.zeros(n, np.float32))

        for kind, (dt, get_i, rel_i) in KINDS.items():
            self._install_elements(dt, get_i, rel_i)

        def set_region(dt):
            def fn(env, a, start, n, buf):
                src = np.ctypeslib.as_array(ctypes.cast(buf, POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,))
                self.objects[a][start:start + n] = src
            return fn

        self._slot(208, None, c_void_p, c_void_p, c_int32, c_int32, c_void_p)(set_region(np.int8))
        self._slot(211, None, c_void_p, c_void_p, c_int32, c_int32, c_void_p)(set_region(np.int32))
        self._slot(213, None, c_void_p, c_void_p, c_int32, c_int32, c_void_p)(set_region(np.float32))

    def _install_elements(self, dt, get_i, rel_i):
        @self._slot(get_i, c_void_p, c_void_p, c_void_p, POINTER(c_uint8))
        def get_elements(env, a, is_copy):
            buf = self.objects[a].copy()
            ptr = buf.ctypes.data
            self.copies[ptr] = (a, buf)
            if is_copy:
                is_copy[0] = 1
            return ptr

        @self._slot(rel_i, None, c_void_p, c_void_p, c_void_p, c_int32)
        def release_elements(env, a, elems, mode):
            h, buf = self.copies[elems]
            assert h == a
            if mode in (0, JNI_COMMIT):
                self.objects[a][:] = buf
            if mode != JNI_COMMIT:
                del self.copies[elems]
            self.releases.append((a, mode))


NSB = "Java_com_aerodynamics4mc_runtime_NativeSimulationBridge_"
MNB = "Java_com_aerodynamics4mc_runtime_MesoscaleNativeBridge_"

# (return type, argument types after (JNIEnv*, jclass)); object arguments are handles
J = c_void_p
SIGNATURES = {
    NSB + "nativeCreateService": (c_int64, []),
    NSB + "nativeReleaseService": (None, [c_int64]),
    NSB + "nativeEnsureBrickWorldRuntime": (c_uint8, [c_int64, c_int64, c_int32, ctypes.c_float, ctypes.c_float]),
    NSB + "nativeSetBrickWorldSolverMode": (c_uint8, [c_int64, c_int64, c_int32]),
    NSB + "nativeSetBrickWorldExactActiveHints": (c_uint8, [c_int64, c_int64, c_int32, J, c_int32]),
    NSB + "nativeSubmitWorldDeltas": (c_uint8, [c_int64, J, J]),
    NSB + "nativeUploadBrickWorldStaticBrickWithSources": (c_uint8, [c_int64, c_int64] + [c_int32] * 4 + [J] * 8),
    NSB + "nativeUploadBrickWorldStaticBrick": (c_uint8, [c_int64, c_int64] + [c_int32] * 4 + [J] * 6),
    NSB + "nativeUploadBrickWorldDynamicBrick": (c_uint8, [c_int64, c_int64] + [c_int32] * 4 + [J] * 3),
    NSB + "nativeUploadBrickWorldBoundaryReferenceBrick": (c_uint8, [c_int64, c_int64] + [c_int32] * 4 + [J] * 3),
    NSB + "nativeCopyBrickWorldDynamicBrick": (c_uint8, [c_int64, c_int64] + [c_int32] * 4 + [J] * 3),
    NSB + "nativeStepBrickWorldRuntime": (c_uint8, [c_int64, c_int64, c_int32]),
    NSB + "nativeCopyBrickWorldPackedFlowAtlas": (c_uint8, [c_int64, c_int64] + [c_int32] * 5 + [J]),
    NSB + "nativeGetBrickWorldRuntimeStatus": (c_uint8, [c_int64, c_int64, J]),
    NSB + "nativeSetPackedFlowAtlas": (c_uint8, [c_int64, c_int64, J]),
    NSB + "nativePollPackedFlowAtlas": (c_uint8, [c_int64, c_int64, J]),
    NSB + "nativeCompressFloatGrid3d": (J, [J, c_int32, c_int32, c_int32, ctypes.c_double]),
    NSB + "nativeDecompressFloatGrid3d": (c_uint8, [J, c_int32, c_int32, c_int32, J]),
    NSB + "nativeRuntimeInfo": (J, []),
    NSB + "nativeLastError": (J, []),
    MNB + "nativeDeriveTransport": (J, [c_int32] * 3 + [ctypes.c_float] * 5),
    MNB + "nativeCreateContext": (c_int64, [c_int32] * 3 + [ctypes.c_float] * 5),
    MNB + "nativeStepContext": (c_uint8, [c_int64] + [c_int32] * 3 + [ctypes.c_float] * 5 + [J, J]),
    MNB + "nativeReleaseContext": (None, [c_int64]),
}


class Natives:
    """The library's Java_* exports called with a FakeJVM's JNIEnv."""

    def __init__(self, lib, jvm):
        self.lib, self.jvm = lib, jvm
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = [c_void_p, c_void_p] + args

    def __call__(self, short, *args):
        name = (NSB if hasattr(self.lib, NSB + short) else MNB) + short
        return getattr(self.lib, name)(self.jvm.env, None, *args)
