// JNI bridge of libaero_lbm.so: the native methods that the mod's Java classes
// declare (src/main/java/com/aerodynamics4mc/runtime/NativeSimulationBridge.java:1239-1656,
// MesoscaleNativeBridge.java:129-165), each a thin wrapper over the C ABI in
// include/aero_lbm_capi.h. Every array argument must be present and have exactly the
// length the call needs, checked before any array is pinned (the reference's
// contract, native/src/aero_lbm_simulation_bridge.cpp:7026-8985); arrays are then
// pinned with Get<Type>ArrayElements and released with mode 0 when the call wrote
// them and succeeded, JNI_ABORT otherwise (inputs are never written back,
// simulation_bridge.cpp:7717-7727).
//
// Built into the library unconditionally against jni.h beside this file (the JNI
// specification's types and function table); the library needs no JDK to build or load.
#include "jni.h"

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/aero_lbm_capi.h"

namespace {

// pinned view of a Java primitive array
template <class JArray, class JElem>
struct Pinned {
    JNIEnv* env;
    JArray array;
    JElem* data = nullptr;
    jsize length = 0;
    bool write_back = false;
    Pinned(JNIEnv* e, JArray a) : env(e), array(a) {
        if (!a) return;
        data = get(e, a);
    }
    ~Pinned() {
        if (data) release(env, array, data, write_back ? 0 : JNI_ABORT);
    }
    explicit operator bool() const { return data != nullptr; }
    static JElem* get(JNIEnv* e, jbyteArray a) { return e->GetByteArrayElements(a, nullptr); }
    static JElem* get(JNIEnv* e, jshortArray a) { return e->GetShortArrayElements(a, nullptr); }
    static JElem* get(JNIEnv* e, jintArray a) { return e->GetIntArrayElements(a, nullptr); }
    static JElem* get(JNIEnv* e, jfloatArray a) { return e->GetFloatArrayElements(a, nullptr); }
    static void release(JNIEnv* e, jbyteArray a, jbyte* p, jint m) { e->ReleaseByteArrayElements(a, p, m); }
    static void release(JNIEnv* e, jshortArray a, jshort* p, jint m) { e->ReleaseShortArrayElements(a, p, m); }
    static void release(JNIEnv* e, jintArray a, jint* p, jint m) { e->ReleaseIntArrayElements(a, p, m); }
    static void release(JNIEnv* e, jfloatArray a, jfloat* p, jint m) { e->ReleaseFloatArrayElements(a, p, m); }
};
using Bytes = Pinned<jbyteArray, jbyte>;
using Shorts = Pinned<jshortArray, jshort>;
using Ints = Pinned<jintArray, jint>;
using Floats = Pinned<jfloatArray, jfloat>;

const uint8_t* u8(const Bytes& b) { return reinterpret_cast<const uint8_t*>(b.data); }
const uint16_t* u16(const Shorts& s) { return reinterpret_cast<const uint16_t*>(s.data); }

jboolean jbool(int ok) { return ok ? JNI_TRUE : JNI_FALSE; }

// length of a Java array, -1 for null (every length test below then fails)
jsize len(JNIEnv* env, jarray a) { return a ? env->GetArrayLength(a) : -1; }

jintArray coords_array(JNIEnv* env, long long service, long long world, bool active) {
    const int n = active ? aero_lbm_simulation_get_brick_world_active_brick_count(service, world)
                         : aero_lbm_simulation_get_brick_world_resident_brick_count(service, world);
    if (n < 0) return nullptr;
    std::vector<jint> coords((std::size_t)n * 3 + 1);
    const int got = active ? aero_lbm_simulation_copy_brick_world_active_bricks(service, world, coords.data(), n)
                           : aero_lbm_simulation_copy_brick_world_resident_bricks(service, world, coords.data(), n);
    if (got < 0) return nullptr;
    jintArray out = env->NewIntArray(got * 3);
    if (out && got > 0) env->SetIntArrayRegion(out, 0, got * 3, coords.data());
    return out;
}

AeroLbmMesoscaleConfig meso_config(jint nx, jint ny, jint nz, jfloat dx, jfloat dt, jfloat nu, jfloat pr, jfloat prt) {
    AeroLbmMesoscaleConfig cfg{};
    aero_lbm_mesoscale_default_config(&cfg);
    cfg.nx = nx;
    cfg.ny = ny;
    cfg.nz = nz;
    cfg.dx_m = dx;
    cfg.dt_s = dt;
    cfg.molecular_nu_m2_s = nu;
    cfg.prandtl_air = pr;
    cfg.turbulent_prandtl = prt;
    return cfg;
}

constexpr int kDeltaInts = 8, kDeltaFloats = 4;  // NativeSimulationBridge.java:30-31

}  // namespace

#define NSB(name) Java_com_aerodynamics4mc_runtime_NativeSimulationBridge_##name
#define MNB(name) Java_com_aerodynamics4mc_runtime_MesoscaleNativeBridge_##name

extern "C" {

JNIEXPORT jlong JNICALL NSB(nativeCreateService)(JNIEnv*, jclass) {
    long long key = 0;
    return aero_lbm_simulation_create_service(&key) ? (jlong)key : 0;
}

JNIEXPORT void JNICALL NSB(nativeReleaseService)(JNIEnv*, jclass, jlong service) {
    aero_lbm_simulation_release_service(service);
}

JNIEXPORT jboolean JNICALL NSB(nativeSetFocus)(JNIEnv*, jclass, jlong service, jint bx, jint by, jint bz, jint r) {
    return jbool(aero_lbm_simulation_set_focus(service, bx, by, bz, r));
}

JNIEXPORT jboolean JNICALL NSB(nativeSubmitWorldDeltas)(JNIEnv* env, jclass, jlong service, jintArray ints,
                                                       jfloatArray floats) {
    if (!ints || !floats) return JNI_FALSE;
    const jsize ni = len(env, ints), nf = len(env, floats);
    if (ni % kDeltaInts != 0 || nf % kDeltaFloats != 0 || ni / kDeltaInts != nf / kDeltaFloats) return JNI_FALSE;
    const int count = ni / kDeltaInts;
    Ints i(env, ints);
    Floats f(env, floats);
    if (!i || !f) return JNI_FALSE;
    std::vector<AeroLbmWorldDelta> deltas((std::size_t)count);
    for (int k = 0; k < count; ++k) {
        const jint* a = i.data + k * kDeltaInts;
        const jfloat* v = f.data + k * kDeltaFloats;
        deltas[k] = AeroLbmWorldDelta{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], v[0], v[1], v[2], v[3]};
    }
    return jbool(aero_lbm_simulation_submit_world_deltas(service, deltas.data(), count));
}

JNIEXPORT jboolean JNICALL NSB(nativeEnsureBrickWorldRuntime)(JNIEnv*, jclass, jlong service, jlong world, jint size,
                                                             jfloat dx, jfloat dt) {
    return jbool(aero_lbm_simulation_ensure_brick_world_runtime(service, world, size, dx, dt));
}

JNIEXPORT jboolean JNICALL NSB(nativeSetBrickWorldSolverMode)(JNIEnv*, jclass, jlong service, jlong world, jint mode) {
    return jbool(aero_lbm_simulation_set_brick_world_solver_mode(service, world, mode));
}

JNIEXPORT jboolean JNICALL NSB(nativeSetBrickWorldActiveHints)(JNIEnv* env, jclass, jlong service, jlong world,
                                                              jint size, jintArray coords, jint count) {
    if (!coords || count < 0 || len(env, coords) < count * 3) return JNI_FALSE;
    Ints c(env, coords);
    if (!c) return JNI_FALSE;
    return jbool(aero_lbm_simulation_set_brick_world_active_hints(service, world, size, c.data, count));
}

JNIEXPORT jboolean JNICALL NSB(nativeSetBrickWorldExactActiveHints)(JNIEnv* env, jclass, jlong service, jlong world,
                                                                   jint size, jintArray coords, jint count) {
    if (!coords || count < 0 || len(env, coords) < count * 3) return JNI_FALSE;
    Ints c(env, coords);
    if (!c) return JNI_FALSE;
    return jbool(aero_lbm_simulation_set_brick_world_exact_active_hints(service, world, size, c.data, count));
}

JNIEXPORT jboolean JNICALL NSB(nativeGetBrickWorldRuntimeStatus)(JNIEnv* env, jclass, jlong service, jlong world,
                                                                jintArray out) {
    if (len(env, out) != AERO_LBM_SIMULATION_BRICK_RUNTIME_STATUS_FIELDS) return JNI_FALSE;
    Ints o(env, out);
    if (!o) return JNI_FALSE;
    const int ok = aero_lbm_simulation_get_brick_world_runtime_status(service, world, o.data,
                                                                     AERO_LBM_SIMULATION_BRICK_RUNTIME_STATUS_FIELDS);
    o.write_back = ok != 0;
    return jbool(ok);
}

JNIEXPORT jboolean JNICALL NSB(nativeStepBrickWorldRuntime)(JNIEnv*, jclass, jlong service, jlong world, jint steps) {
    return jbool(aero_lbm_simulation_step_brick_world_runtime(service, world, steps));
}

JNIEXPORT jintArray JNICALL NSB(nativeGetBrickWorldResidentBrickCoords)(JNIEnv* env, jclass, jlong service,
                                                                        jlong world) {
    return coords_array(env, service, world, false);
}

JNIEXPORT jintArray JNICALL NSB(nativeGetBrickWorldActiveBrickCoords)(JNIEnv* env, jclass, jlong service,
                                                                      jlong world) {
    return coords_array(env, service, world, true);
}

static jboolean upload_static(JNIEnv* env, jlong service, jlong world, jint size, jint bx, jint by, jint bz,
                              jbyteArray obstacle, jbyteArray kind, jshortArray open, jfloatArray emitter,
                              bool with_sources, jbyteArray fan_src, jfloatArray emitter_src, jbyteArray sky,
                              jbyteArray direct) {
    const long long cells = (long long)size * size * size;
    if (cells <= 0 || len(env, obstacle) != cells || len(env, kind) != cells || len(env, open) != cells ||
        len(env, emitter) != cells || len(env, sky) != cells * 6 || len(env, direct) != cells * 6)
        return JNI_FALSE;
    if (with_sources && (len(env, fan_src) != cells || len(env, emitter_src) != cells)) return JNI_FALSE;
    Bytes ob(env, obstacle), kd(env, kind), sk(env, sky), dr(env, direct), fs(env, fan_src);
    Shorts op(env, open);
    Floats em(env, emitter), es(env, emitter_src);
    if (!ob || !kd || !op || !em || !sk || !dr || (with_sources && (!fs || !es))) return JNI_FALSE;
    return jbool(aero_lbm_simulation_upload_brick_world_static_brick_with_sources(
        service, world, size, bx, by, bz, u8(ob), u8(kd), u16(op), em.data, with_sources ? u8(fs) : nullptr,
        with_sources ? es.data : nullptr, u8(sk), u8(dr)));
}

JNIEXPORT jboolean JNICALL NSB(nativeUploadBrickWorldStaticBrick)(JNIEnv* env, jclass, jlong service, jlong world,
                                                                 jint size, jint bx, jint by, jint bz,
                                                                 jbyteArray obstacle, jbyteArray kind,
                                                                 jshortArray open, jfloatArray emitter,
                                                                 jbyteArray sky, jbyteArray direct) {
    return upload_static(env, service, world, size, bx, by, bz, obstacle, kind, open, emitter, false, nullptr, nullptr,
                         sky, direct);
}

JNIEXPORT jboolean JNICALL NSB(nativeUploadBrickWorldStaticBrickWithSources)(
    JNIEnv* env, jclass, jlong service, jlong world, jint size, jint bx, jint by, jint bz, jbyteArray obstacle,
    jbyteArray kind, jshortArray open, jfloatArray emitter, jbyteArray fan_src, jfloatArray emitter_src,
    jbyteArray sky, jbyteArray direct) {
    return upload_static(env, service, world, size, bx, by, bz, obstacle, kind, open, emitter, true, fan_src,
                         emitter_src, sky, direct);
}

JNIEXPORT jboolean JNICALL NSB(nativeQueueBrickWorldStaticBrickUpload)(JNIEnv* env, jclass, jlong service, jlong world,
                                                                      jint size, jint bx, jint by, jint bz,
                                                                      jbyteArray obstacle, jbyteArray kind,
                                                                      jshortArray open, jfloatArray emitter,
                                                                      jbyteArray sky, jbyteArray direct) {
    const long long cells = (long long)size * size * size;
    if (cells <= 0 || len(env, obstacle) != cells || len(env, kind) != cells || len(env, open) != cells ||
        len(env, emitter) != cells || len(env, sky) != cells * 6 || len(env, direct) != cells * 6)
        return JNI_FALSE;
    Bytes ob(env, obstacle), kd(env, kind), sk(env, sky), dr(env, direct);
    Shorts op(env, open);
    Floats em(env, emitter);
    if (!ob || !kd || !op || !em || !sk || !dr) return JNI_FALSE;
    return jbool(aero_lbm_simulation_queue_brick_world_static_brick_upload(service, world, size, bx, by, bz, u8(ob),
                                                                          u8(kd), u16(op), em.data, u8(sk), u8(dr)) >
                 0);
}

static jboolean dynamic_io(JNIEnv* env, int op, jlong service, jlong world, jint size, jint bx, jint by, jint bz,
                           jfloatArray flow, jfloatArray air, jfloatArray surface) {
    const long long cells = (long long)size * size * size;
    if (cells <= 0 || len(env, flow) != cells * AERO_LBM_SIMULATION_FLOW_STATE_CHANNELS || len(env, air) != cells ||
        len(env, surface) != cells)
        return JNI_FALSE;
    Floats f(env, flow), a(env, air), s(env, surface);
    if (!f || !a || !s) return JNI_FALSE;
    int ok = 0;
    if (op == 0) ok = aero_lbm_simulation_upload_brick_world_dynamic_brick(service, world, size, bx, by, bz, f.data,
                                                                          a.data, s.data);
    if (op == 1) ok = aero_lbm_simulation_upload_brick_world_boundary_reference_brick(service, world, size, bx, by, bz,
                                                                                     f.data, a.data, s.data);
    if (op == 2) {
        ok = aero_lbm_simulation_copy_brick_world_dynamic_brick(service, world, size, bx, by, bz, f.data, a.data,
                                                                s.data);
        f.write_back = a.write_back = s.write_back = ok != 0;
    }
    return jbool(ok);
}

JNIEXPORT jboolean JNICALL NSB(nativeUploadBrickWorldDynamicBrick)(JNIEnv* env, jclass, jlong service, jlong world,
                                                                  jint size, jint bx, jint by, jint bz,
                                                                  jfloatArray flow, jfloatArray air,
                                                                  jfloatArray surface) {
    return dynamic_io(env, 0, service, world, size, bx, by, bz, flow, air, surface);
}

JNIEXPORT jboolean JNICALL NSB(nativeUploadBrickWorldBoundaryReferenceBrick)(JNIEnv* env, jclass, jlong service,
                                                                            jlong world, jint size, jint bx, jint by,
                                                                            jint bz, jfloatArray flow,
                                                                            jfloatArray air, jfloatArray surface) {
    return dynamic_io(env, 1, service, world, size, bx, by, bz, flow, air, surface);
}

JNIEXPORT jboolean JNICALL NSB(nativeCopyBrickWorldDynamicBrick)(JNIEnv* env, jclass, jlong service, jlong world,
                                                                jint size, jint bx, jint by, jint bz,
                                                                jfloatArray flow, jfloatArray air,
                                                                jfloatArray surface) {
    return dynamic_io(env, 2, service, world, size, bx, by, bz, flow, air, surface);
}

JNIEXPORT jboolean JNICALL NSB(nativeCopyBrickWorldPackedFlowAtlas)(JNIEnv* env, jclass, jlong service, jlong world,
                                                                   jint size, jint bx, jint by, jint bz, jint stride,
                                                                   jshortArray out) {
    if (!out || stride <= 0) return JNI_FALSE;
    const int res = (size + stride - 1) / stride;
    const long long values = (long long)res * res * res * AERO_LBM_SIMULATION_PACKED_ATLAS_CHANNELS;
    if (res <= 0 || len(env, out) != values) return JNI_FALSE;
    Shorts o(env, out);
    if (!o) return JNI_FALSE;
    const int ok = aero_lbm_simulation_copy_brick_world_packed_flow_atlas(
        service, world, size, bx, by, bz, stride, reinterpret_cast<int16_t*>(o.data), (int)values);
    o.write_back = ok != 0;
    return jbool(ok);
}

// a caller's packed flow atlas under a key of the service (ref:
// native/src/aero_lbm_simulation_bridge.cpp:8789-8840): the input goes back with
// JNI_ABORT, the polled output with mode 0 whatever the result, as the reference
JNIEXPORT jboolean JNICALL NSB(nativeSetPackedFlowAtlas)(JNIEnv* env, jclass, jlong service, jlong atlas,
                                                        jshortArray values) {
    if (!values) return JNI_FALSE;
    const jsize n = len(env, values);
    Shorts v(env, values);
    if (!v) return JNI_FALSE;
    return jbool(aero_lbm_simulation_set_packed_flow_atlas(service, atlas, reinterpret_cast<const int16_t*>(v.data),
                                                          (int)n));
}

JNIEXPORT jboolean JNICALL NSB(nativePollPackedFlowAtlas)(JNIEnv* env, jclass, jlong service, jlong atlas,
                                                         jshortArray out) {
    if (!out) return JNI_FALSE;
    const jsize n = len(env, out);
    Shorts o(env, out);
    if (!o) return JNI_FALSE;
    o.write_back = true;
    return jbool(aero_lbm_simulation_poll_packed_flow_atlas(service, atlas, reinterpret_cast<int16_t*>(o.data),
                                                           (int)n));
}

// ---- analysis codec (ref: simulation_bridge.cpp:8880-8927 and 8929-8974): a zfp
// stream of a 3D float grid as a new byte[], null on any failure; decompression into
// a float[] of exactly nx * ny * nz values, written back only on success
JNIEXPORT jbyteArray JNICALL NSB(nativeCompressFloatGrid3d)(JNIEnv* env, jclass, jfloatArray values, jint nx, jint ny,
                                                           jint nz, jdouble tolerance) {
    const long long cells = (long long)nx * ny * nz;
    if (!values || nx <= 0 || ny <= 0 || nz <= 0 || len(env, values) != cells) return nullptr;
    const long long cap = aero_lbm_analysis_max_compressed_size_3d(nx, ny, nz);
    if (cap <= 0) return nullptr;
    std::vector<uint8_t> bytes((std::size_t)cap);
    long long n = 0;
    {
        Floats v(env, values);
        if (!v) return nullptr;
        n = aero_lbm_analysis_compress_float_grid_3d(v.data, nx, ny, nz, (double)tolerance, bytes.data(), cap);
    }
    if (n <= 0 || n > 0x7fffffffll) return nullptr;
    jbyteArray out = env->NewByteArray((jsize)n);
    if (!out) return nullptr;
    env->SetByteArrayRegion(out, 0, (jsize)n, reinterpret_cast<const jbyte*>(bytes.data()));
    return out;
}

JNIEXPORT jboolean JNICALL NSB(nativeDecompressFloatGrid3d)(JNIEnv* env, jclass, jbyteArray compressed, jint nx,
                                                           jint ny, jint nz, jfloatArray out_values) {
    if (!compressed || !out_values) return JNI_FALSE;
    const long long cells = (long long)nx * ny * nz;
    if (nx <= 0 || ny <= 0 || nz <= 0 || len(env, out_values) != cells) return JNI_FALSE;
    const jsize size = env->GetArrayLength(compressed);
    std::vector<uint8_t> bytes((std::size_t)size);
    if (size > 0) {
        env->GetByteArrayRegion(compressed, 0, size, reinterpret_cast<jbyte*>(bytes.data()));
        if (env->ExceptionCheck()) return JNI_FALSE;
    }
    Floats o(env, out_values);
    if (!o) return JNI_FALSE;
    const int ok = aero_lbm_analysis_decompress_float_grid_3d(bytes.data(), (long long)size, nx, ny, nz, o.data);
    o.write_back = ok != 0;
    return jbool(ok);
}

// ref: simulation_bridge.cpp:8976-8979
JNIEXPORT jstring JNICALL NSB(nativeRuntimeInfo)(JNIEnv* env, jclass) {
    const char* s = aero_lbm_simulation_runtime_info();
    return env->NewStringUTF(s ? s : "");
}

JNIEXPORT jstring JNICALL NSB(nativeLastError)(JNIEnv* env, jclass) {
    const char* s = aero_lbm_simulation_last_error();
    if (!s || !*s) s = aero_lbm_last_error();
    return env->NewStringUTF(s ? s : "");
}

// ---- mesoscale (ref: native/src/aero_lbm_mesoscale.cpp:1511-1603, aero_lbm_jni.cpp:10493-10533)
JNIEXPORT jfloatArray JNICALL MNB(nativeDeriveTransport)(JNIEnv* env, jclass, jint nx, jint ny, jint nz, jfloat dx,
                                                        jfloat dt, jfloat nu, jfloat pr, jfloat prt) {
    const AeroLbmMesoscaleConfig cfg = meso_config(nx, ny, nz, dx, dt, nu, pr, prt);
    AeroLbmMesoscaleTransport t{};
    if (!aero_lbm_mesoscale_derive_transport(&cfg, &t)) return nullptr;
    const jfloat v[5] = {t.velocity_scale_m_s_per_lattice, t.nu_molecular_lattice, t.alpha_molecular_lattice,
                         t.tau_shear_molecular, t.tau_thermal_molecular};
    jfloatArray out = env->NewFloatArray(5);
    if (out) env->SetFloatArrayRegion(out, 0, 5, v);
    return out;
}

JNIEXPORT jlong JNICALL MNB(nativeCreateContext)(JNIEnv*, jclass, jint nx, jint ny, jint nz, jfloat dx, jfloat dt,
                                                jfloat nu, jfloat pr, jfloat prt) {
    const AeroLbmMesoscaleConfig cfg = meso_config(nx, ny, nz, dx, dt, nu, pr, prt);
    long long key = 0;
    return aero_lbm_mesoscale_create_context(&cfg, &key) ? (jlong)key : 0;
}

JNIEXPORT jboolean JNICALL MNB(nativeStepContext)(JNIEnv* env, jclass, jlong key, jint nx, jint ny, jint nz,
                                                 jfloat dx, jfloat dt, jfloat nu, jfloat pr, jfloat prt,
                                                 jfloatArray forcing, jfloatArray state) {
    const AeroLbmMesoscaleConfig cfg = meso_config(nx, ny, nz, dx, dt, nu, pr, prt);
    const long long cells = (long long)nx * ny * nz;
    if (len(env, forcing) != cells * AERO_LBM_MESOSCALE_FORCING_CHANNELS ||
        len(env, state) != cells * AERO_LBM_MESOSCALE_STATE_CHANNELS)
        return JNI_FALSE;
    Floats f(env, forcing), s(env, state);
    if (!f || !s) return JNI_FALSE;
    const int ok = aero_lbm_mesoscale_step_context(key, &cfg, f.data, s.data);
    s.write_back = ok != 0;
    return jbool(ok);
}

JNIEXPORT void JNICALL MNB(nativeReleaseContext)(JNIEnv*, jclass, jlong key) { aero_lbm_mesoscale_release_context(key); }

}  // extern "C"
