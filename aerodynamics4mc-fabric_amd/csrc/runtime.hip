// Solver-context runtime and the core aero_lbm_* C ABI.
//
// Mirrors the reference's context registry and entry points
// (ref: native/src/aero_lbm_jni.cpp:8592-10368): one global configuration set
// by aero_lbm_init_rect, contexts keyed by a caller-chosen long long and
// created on first step, 1/0 returns with the reason in aero_lbm_last_error().
// Unlike the reference there is no CPU fallback: every context lives on the
// GPU (device 0, one HIP stream) and an API call fails with an error when HIP
// is unavailable.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aero_lbm_capi.h"
#include "fast_kernels.h"
#include "kernels.h"
#include "runtime_internal.h"

static_assert(sizeof(AeroLbmBoundaryFaceConfig) == 32, "face layout");
static_assert(sizeof(AeroLbmBenchmarkConfig) == 320, "benchmark config layout");
static_assert(sizeof(AeroLbmTimingSnapshot) == 72, "timing snapshot layout");
static_assert(sizeof(AeroLbmWorldDelta) == 48, "world delta layout");
static_assert(sizeof(AeroLbmMesoscaleConfig) == 96, "mesoscale config layout");
static_assert(sizeof(AeroLbmMesoscaleTransport) == 20, "mesoscale transport layout");

namespace aero {

RuntimeState& rt() {
    static RuntimeState s;
    return s;
}

void set_error(const std::string& e) { rt().last_error = e; }
void clear_error() { rt().last_error.clear(); }

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

bool ensure_device() {
    RuntimeState& s = rt();
    if (s.device_ready) return true;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error("no HIP device available (libaero_lbm has no CPU fallback)");
        return false;
    }
    if (!hip_ok(hipSetDevice(0), "hipSetDevice")) return false;
    hipDeviceProp_t prop{};
    if (!hip_ok(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties")) return false;
    s.device_name = prop.name;
    s.device_arch = prop.gcnArchName;
    s.cu_count = prop.multiProcessorCount;
    if (!hip_ok(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate")) return false;
    s.device_ready = true;
    return true;
}

// ---- benchmark configuration (ref: native/src/aero_lbm_jni.cpp:298-441) ----
static AeroLbmBoundaryFaceConfig make_face(int hk, int tk) {
    AeroLbmBoundaryFaceConfig f{};
    f.hydrodynamic_kind = hk;
    f.thermal_kind = tk;
    return f;
}
static void set_all_faces(AeroLbmBenchmarkConfig& c, int hk, int tk) {
    c.x_min = c.x_max = c.y_min = c.y_max = c.z_min = c.z_max = make_face(hk, tk);
}

AeroLbmBenchmarkConfig default_benchmark_config() {
    AeroLbmBenchmarkConfig c{};
    c.abi_version = AERO_LBM_BENCHMARK_ABI_VERSION;
    c.struct_size = sizeof(AeroLbmBenchmarkConfig);
    c.flags = AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING | AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_NOISE |
              AERO_LBM_BENCHMARK_FLAG_DISABLE_SPONGE | AERO_LBM_BENCHMARK_FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND |
              AERO_LBM_BENCHMARK_FLAG_DISABLE_SGS;
    c.enabled = 0;
    c.preset = AERO_LBM_BENCHMARK_PRESET_NONE;
    c.reynolds_number = 100.0f;
    c.rayleigh_number = 1.0e5f;
    c.prandtl_number = 0.71f;
    c.mach_number = 0.05f;
    c.reference_density = 1.0f;
    c.reference_temperature = 0.0f;
    c.reference_length = 1.0f;
    c.gravity[1] = -1.0f;
    set_all_faces(c, AERO_LBM_HYDRO_BOUNDARY_INHERIT_GAME, AERO_LBM_THERMAL_BOUNDARY_INHERIT_GAME);
    return c;
}

static bool valid_preset(int p) { return p >= 0 && p <= 6; }

static void apply_preset(AeroLbmBenchmarkConfig& c, int preset) {
    c = default_benchmark_config();
    c.enabled = 1;
    c.preset = preset;
    const uint32_t no_src_buoy =
        AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE | AERO_LBM_BENCHMARK_FLAG_DISABLE_BUOYANCY;
    switch (preset) {
        case AERO_LBM_BENCHMARK_PRESET_TAYLOR_GREEN_3D:
            c.flags |= AERO_LBM_BENCHMARK_FLAG_DISABLE_CONVECTIVE_OUTFLOW | no_src_buoy;
            set_all_faces(c, AERO_LBM_HYDRO_BOUNDARY_PERIODIC, AERO_LBM_THERMAL_BOUNDARY_DISABLED);
            break;
        case AERO_LBM_BENCHMARK_PRESET_LID_DRIVEN_CAVITY_2D:
        case AERO_LBM_BENCHMARK_PRESET_LID_DRIVEN_CAVITY_3D:
            c.reynolds_number = 1000.0f;
            set_all_faces(c, AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK, AERO_LBM_THERMAL_BOUNDARY_ADIABATIC);
            c.y_max = make_face(AERO_LBM_HYDRO_BOUNDARY_MOVING_WALL, AERO_LBM_THERMAL_BOUNDARY_ADIABATIC);
            c.y_max.velocity[0] = 0.10f;
            c.flags |= no_src_buoy;
            if (preset == AERO_LBM_BENCHMARK_PRESET_LID_DRIVEN_CAVITY_2D) {
                c.z_min = c.z_max = make_face(AERO_LBM_HYDRO_BOUNDARY_PERIODIC, AERO_LBM_THERMAL_BOUNDARY_PERIODIC);
            }
            break;
        case AERO_LBM_BENCHMARK_PRESET_CYLINDER_CROSSFLOW_2D:
            c.reynolds_number = 100.0f;
            c.flags |= no_src_buoy;
            set_all_faces(c, AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK, AERO_LBM_THERMAL_BOUNDARY_DISABLED);
            c.x_min = make_face(AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET, AERO_LBM_THERMAL_BOUNDARY_DISABLED);
            c.initial_velocity[0] = 0.08f;
            c.x_min.velocity[0] = 1.5f * 0.08f;
            c.x_max = make_face(AERO_LBM_HYDRO_BOUNDARY_CONVECTIVE_OUTFLOW, AERO_LBM_THERMAL_BOUNDARY_DISABLED);
            c.z_min = c.z_max = make_face(AERO_LBM_HYDRO_BOUNDARY_PERIODIC, AERO_LBM_THERMAL_BOUNDARY_PERIODIC);
            break;
        case AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_2D:
        case AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_3D:
            c.rayleigh_number = 1.0e5f;
            c.prandtl_number = 0.71f;
            c.mach_number = 0.15f;
            c.flags |= AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE;
            set_all_faces(c, AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK, AERO_LBM_THERMAL_BOUNDARY_ADIABATIC);
            c.x_min = make_face(AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK, AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET);
            c.x_min.temperature = 0.5f;
            c.x_max = make_face(AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK, AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET);
            c.x_max.temperature = -0.5f;
            if (preset == AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_2D) {
                c.z_min = c.z_max = make_face(AERO_LBM_HYDRO_BOUNDARY_PERIODIC, AERO_LBM_THERMAL_BOUNDARY_PERIODIC);
            }
            break;
        default:
            c.enabled = 0;
            c.preset = AERO_LBM_BENCHMARK_PRESET_NONE;
            break;
    }
}

static float fin_or(float v, float fb) { return std::isfinite(v) ? v : fb; }
static void sanitize_face(AeroLbmBoundaryFaceConfig& f) {
    if (f.hydrodynamic_kind < 0 || f.hydrodynamic_kind > 7) f.hydrodynamic_kind = 0;
    if (f.thermal_kind < 0 || f.thermal_kind > 5) f.thermal_kind = 0;
    for (float& v : f.velocity) v = fin_or(v, 0.0f);
    f.pressure = fin_or(f.pressure, 0.0f);
    f.temperature = fin_or(f.temperature, 0.0f);
    f.heat_flux = fin_or(f.heat_flux, 0.0f);
}

void sanitize_benchmark_config(AeroLbmBenchmarkConfig& c) {
    c.abi_version = AERO_LBM_BENCHMARK_ABI_VERSION;
    c.struct_size = sizeof(AeroLbmBenchmarkConfig);
    c.flags &= 0xFFu;
    c.enabled = c.enabled != 0 ? 1 : 0;
    if (!valid_preset(c.preset)) c.preset = AERO_LBM_BENCHMARK_PRESET_NONE;
    c.reynolds_number = fin_or(c.reynolds_number, 100.0f);
    c.rayleigh_number = fin_or(c.rayleigh_number, 1.0e5f);
    c.prandtl_number = fin_or(c.prandtl_number, 0.71f);
    const float m = fin_or(c.mach_number, 0.05f);
    c.mach_number = m < 1.0e-4f ? 1.0e-4f : (m > 0.30f ? 0.30f : m);
    c.reference_density = fin_or(c.reference_density, 1.0f);
    c.reference_temperature = fin_or(c.reference_temperature, 0.0f);
    const float rl = fin_or(c.reference_length, 1.0f);
    c.reference_length = rl < 1.0e-6f ? 1.0e-6f : rl;
    for (int i = 0; i < 3; ++i) {
        c.body_force[i] = fin_or(c.body_force[i], 0.0f);
        c.initial_velocity[i] = fin_or(c.initial_velocity[i], 0.0f);
    }
    c.gravity[0] = fin_or(c.gravity[0], 0.0f);
    c.gravity[1] = fin_or(c.gravity[1], -1.0f);
    c.gravity[2] = fin_or(c.gravity[2], 0.0f);
    c.initial_pressure = fin_or(c.initial_pressure, 0.0f);
    c.initial_temperature = fin_or(c.initial_temperature, 0.0f);
    sanitize_face(c.x_min);
    sanitize_face(c.x_max);
    sanitize_face(c.y_min);
    sanitize_face(c.y_max);
    sanitize_face(c.z_min);
    sanitize_face(c.z_max);
}

AeroLbmSolverOverrides default_overrides() {
    AeroLbmSolverOverrides o{};
    o.struct_size = sizeof(AeroLbmSolverOverrides);
    o.sgs_constant = NAN;
    o.sgs_nut_cap_ratio = NAN;
    o.boussinesq_beta = NAN;
    o.thermal_diffusivity = NAN;
    o.gravity_axis = -1;
    o.gravity_sign = 0;
    o.kernel_variant = AERO_LBM_KERNEL_FAST;
    return o;
}

// ---- parameter resolution (ref: native/src/aero_lbm_jni.cpp:787-993, benchmark mode) ----
static float rmaxf_(float a, float b) { return (a < b) ? b : a; }
static float rminf_(float a, float b) { return (b < a) ? b : a; }
static float clampr_(float v, float lo, float hi) { return rminf_(hi, rmaxf_(lo, v)); }
static float clamp_nu(float nu) { return clampr_(nu, 1.0e-8f, (0.95f - 0.5f) / 3.0f); }
static float tau_from_nu(float nu) { return 0.5f + 3.0f * clamp_nu(nu); }

static bool heated_cavity(const AeroLbmBenchmarkConfig& g) {
    return g.preset == AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_2D ||
           g.preset == AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_3D;
}

// ref: jni.cpp:881-899: span of the Dirichlet face temperatures (1 without any)
static float dirichlet_temperature_span(const AeroLbmBenchmarkConfig& g) {
    const AeroLbmBoundaryFaceConfig* faces[6] = {&g.x_min, &g.x_max, &g.y_min, &g.y_max, &g.z_min, &g.z_max};
    float lo = 1.0e9f, hi = -1.0e9f;
    bool any = false;
    for (auto* f : faces) {
        if (f->thermal_kind != AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET) continue;
        any = true;
        lo = std::min(lo, f->temperature);
        hi = std::max(hi, f->temperature);
    }
    return any ? rmaxf_(1.0e-6f, hi - lo) : 1.0f;
}

// ref: jni.cpp:823-859 benchmark_reference_speed (cylinder preset not supported)
static float reference_speed(const AeroLbmBenchmarkConfig& g) {
    if (heated_cavity(g)) return clampr_(fin_or(g.mach_number, 0.05f), 1.0e-4f, 0.30f) * std::sqrt(1.0f / 3.0f);
    const float* iv = g.initial_velocity;
    float speed = std::sqrt(iv[0] * iv[0] + iv[1] * iv[1] + iv[2] * iv[2]);
    if (g.preset == AERO_LBM_BENCHMARK_PRESET_CYLINDER_CROSSFLOW_2D) {
        if (speed > 1.0e-8f) return speed;
        const float* v = g.x_min.velocity;
        return (2.0f / 3.0f) * std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    const AeroLbmBoundaryFaceConfig* faces[6] = {&g.x_min, &g.x_max, &g.y_min, &g.y_max, &g.z_min, &g.z_max};
    for (auto* f : faces) {
        const float* v = f->velocity;
        speed = rmaxf_(speed, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    return speed;
}

bool resolve_params(int nx, int ny, int nz, int in_ch, StepParams& P) {
    RuntimeState& s = rt();
    const AeroLbmBenchmarkConfig& g = s.bench;
    const bool runtime = !g.enabled;
    std::memset(&P, 0, sizeof(P));
    P.nx = nx;
    P.ny = ny;
    P.nz = nz;
    P.cells = nx * ny * nz;
    P.plane = fast_plane_stride(P.cells);
    P.in_ch = in_ch;
    const AeroLbmBoundaryFaceConfig* faces[6] = {&g.x_min, &g.x_max, &g.y_min, &g.y_max, &g.z_min, &g.z_max};
    for (int i = 0; i < 6; ++i) {
        FaceParams& f = P.face[i];
        f.hydro_kind = faces[i]->hydrodynamic_kind;
        f.thermal_kind = faces[i]->thermal_kind;
        for (int k = 0; k < 3; ++k) f.vel[k] = faces[i]->velocity[k];
        f.pressure = faces[i]->pressure;
        f.temperature = faces[i]->temperature;
        f.heat_flux = faces[i]->heat_flux;
    }
    for (int a = 0; a < 3; ++a) {
        P.hydro_wrap_lo[a] = faces[2 * a]->hydrodynamic_kind == AERO_LBM_HYDRO_BOUNDARY_PERIODIC;
        P.hydro_wrap_hi[a] = faces[2 * a + 1]->hydrodynamic_kind == AERO_LBM_HYDRO_BOUNDARY_PERIODIC;
        P.thermal_wrap_lo[a] = faces[2 * a]->thermal_kind == AERO_LBM_THERMAL_BOUNDARY_PERIODIC;
        P.thermal_wrap_hi[a] = faces[2 * a + 1]->thermal_kind == AERO_LBM_THERMAL_BOUNDARY_PERIODIC;
        P.hydro_periodic[a] = P.hydro_wrap_lo[a] && P.hydro_wrap_hi[a];
    }
    // benchmark flags apply in benchmark mode only (ref: jni.cpp:722-724)
    auto flag = [&](uint32_t b) { return !runtime && (g.flags & b) != 0; };
    P.sgs_on = !flag(AERO_LBM_BENCHMARK_FLAG_DISABLE_SGS);
    P.buoyancy_on = !flag(AERO_LBM_BENCHMARK_FLAG_DISABLE_BUOYANCY);
    P.sponge_on = !flag(AERO_LBM_BENCHMARK_FLAG_DISABLE_SPONGE);
    P.source_on = !flag(AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE);
    P.convective_disabled = flag(AERO_LBM_BENCHMARK_FLAG_DISABLE_CONVECTIVE_OUTFLOW);
    const float re = fin_or(g.reynolds_number, 0.0f);
    const float ref_speed = reference_speed(g);
    const float ref_len = rmaxf_(1.0e-6f, fin_or(g.reference_length, 1.0f));
    const float nu_rt = clamp_nu(1.50e-5f * (1.0f / 20.0f) / (1.0f * 1.0f));
    const bool heated = heated_cavity(g);
    const float ra = rmaxf_(1.0e-6f, fin_or(g.rayleigh_number, 1.0e5f));
    const float pr = rmaxf_(1.0e-6f, fin_or(g.prandtl_number, 0.71f));
    float nu;
    if (heated) nu = clamp_nu(ref_speed * ref_len * std::sqrt(pr / ra));  // ref: jni.cpp:937-943
    else if (re <= 1.0e-6f || ref_speed <= 1.0e-8f) nu = nu_rt;
    else nu = clamp_nu(ref_speed * ref_len / re);
    P.nu_base = nu;
    const float tau_min = 0.5f + 3.0f * 1.0e-8f;
    P.tau_shear_base = clampr_(tau_from_nu(nu), tau_min, 0.95f);
    P.tau_normal_base = clampr_(tau_from_nu(clamp_nu(nu)), tau_min, 0.95f);
    P.nu_rt0 = nu_rt;
    P.thermal_dt = 1.0f;
    P.thermal_update = 1;
    // ref: jni.cpp:799-809: fan forcing and its noise follow the flags in benchmark mode
    P.fan_on = !flag(AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING);
    P.fan_noise_amp = flag(AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_NOISE) ? 0.0f : 0.02f;
    // ref: jni.cpp:966-971 (heated cavity: nu / Pr), 986-993 (heated cavity: U^2 / (L dT))
    P.alpha = heated ? nu / pr : 0.035f;
    P.beta = heated ? (ref_speed * ref_speed) / rmaxf_(1.0e-6f, ref_len * dirichlet_temperature_span(g))
                    : ((9.81f * (1.0f / 300.0f)) * 20.0f) * (((1.0f / 20.0f) * (1.0f / 20.0f)) / 1.0f);
    P.thermal_bfecc = heated ? 1 : 0;  // ref: jni.cpp:871-873
    P.cylinder = g.preset == AERO_LBM_BENCHMARK_PRESET_CYLINDER_CROSSFLOW_2D ? 1 : 0;
    if (runtime) {
        // runtime mode (ref: jni.cpp:896-993): molecular viscosity nu0 and diffusivity of
        // air at 1 m cells and 0.05 s steps, tau_normal = 0.5 + 3 nu0, BFECC every second
        // step with dt 2, the Boussinesq beta above; cylinder and heated-cavity presets
        // belong to benchmark mode
        P.runtime_mode = 1;
        P.nu_base = nu_rt;
        P.tau_shear_base = clampr_(tau_from_nu(nu_rt), tau_min, 0.95f);
        P.tau_normal_base = clampr_(0.5f + 3.0f * (1.50e-5f * (1.0f / 20.0f) / (1.0f * 1.0f)), tau_min, 0.95f);
        P.alpha = rmaxf_(1.0e-8f, (1.50e-5f / 0.71f) * (1.0f / 20.0f) / (1.0f * 1.0f));
        P.beta = ((9.81f * (1.0f / 300.0f)) * 20.0f) * (((1.0f / 20.0f) * (1.0f / 20.0f)) / 1.0f);
        P.thermal_bfecc = 1;
        P.thermal_dt = 2.0f;
        P.cylinder = 0;
    }
    float sgs_c = 0.025f;
    P.nut_cap_ratio = 5.0f;
    P.g_axis = 1;
    P.g_sign = -1;
    const AeroLbmSolverOverrides& o = s.overrides;
    if (!std::isnan(o.sgs_constant)) sgs_c = o.sgs_constant;
    if (!std::isnan(o.sgs_nut_cap_ratio)) P.nut_cap_ratio = o.sgs_nut_cap_ratio;
    if (!std::isnan(o.boussinesq_beta)) P.beta = o.boussinesq_beta;
    if (!std::isnan(o.thermal_diffusivity)) P.alpha = o.thermal_diffusivity;
    if (o.gravity_axis >= 0 && o.gravity_axis <= 2) P.g_axis = o.gravity_axis;
    if (o.gravity_sign == 1 || o.gravity_sign == -1) P.g_sign = o.gravity_sign;
    P.sgs_c2 = sgs_c * sgs_c;
    P.kmax_speed = 0.57735026919f * 0.60f;
    P.faces_dev = nullptr;
    // faces whose rule reads other cells' post-collision values (symmetry, convective,
    // inherit -> convective) need the A-A face store; periodic sides never leave the domain
    long long off = 0;
    P.face_store_faces = 0;
    for (int f = 0; f < 6; ++f) {
        const int axis = f >> 1;
        const bool wraps = (f & 1) ? P.hydro_wrap_hi[axis] : P.hydro_wrap_lo[axis];
        const int k = P.face[f].hydro_kind;
        const bool needs = !wraps && (k == AERO_LBM_HYDRO_BOUNDARY_CONVECTIVE_OUTFLOW ||
                                      k == AERO_LBM_HYDRO_BOUNDARY_SYMMETRY ||
                                      k == AERO_LBM_HYDRO_BOUNDARY_PERIODIC ||
                                      (k == AERO_LBM_HYDRO_BOUNDARY_INHERIT_GAME && !P.convective_disabled));
        P.face_store_off[f] = off;
        P.face[f].store_off = off;
        if (needs) {
            P.face_store_faces |= 1 << f;
            const long long A = axis == 0 ? ny : nx;
            const long long B = axis == 2 ? ny : nz;
            off += 3LL * 9LL * A * B;
        }
    }
    P.face_store_floats = off;
    P.thermal_wrap_matches = 1;
    for (int a = 0; a < 3; ++a)
        if (P.thermal_wrap_lo[a] != P.hydro_wrap_lo[a] || P.thermal_wrap_hi[a] != P.hydro_wrap_hi[a])
            P.thermal_wrap_matches = 0;
    for (int side = 0; side < 2; ++side) {
        const bool wraps = side ? P.hydro_wrap_hi[0] : P.hydro_wrap_lo[0];
        const int k = P.face[side].hydro_kind, tk = P.face[side].thermal_kind;
        const bool plain_bounce = k == AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK ||
                                  (k == AERO_LBM_HYDRO_BOUNDARY_INHERIT_GAME && P.convective_disabled);
        const bool therm = tk == AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET || tk == AERO_LBM_THERMAL_BOUNDARY_HEAT_FLUX_NEUMANN;
        P.xface_bulk[side] = !wraps && (!plain_bounce || therm) ? 1 : 0;
    }
    for (int side = 0; side < 2; ++side) {
        const bool wraps = side ? P.hydro_wrap_hi[1] : P.hydro_wrap_lo[1];
        const int tk = P.face[2 + side].thermal_kind;
        const bool therm = tk == AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET || tk == AERO_LBM_THERMAL_BOUNDARY_HEAT_FLUX_NEUMANN;
        P.yface_bulk[side] =
            !wraps && !therm && P.face[2 + side].hydro_kind == AERO_LBM_HYDRO_BOUNDARY_SYMMETRY ? 1 : 0;
    }
    return true;
}

bool prepare_params(Context& c, StepParams& P) {
    if (!resolve_params(c.nx, c.ny, c.nz, rt().cfg.in_ch, P)) return false;
    if (!c.faces_valid || std::memcmp(c.faces_host, P.face, sizeof(P.face)) != 0) {
        // the previous copy may still be in flight from faces_host: wait before overwriting it
        if (!hip_ok(hipStreamSynchronize(rt().stream), "sync")) return false;
        std::memcpy(c.faces_host, P.face, sizeof(P.face));
        if (!hip_ok(hipMemcpyAsync(c.faces_dev, c.faces_host, sizeof(P.face), hipMemcpyHostToDevice, rt().stream),
                    "upload faces"))
            return false;
        c.faces_valid = true;
    }
    P.faces_dev = c.faces_dev;
    P.thermal_on = c.thermal_on ? 1 : 0;
    P.shell_on = c.shell_on ? 1 : 0;
    P.shell_layers = c.shell_on ? c.shell_layers : 0;
    P.shell_relax = c.shell_on ? c.shell_relax : 0.0f;
    // bulk cells index their neighbours without bounds checks, so the classes must
    // have been built with the periodic faces in force now
    if (c.has_payload) {
        const int w[6] = {P.hydro_wrap_lo[0], P.hydro_wrap_hi[0], P.hydro_wrap_lo[1],
                          P.hydro_wrap_hi[1], P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]};
        if (std::memcmp(w, c.class_wrap, sizeof(w)) != 0) {
            launch_classify(P, c.flags, rt().stream);
            std::memcpy(c.class_wrap, w, sizeof(w));
            c.mask_valid = false;
        }
        // the fast kernels' link mask also depends on face rules, sponge and face store
        int sig[kMaskSigLen];
        mask_signature(P, sig);
        if (c.mask && (!c.mask_valid || std::memcmp(sig, c.mask_sig, sizeof(sig)) != 0)) {
            if (!storage_rebuild_masks(c, P)) return false;
        }
    }
    return sync_params(c, P);
}

// keep the context's host and device copies of its step parameters equal to P
bool sync_params(Context& c, StepParams& P) {
    P.self_dev = c.params_dev;
    if (!c.params_valid || std::memcmp(&c.params_host, &P, sizeof(P)) != 0) {
        if (!hip_ok(hipStreamSynchronize(rt().stream), "sync")) return false;
        c.generation++;
        std::memcpy(&c.params_host, &P, sizeof(P));
        if (!hip_ok(hipMemcpyAsync(c.params_dev, &c.params_host, sizeof(P), hipMemcpyHostToDevice, rt().stream),
                    "upload step parameters"))
            return false;
        c.params_valid = true;
    }
    return true;
}

// ---- contexts ----
static void free_context_buffers(Context& c) {
    for (void* p : c.allocations) (void)hipFree(p);
    c.allocations.clear();
}

Context::~Context() { free_context_buffers(*this); }

bool Context::alloc(void** p, size_t bytes) {
    if (!hip_ok(hipMalloc(p, bytes), "hipMalloc")) return false;
    allocations.push_back(*p);
    return true;
}

// The reference's AERO_LBM_D3Q27_FP16_INPLACE switch (native/src/aero_lbm_jni.cpp:5730-5764):
// with the solver mode left at auto, a benchmark-mode context without a preset
// takes the FP16 in-place storage. Any value other than 0 / false / off turns it on.
static bool fp16_inplace_env() {
    const char* v = std::getenv("AERO_LBM_D3Q27_FP16_INPLACE");
    return v && std::strcmp(v, "0") && std::strcmp(v, "false") && std::strcmp(v, "FALSE") && std::strcmp(v, "off") &&
           std::strcmp(v, "OFF");
}

static Storage storage_for_mode(int mode, int variant) {
    // runtime (non-benchmark) mode runs on the exact FP32 kernels
    if (!rt().bench.enabled && (mode == AERO_LBM_REALTIME_SOLVER_AUTO || mode == AERO_LBM_REALTIME_SOLVER_CLASSIC_D3Q27))
        return Storage::F32Exact;
    if (mode == AERO_LBM_REALTIME_SOLVER_AUTO && rt().bench.preset == AERO_LBM_BENCHMARK_PRESET_NONE &&
        fp16_inplace_env())
        return Storage::F16;
    switch (mode) {
        case AERO_LBM_REALTIME_SOLVER_D3Q27_FP16_INPLACE_EXPERIMENTAL: return Storage::F16;
        case AERO_LBM_REALTIME_SOLVER_D3Q27_BF16_INPLACE: return Storage::BF16;
        case AERO_LBM_REALTIME_SOLVER_D3Q27_S16_INPLACE: return Storage::S16;
        case AERO_LBM_REALTIME_SOLVER_D3Q27_FP32_INPLACE: return Storage::F32AA;
        default: return variant == AERO_LBM_KERNEL_EXACT ? Storage::F32Exact : Storage::F32Fast;
    }
}

static bool allocate_context(Context& c, int nx, int ny, int nz) {
    c.nx = nx;
    c.ny = ny;
    c.nz = nz;
    c.cells = (size_t)nx * ny * nz;
    c.storage = storage_for_mode(rt().solver_mode, rt().overrides.kernel_variant);
    const size_t N = c.cells;
    bool ok = true;
    ok &= c.alloc((void**)&c.flags, N);
    ok &= c.alloc((void**)&c.ref_p, N * 4);
    ok &= c.alloc((void**)&c.ref_u, 3 * N * 4);
    ok &= c.alloc((void**)&c.ref_t, N * 4);
    ok &= c.alloc((void**)&c.src, N * 4);
    ok &= c.alloc((void**)&c.T[0], N * 4);
    ok &= c.alloc((void**)&c.T[1], N * 4);
    ok &= c.alloc((void**)&c.u_state, 3 * N * 4);
    ok &= c.alloc((void**)&c.out_dev, N * 4 * 4);
    ok &= c.alloc((void**)&c.sample_dev, 64);
    ok &= c.alloc((void**)&c.faces_dev, sizeof(FaceParams) * 6);
    ok &= c.alloc((void**)&c.params_dev, sizeof(StepParams));
    if (!ok) return false;
    if (!storage_alloc(c)) return false;
    hipStream_t st = rt().stream;
    ok &= hip_ok(hipMemsetAsync(c.flags, 0, N, st), "hipMemset");
    ok &= hip_ok(hipMemsetAsync(c.T[0], 0, N * 4, st), "hipMemset");
    ok &= hip_ok(hipMemsetAsync(c.T[1], 0, N * 4, st), "hipMemset");
    ok &= hip_ok(hipMemsetAsync(c.u_state, 0, 3 * N * 4, st), "hipMemset");
    ok &= hip_ok(hipMemsetAsync(c.ref_p, 0, N * 4, st), "hipMemset");
    ok &= hip_ok(hipMemsetAsync(c.ref_u, 0, 3 * N * 4, st), "hipMemset");
    ok &= hip_ok(hipMemsetAsync(c.ref_t, 0, N * 4, st), "hipMemset");
    ok &= hip_ok(hipMemsetAsync(c.src, 0, N * 4, st), "hipMemset");
    return ok;
}

Context* find_context(long long key) {
    auto it = rt().contexts.find(key);
    return it == rt().contexts.end() ? nullptr : it->second.get();
}

Context* get_context(long long key, int nx, int ny, int nz, bool create) {
    Context* c = find_context(key);
    if (c && (c->nx != nx || c->ny != ny || c->nz != nz)) {
        rt().contexts.erase(key);  // shape change: clear the context (ref: jni.cpp:8008-8015)
        c = nullptr;
    }
    if (!c && create) {
        if (!ensure_device()) return nullptr;
        auto nc = std::make_unique<Context>();
        if (!allocate_context(*nc, nx, ny, nz)) return nullptr;
        c = nc.get();
        rt().contexts[key] = std::move(nc);
    }
    return c;
}

CellFields cell_fields(const Context& c) {
    return CellFields{c.flags, c.ref_p, c.ref_u, c.ref_t, c.src, c.shell, c.shell_state, c.fan};
}

static bool host_any_nonzero(const float* p, size_t n, size_t stride, size_t off) {
    for (size_t i = 0; i < n; ++i)
        if (p[i * stride + off] != 0.0f) return true;
    return false;
}

void update_thermal_activity(Context& c, const StepParams& P) {
    if (c.thermal_on) return;
    const AeroLbmBenchmarkConfig& g = rt().bench;
    bool on = c.temperature_nonzero || g.initial_temperature != 0.0f;
    const AeroLbmBoundaryFaceConfig* faces[6] = {&g.x_min, &g.x_max, &g.y_min, &g.y_max, &g.z_min, &g.z_max};
    for (auto* f : faces) {
        if (f->thermal_kind == AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET && f->temperature != 0.0f) on = true;
        if (f->thermal_kind == AERO_LBM_THERMAL_BOUNDARY_HEAT_FLUX_NEUMANN && f->heat_flux != 0.0f) on = true;
    }
    if (P.source_on && c.payload_src_nonzero) on = true;
    if (c.payload_t_nonzero) on = true;
    c.thermal_on = on;
}

// The device half of a payload ingest, with the packet in c.payload_dev: cell
// flags, reference fields and fan planes from the packet, then the cell classes
// and the fast kernels' link masks.
static bool ingest_resident_payload(Context& c, const StepParams& P, bool fans) {
    hipStream_t st = rt().stream;
    if (c.storage == Storage::F32Exact && c.initialized && c.post_collision) {
        // the reference streams step n with step n's payload; run that stream now,
        // before the new payload replaces flags and reference pressure
        launch_exact_stream(P, cell_fields(c), c.f32[c.f_cur], c.f32[1 - c.f_cur], st);
        c.f_cur ^= 1;
        c.post_collision = false;
    }
    if (fans && !c.fan && !c.alloc((void**)&c.fan, 4 * c.cells * sizeof(float))) return false;
    c.fan_active = fans;
    launch_ingest(P, c.payload_dev, c.flags, c.ref_p, c.ref_u, c.ref_t, c.src, fans ? c.fan : nullptr, st);
    launch_classify(P, c.flags, st);
    if (!storage_rebuild_masks(c, P)) return false;
    const int w[6] = {P.hydro_wrap_lo[0], P.hydro_wrap_hi[0], P.hydro_wrap_lo[1],
                      P.hydro_wrap_hi[1], P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]};
    std::memcpy(c.class_wrap, w, sizeof(w));
    c.has_payload = true;
    return hip_ok(hipGetLastError(), "ingest");
}

static bool ensure_payload_buffer(Context& c) {
    return c.payload_dev || c.alloc((void**)&c.payload_dev, c.cells * (size_t)rt().cfg.in_ch * sizeof(float));
}

// A host packet: skipped when it equals the last one, else uploaded and ingested.
static bool ingest_payload(Context& c, const StepParams& P, const float* packet) {
    const size_t n = c.cells * (size_t)rt().cfg.in_ch;
    if (c.payload_host.size() == n && std::memcmp(c.payload_host.data(), packet, n * sizeof(float)) == 0) return true;
    const int ch = rt().cfg.in_ch;
    // fan cells in a payload while fan forcing is on: the force needs the fan planes
    const bool fans = P.fan_on && host_any_nonzero(packet, c.cells, ch, 1);
    if (fans && c.storage != Storage::F32Exact) {
        set_error("fan forcing runs on the exact FP32 kernels: select AERO_LBM_KERNEL_EXACT with the classic D3Q27 "
                  "solver mode, or set AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING");
        return false;
    }
    c.payload_host.assign(packet, packet + n);
    if (ch > 9 && host_any_nonzero(packet, c.cells, ch, 9)) c.payload_src_nonzero = true;
    if (ch > 10 && host_any_nonzero(packet, c.cells, ch, 10)) c.payload_t_nonzero = true;
    if (!ensure_payload_buffer(c)) return false;
    if (!hip_ok(hipMemcpyAsync(c.payload_dev, packet, n * sizeof(float), hipMemcpyHostToDevice, rt().stream),
                "upload payload"))
        return false;
    c.timing_upload_bytes += n * sizeof(float);
    return ingest_resident_payload(c, P, fans);
}

// A packet already in device memory whose fan, source and temperature channels are
// zero (the wind-tunnel layer's packets): copied into the context and ingested.
static bool ingest_device_packet(Context& c, const StepParams& P, const float* packet_dev) {
    c.payload_host.clear();  // the host copy no longer describes the resident payload
    if (!ensure_payload_buffer(c)) return false;
    const size_t bytes = c.cells * (size_t)rt().cfg.in_ch * sizeof(float);
    if (!hip_ok(hipMemcpyAsync(c.payload_dev, packet_dev, bytes, hipMemcpyDeviceToDevice, rt().stream),
                "copy payload"))
        return false;
    return ingest_resident_payload(c, P, false);
}

// Per-step switches that depend on the context's step counter: the temperature
// update stride (ref: jni.cpp:2191-2196) and the fan forcing's noise seed.
void step_switches(const Context& c, StepParams& P) {
    const uint64_t stride = P.runtime_mode ? 2u : 1u;
    P.thermal_update = (c.steps % stride) == 0 ? 1 : 0;
    P.fan_on = (P.fan_on && c.fan_active) ? 1 : 0;
    P.tick = P.fan_on ? (unsigned long long)c.steps : 0ull;
}

// One solver step of a context. packet == nullptr means "resident payload";
// packet_on_device says the packet is in device memory.
bool step_context(Context& c, const float* packet, float* out, int out_ch, float scale, bool packet_on_device) {
    c.generation++;  // stepped outside any batch
    StepParams P;
    if (!prepare_params(c, P)) return false;
    if (packet) {
        if (!(packet_on_device ? ingest_device_packet(c, P, packet) : ingest_payload(c, P, packet))) return false;
    } else if (!c.has_payload) {
        set_error("cached step requested before any payload was uploaded");
        return false;
    }
    update_thermal_activity(c, P);
    P.thermal_on = c.thermal_on ? 1 : 0;
    step_switches(c, P);
    if (!sync_params(c, P)) return false;  // the step runs with (and records) the thermal switch just decided
    if (!storage_step(c, P)) return false;
    c.steps += 1;
    if (out) {
        if (!storage_readout(c, P, 1, out_ch, scale, c.out_dev)) return false;
        const size_t bytes = c.cells * (size_t)out_ch * sizeof(float);
        if (bytes > c.cells * 16) {
            // wider outputs than 4 channels: use a temporary device buffer
            float* tmp = nullptr;
            if (!hip_ok(hipMalloc(&tmp, bytes), "hipMalloc")) return false;
            storage_readout(c, P, 1, out_ch, scale, tmp);
            const bool ok = hip_ok(hipMemcpyAsync(out, tmp, bytes, hipMemcpyDeviceToHost, rt().stream), "readback") &&
                            hip_ok(hipStreamSynchronize(rt().stream), "sync");
            hipFree(tmp);
            return ok;
        }
        if (!hip_ok(hipMemcpyAsync(out, c.out_dev, bytes, hipMemcpyDeviceToHost, rt().stream), "readback"))
            return false;
        if (!hip_ok(hipStreamSynchronize(rt().stream), "sync")) return false;
    }
    return hip_ok(hipGetLastError(), "step");
}

void record_timing(double total_ms, double solver_ms, double readback_ms, double copy_ms) {
    Timing& t = rt().timing;
    t.ticks += 1;
    t.last_total = total_ms;
    t.last_solver = solver_ms;
    t.last_readback = readback_ms;
    t.last_copy = copy_ms;
    t.sum_total += total_ms;
    t.sum_solver += solver_ms;
    t.sum_readback += readback_ms;
    t.sum_copy += copy_ms;
}

}  // namespace aero

using namespace aero;
using Clock = std::chrono::steady_clock;

static double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

extern "C" {

AERO_LBM_CAPI_EXPORT int aero_lbm_init_rect(int nx, int ny, int nz, int input_channels, int output_channels) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    RuntimeState& s = rt();
    s.contexts.clear();
    s.timing = Timing{};
    if (nx <= 0 || ny <= 0 || nz <= 0 || input_channels < 9 || output_channels < 4) {
        s.cfg = RuntimeConfig{};
        set_error("aero_lbm_init_rect: invalid dimensions or channel counts");
        return 0;
    }
    s.cfg.grid_size = (nx == ny && ny == nz) ? nx : 0;
    s.cfg.nx = nx;
    s.cfg.ny = ny;
    s.cfg.nz = nz;
    s.cfg.in_ch = input_channels;
    s.cfg.out_ch = output_channels;
    s.cfg.initialized = true;
    if (!ensure_device()) {
        s.runtime_info = "hip|unavailable (" + s.last_error + ")";
        s.cfg.initialized = false;
        return 0;
    }
    s.runtime_info = "hip|cumulant-d3q27+sgs+bouss+fp16/bf16-shifted-aa:" + s.device_name + " " + s.device_arch;
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_init(int grid_size, int input_channels, int output_channels) {
    return aero_lbm_init_rect(grid_size, grid_size, grid_size, input_channels, output_channels);
}

static int step_impl(const float* packet, int nx, int ny, int nz, long long key, float scale, float* out,
                     bool packet_on_device = false) {
    auto t0 = Clock::now();
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!rt().cfg.initialized) {
        set_error("native_step_raw_dims: runtime not initialized");
        return 0;
    }
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        set_error("native_step_raw_dims: invalid dimensions");
        return 0;
    }
    if (!std::isfinite(scale)) scale = 1.0f;
    Context* c = get_context(key, nx, ny, nz, packet != nullptr);
    if (!c) {
        if (rt().last_error.empty()) set_error("cached solver step requested before context initialization");
        return 0;
    }
    const bool ok = step_context(*c, packet, out, rt().cfg.out_ch, scale, packet_on_device);
    const double total = ms_since(t0);
    record_timing(total, total, 0.0, 0.0);
    return ok ? 1 : 0;
}

}  // extern "C"

namespace aero {
bool context_resident(long long key, int nx, int ny, int nz) {
    std::lock_guard<std::mutex> lk(rt().mu);
    const Context* c = find_context(key);
    return c && c->has_payload && c->nx == nx && c->ny == ny && c->nz == nz;
}

int step_rect_device_packet(const float* packet_dev, int nx, int ny, int nz, long long key, float scale, float* out) {
    return step_impl(packet_dev, nx, ny, nz, key, scale, out, packet_dev != nullptr);
}
}  // namespace aero

extern "C" {

AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect(const float* packet, int nx, int ny, int nz, long long key, float* out) {
    if (!packet) {
        set_error("native_step_raw_dims: missing packet");
        return 0;
    }
    return step_impl(packet, nx, ny, nz, key, 1.0f, out);
}

AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_scaled(const float* packet, int nx, int ny, int nz, long long key,
                                                   float scale, float* out) {
    if (!packet) {
        set_error("native_step_raw_dims: missing packet");
        return 0;
    }
    return step_impl(packet, nx, ny, nz, key, scale, out);
}

AERO_LBM_CAPI_EXPORT int aero_lbm_step(const float* packet, int grid_size, long long key, float* out) {
    return aero_lbm_step_rect(packet, grid_size, grid_size, grid_size, key, out);
}

AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_cached(int nx, int ny, int nz, long long key, float* out) {
    return step_impl(nullptr, nx, ny, nz, key, 1.0f, out);
}

AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_cached_scaled(int nx, int ny, int nz, long long key, float scale,
                                                          float* out) {
    return step_impl(nullptr, nx, ny, nz, key, scale, out);
}

AERO_LBM_CAPI_EXPORT int aero_lbm_run_cached_steps_timed(int nx, int ny, int nz, long long key, int steps,
                                                         float* out_ms) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!out_ms || steps <= 0) {
        set_error("run_cached_steps_timed: invalid arguments");
        return 0;
    }
    Context* c = get_context(key, nx, ny, nz, false);
    if (!c || !c->has_payload) {
        set_error("run_cached_steps_timed: context has no resident payload");
        return 0;
    }
    hipEvent_t e0, e1;
    if (!hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate")) return 0;
    bool ok = hip_ok(hipEventRecord(e0, rt().stream), "hipEventRecord");
    for (int i = 0; ok && i < steps; ++i) ok = step_context(*c, nullptr, nullptr, 0, 1.0f);
    ok = ok && hip_ok(hipEventRecord(e1, rt().stream), "hipEventRecord") &&
         hip_ok(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.0f;
    if (ok) ok = hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *out_ms = ms;
    return ok ? 1 : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_has_context(long long key) {
    std::lock_guard<std::mutex> lk(rt().mu);
    return find_context(key) ? 1 : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_context_realtime_cached_initialized(long long key) {
    std::lock_guard<std::mutex> lk(rt().mu);
    Context* c = find_context(key);
    return (c && c->has_payload && c->initialized) ? 1 : 0;
}

// A stepped context is read with the parameters of its last step (the face rules
// its stored state was streamed with), whatever the global configuration is now.
static bool readout_params(Context& c, StepParams& P) {
    if (c.params_valid) {
        P = c.params_host;
        return true;
    }
    return prepare_params(c, P);
}

static int readout_impl(int nx, int ny, int nz, long long key, int stride, float* out, int value_count) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!rt().cfg.initialized || !out || stride <= 0 || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    const int sx = (nx + stride - 1) / stride, sy = (ny + stride - 1) / stride, sz = (nz + stride - 1) / stride;
    if (value_count != sx * sy * sz * 4) {
        set_error("extract_flow_atlas: value_count mismatch");
        return 0;
    }
    Context* c = get_context(key, nx, ny, nz, false);
    if (!c || !c->initialized) {
        set_error("extract_flow_atlas: unknown or uninitialized context");
        return 0;
    }
    StepParams P;
    if (!readout_params(*c, P)) return 0;
    if (!storage_readout(*c, P, stride, 4, 1.0f, c->out_dev)) return 0;
    const bool ok = hip_ok(hipMemcpyAsync(out, c->out_dev, (size_t)value_count * 4, hipMemcpyDeviceToHost,
                                          rt().stream), "readback") &&
                    hip_ok(hipStreamSynchronize(rt().stream), "sync");
    return ok ? 1 : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_extract_flow_atlas_rect(int nx, int ny, int nz, long long key, int stride,
                                                          float* out, int value_count) {
    return readout_impl(nx, ny, nz, key, stride, out, value_count);
}

AERO_LBM_CAPI_EXPORT int aero_lbm_get_flow_state_rect(int nx, int ny, int nz, long long key, float* out) {
    const long long cells = (long long)nx * ny * nz;
    if (cells <= 0 || cells > 0x7fffffff / 4) return 0;
    return readout_impl(nx, ny, nz, key, 1, out, (int)(cells * 4));
}

AERO_LBM_CAPI_EXPORT int aero_lbm_sample_flow_point_rect(int nx, int ny, int nz, long long key, int x, int y, int z,
                                                         float* out) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!rt().cfg.initialized || !out || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0;
    Context* c = get_context(key, nx, ny, nz, false);
    if (!c || !c->initialized) return 0;
    StepParams P;
    if (!readout_params(*c, P)) return 0;
    if (!storage_sample(*c, P, x, y, z, c->sample_dev)) return 0;
    return (hip_ok(hipMemcpyAsync(out, c->sample_dev, 16, hipMemcpyDeviceToHost, rt().stream), "readback") &&
            hip_ok(hipStreamSynchronize(rt().stream), "sync"))
               ? 1
               : 0;
}

static bool copy_temperature(Context& c, float* out) {
    return hip_ok(hipMemcpyAsync(out, c.T[c.t_cur], c.cells * 4, hipMemcpyDeviceToHost, rt().stream), "readback") &&
           hip_ok(hipStreamSynchronize(rt().stream), "sync");
}

AERO_LBM_CAPI_EXPORT int aero_lbm_get_temperature_state_rect(int nx, int ny, int nz, long long key, float* out) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!rt().cfg.initialized || !out || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    Context* c = get_context(key, nx, ny, nz, false);
    if (!c) return 0;
    return copy_temperature(*c, out) ? 1 : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_sample_temperature_point_rect(int nx, int ny, int nz, long long key, int x, int y,
                                                                int z, float* out) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!rt().cfg.initialized || !out || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0;
    Context* c = get_context(key, nx, ny, nz, false);
    if (!c) return 0;
    const size_t cell = ((size_t)x * ny + y) * nz + z;
    return (hip_ok(hipMemcpyAsync(out, c->T[c->t_cur] + cell, 4, hipMemcpyDeviceToHost, rt().stream), "readback") &&
            hip_ok(hipStreamSynchronize(rt().stream), "sync"))
               ? 1
               : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_set_temperature_state_rect(int nx, int ny, int nz, long long key, const float* t) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!rt().cfg.initialized || !t || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    Context* c = get_context(key, nx, ny, nz, true);
    if (!c) return 0;
    // assign_temperature_state clamps to [-1, 1] (ref: jni.cpp:8026-8033)
    std::vector<float> tmp(c->cells);
    for (size_t i = 0; i < c->cells; ++i) {
        const float v = t[i];
        const float m = (-1.0f < v) ? v : -1.0f;
        tmp[i] = (m < 1.0f) ? m : 1.0f;
        if (tmp[i] != 0.0f) c->temperature_nonzero = true;
    }
    c->generation++;
    return (hip_ok(hipMemcpyAsync(c->T[c->t_cur], tmp.data(), c->cells * 4, hipMemcpyHostToDevice, rt().stream),
                   "upload") &&
            hip_ok(hipStreamSynchronize(rt().stream), "sync"))
               ? 1
               : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_copy_flow_temperature_subrect(int nx, int ny, int nz, long long key, int ox, int oy,
                                                                int oz, int cnx, int cny, int cnz, float* out_flow,
                                                                float* out_t) {
    if (!out_flow || !out_t || nx <= 0 || ny <= 0 || nz <= 0 || cnx <= 0 || cny <= 0 || cnz <= 0) return 0;
    if (ox < 0 || oy < 0 || oz < 0 || ox + cnx > nx || oy + cny > ny || oz + cnz > nz) return 0;
    const size_t cells = (size_t)nx * ny * nz;
    std::vector<float> flow(cells * 4), temp(cells);
    if (!aero_lbm_get_flow_state_rect(nx, ny, nz, key, flow.data())) return 0;
    if (!aero_lbm_get_temperature_state_rect(nx, ny, nz, key, temp.data())) return 0;
    std::vector<uint8_t> flags(cells);
    {
        std::lock_guard<std::mutex> lk(rt().mu);
        Context* c = find_context(key);
        if (!c) return 0;
        if (!hip_ok(hipMemcpy(flags.data(), c->flags, cells, hipMemcpyDeviceToHost), "readback")) return 0;
    }
    size_t dst = 0;
    for (int x = 0; x < cnx; ++x)
        for (int y = 0; y < cny; ++y)
            for (int z = 0; z < cnz; ++z, ++dst) {
                const size_t src = ((size_t)(ox + x) * ny + (oy + y)) * nz + (oz + z);
                const bool solid = flags[src] & kCellSolid;
                float t = temp[src];
                t = (-1.0f < t) ? t : -1.0f;
                t = (t < 1.0f) ? t : 1.0f;
                out_t[dst] = solid ? 0.0f : (std::isfinite(t) ? t : 0.0f);
                for (int k = 0; k < 4; ++k) out_flow[dst * 4 + k] = flow[src * 4 + k];
            }
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_get_last_force(long long key, float* fx, float* fy, float* fz) {
    // drag integral of the last step: produced for the cylinder preset, zero otherwise
    // (ref: jni.cpp:2452-2492, ctx.last_force)
    std::lock_guard<std::mutex> lk(rt().mu);
    Context* c = find_context(key);
    if (!c || !fx || !fy || !fz) return 0;
    float f[3] = {0.0f, 0.0f, 0.0f};
    if (c->force_valid && c->params_valid && c->params_host.cylinder) {
        const int blocks = (int)((c->cells + 255) / 256);
        if (!hip_ok(hipMemcpyAsync(f, c->force_part + 3 * (size_t)blocks, sizeof(f), hipMemcpyDeviceToHost,
                                   rt().stream),
                    "readback") ||
            !hip_ok(hipStreamSynchronize(rt().stream), "sync"))
            return 0;
    }
    *fx = f[0];
    *fy = f[1];
    *fz = f[2];
    return 1;
}

AERO_LBM_CAPI_EXPORT void aero_lbm_release_context(long long key) {
    std::lock_guard<std::mutex> lk(rt().mu);
    rt().contexts.erase(key);
}

AERO_LBM_CAPI_EXPORT void aero_lbm_shutdown(void) {
    std::lock_guard<std::mutex> lk(rt().mu);
    rt().contexts.clear();
    rt().cfg = RuntimeConfig{};
    rt().runtime_info.clear();
    rt().timing = Timing{};
}

AERO_LBM_CAPI_EXPORT const char* aero_lbm_runtime_info(void) {
    return rt().runtime_info.empty() ? "uninitialized" : rt().runtime_info.c_str();
}

AERO_LBM_CAPI_EXPORT const char* aero_lbm_last_error(void) { return rt().last_error.c_str(); }

AERO_LBM_CAPI_EXPORT const char* aero_lbm_timing_info(void) {
    static std::string text;
    const Timing& t = rt().timing;
    if (t.ticks == 0) {
        text = "ticks=0";
        return text.c_str();
    }
    std::ostringstream o;
    o.setf(std::ios::fixed);
    o.precision(3);
    const double inv = 1.0 / (double)t.ticks;
    o << "ticks=" << t.ticks << " path=hip last_ms(copy=" << t.last_copy << ",solver=" << t.last_solver
      << ",readback=" << t.last_readback << ",total=" << t.last_total << ") avg_ms(copy=" << t.sum_copy * inv
      << ",solver=" << t.sum_solver * inv << ",readback=" << t.sum_readback * inv << ",total=" << t.sum_total * inv
      << ")";
    text = o.str();
    return text.c_str();
}

AERO_LBM_CAPI_EXPORT const char* aero_lbm_memory_info(void) {
    static std::string text;
    std::lock_guard<std::mutex> lk(rt().mu);
    size_t bytes = 0;
    for (auto& kv : rt().contexts) bytes += context_device_bytes(*kv.second);
    size_t free_b = 0, total_b = 0;
    if (rt().device_ready) (void)hipMemGetInfo(&free_b, &total_b);
    std::ostringstream o;
    o << "contexts=" << rt().contexts.size() << " context_device_bytes=" << bytes << " device_free=" << free_b
      << " device_total=" << total_b;
    text = o.str();
    return text.c_str();
}

AERO_LBM_CAPI_EXPORT void aero_lbm_reset_timing(void) { rt().timing = Timing{}; }

AERO_LBM_CAPI_EXPORT void aero_lbm_set_realtime_solver_mode(int mode) {
    switch (mode) {
        case AERO_LBM_REALTIME_SOLVER_AUTO:
        case AERO_LBM_REALTIME_SOLVER_CLASSIC_D3Q27:
        case AERO_LBM_REALTIME_SOLVER_COMPACT_EXPERIMENTAL:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_FP16_INPLACE_EXPERIMENTAL:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_BF16_INPLACE:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_S16_INPLACE:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_FP32_INPLACE:
            rt().solver_mode = mode;
            break;
        default:
            rt().solver_mode = AERO_LBM_REALTIME_SOLVER_AUTO;
            break;
    }
}

AERO_LBM_CAPI_EXPORT int aero_lbm_finish(void) {
    clear_error();
    if (!rt().device_ready) return 1;
    return hip_ok(hipStreamSynchronize(rt().stream), "hipStreamSynchronize") ? 1 : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_get_timing_snapshot(AeroLbmTimingSnapshot* o) {
    if (!o) return 0;
    const Timing& t = rt().timing;
    const double inv = t.ticks == 0 ? 0.0 : 1.0 / (double)t.ticks;
    o->ticks = t.ticks;
    o->last_payload_copy_ms = t.last_copy;
    o->last_solver_ms = t.last_solver;
    o->last_readback_ms = t.last_readback;
    o->last_total_ms = t.last_total;
    o->avg_payload_copy_ms = t.sum_copy * inv;
    o->avg_solver_ms = t.sum_solver * inv;
    o->avg_readback_ms = t.sum_readback * inv;
    o->avg_total_ms = t.sum_total * inv;
    return 1;
}

AERO_LBM_CAPI_EXPORT void aero_lbm_benchmark_default_config(AeroLbmBenchmarkConfig* out) {
    if (out) *out = default_benchmark_config();
}

AERO_LBM_CAPI_EXPORT int aero_lbm_benchmark_default_preset_config(int preset, AeroLbmBenchmarkConfig* out) {
    if (!out || !valid_preset(preset)) return 0;
    AeroLbmBenchmarkConfig c{};
    apply_preset(c, preset);
    sanitize_benchmark_config(c);
    *out = c;
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_benchmark_set_config(const AeroLbmBenchmarkConfig* config) {
    if (!config) return 0;
    if (config->abi_version != AERO_LBM_BENCHMARK_ABI_VERSION) return 0;
    if (config->struct_size != sizeof(AeroLbmBenchmarkConfig)) return 0;
    std::lock_guard<std::mutex> lk(rt().mu);
    rt().bench = *config;
    sanitize_benchmark_config(rt().bench);
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_benchmark_get_config(AeroLbmBenchmarkConfig* out) {
    if (!out) return 0;
    *out = rt().bench;
    out->abi_version = AERO_LBM_BENCHMARK_ABI_VERSION;
    out->struct_size = sizeof(AeroLbmBenchmarkConfig);
    return 1;
}

AERO_LBM_CAPI_EXPORT void aero_lbm_benchmark_reset_config(void) { rt().bench = default_benchmark_config(); }

AERO_LBM_CAPI_EXPORT const char* aero_lbm_benchmark_info(void) {
    static std::string text;
    static const char* hydro_names[] = {"inherit_game", "periodic", "bounce_back", "moving_wall",
                                        "velocity", "pressure", "convective_outflow", "symmetry"};
    static const char* preset_names[] = {"none", "taylor_green_3d", "lid_driven_cavity_2d", "lid_driven_cavity_3d",
                                         "cylinder_crossflow_2d", "heated_cavity_2d", "heated_cavity_3d"};
    const AeroLbmBenchmarkConfig& c = rt().bench;
    std::ostringstream o;
    o << "enabled=" << c.enabled << " preset=" << preset_names[valid_preset(c.preset) ? c.preset : 0] << " flags=0x"
      << std::hex << c.flags << std::dec << " Re=" << c.reynolds_number << " Ra=" << c.rayleigh_number
      << " Pr=" << c.prandtl_number << " Ma=" << c.mach_number << " boundaries={x-:"
      << hydro_names[c.x_min.hydrodynamic_kind] << ",x+:" << hydro_names[c.x_max.hydrodynamic_kind]
      << ",y-:" << hydro_names[c.y_min.hydrodynamic_kind] << ",y+:" << hydro_names[c.y_max.hydrodynamic_kind]
      << ",z-:" << hydro_names[c.z_min.hydrodynamic_kind] << ",z+:" << hydro_names[c.z_max.hydrodynamic_kind] << "}";
    text = o.str();
    return text.c_str();
}

AERO_LBM_CAPI_EXPORT void aero_lbm_default_solver_overrides(AeroLbmSolverOverrides* out) {
    if (out) *out = default_overrides();
}

AERO_LBM_CAPI_EXPORT int aero_lbm_set_solver_overrides(const AeroLbmSolverOverrides* o) {
    if (!o || o->struct_size != sizeof(AeroLbmSolverOverrides)) {
        set_error("set_solver_overrides: null or wrong struct_size");
        return 0;
    }
    if (o->kernel_variant != AERO_LBM_KERNEL_FAST && o->kernel_variant != AERO_LBM_KERNEL_EXACT) {
        set_error("set_solver_overrides: unknown kernel_variant");
        return 0;
    }
    std::lock_guard<std::mutex> lk(rt().mu);
    rt().overrides = *o;
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_get_solver_overrides(AeroLbmSolverOverrides* out) {
    if (!out) return 0;
    *out = rt().overrides;
    return 1;
}

}  // extern "C"
