/*
 * D3Q27 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, always as the checker (or as the CPU baseline being timed),
 * never as the product path. libaero_lbm never links or calls it.
 *
 * It restates the reference's single-threaded CPU solver step
 * (native/src/aero_lbm_jni.cpp:1180-2581, driver 8480-8507) in benchmark mode
 * twice: FP32 in the reference's operation order (prefix o32_) and FP64
 * (prefix o64_). Build: oracle/Makefile (-O2 -ffp-contract=off, no fast-math,
 * x86-64 SSE2 so float arithmetic is IEEE binary32 without contraction).
 *
 * Parity pinning (the reference itself cannot be compiled here because
 * aero_lbm_jni.cpp includes <jni.h>, which this image does not have): see the
 * tests in tests/test_oracle_*.py -- lattice tables and feq against the
 * reference's own hydro_core.cpp compiled into oracle/_ref, closed forms of the
 * cumulant collision (tau = 1 gives the factorised equilibrium; all 27
 * post-collision central moments follow the relaxation/closure relations),
 * Guo-source moment identities, conservation, equilibrium invariance, no-slip
 * and the reference debug harness thresholds.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/aero_lbm_capi.h"

#define ORACLE_API __attribute__((visibility("default")))

#ifdef _OPENMP
#include <omp.h>
#endif

/* Worker threads for the per-cell loops (1 = the reference's single thread).
 * Every loop writes disjoint cells, so results do not depend on the count. */
ORACLE_API void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* hydro_core.h:122-205: q = (ix*3 + iy)*3 + iz, c = i - 1, opp(q) = 26 - q */
static const int kCX[27] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
static const int kCY[27] = {-1, -1, -1, 0, 0, 0, 1, 1, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1};
static const int kCZ[27] = {-1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1};
static const int kOPP[27] = {26, 25, 24, 23, 22, 21, 20, 19, 18, 17, 16, 15, 14, 13,
                             12, 11, 10, 9,  8,  7,  6,  5,  4,  3,  2,  1,  0};

/* Sanitise a benchmark config like the reference does on set_config
 * (jni.cpp:409-441). */
static float fin_or(float v, float fb) { return isfinite(v) ? v : fb; }
static void sanitize_face(AeroLbmBoundaryFaceConfig* f) {
    if (f->hydrodynamic_kind < 0 || f->hydrodynamic_kind > 7) f->hydrodynamic_kind = 0;
    if (f->thermal_kind < 0 || f->thermal_kind > 5) f->thermal_kind = 0;
    for (int i = 0; i < 3; ++i) f->velocity[i] = fin_or(f->velocity[i], 0.0f);
    f->pressure = fin_or(f->pressure, 0.0f);
    f->temperature = fin_or(f->temperature, 0.0f);
    f->heat_flux = fin_or(f->heat_flux, 0.0f);
}
ORACLE_API void oracle_sanitize_config(AeroLbmBenchmarkConfig* c) {
    c->abi_version = AERO_LBM_BENCHMARK_ABI_VERSION;
    c->struct_size = sizeof(AeroLbmBenchmarkConfig);
    c->flags &= 0xFFu;
    c->enabled = c->enabled != 0 ? 1 : 0;
    if (c->preset < 0 || c->preset > 6) c->preset = 0;
    c->reynolds_number = fin_or(c->reynolds_number, 100.0f);
    c->rayleigh_number = fin_or(c->rayleigh_number, 1.0e5f);
    c->prandtl_number = fin_or(c->prandtl_number, 0.71f);
    float m = fin_or(c->mach_number, 0.05f);
    c->mach_number = m < 1.0e-4f ? 1.0e-4f : (m > 0.30f ? 0.30f : m);
    c->reference_density = fin_or(c->reference_density, 1.0f);
    c->reference_temperature = fin_or(c->reference_temperature, 0.0f);
    float rl = fin_or(c->reference_length, 1.0f);
    c->reference_length = rl < 1.0e-6f ? 1.0e-6f : rl;
    for (int i = 0; i < 3; ++i) {
        c->body_force[i] = fin_or(c->body_force[i], 0.0f);
        c->initial_velocity[i] = fin_or(c->initial_velocity[i], 0.0f);
    }
    c->gravity[0] = fin_or(c->gravity[0], 0.0f);
    c->gravity[1] = fin_or(c->gravity[1], -1.0f);
    c->gravity[2] = fin_or(c->gravity[2], 0.0f);
    c->initial_pressure = fin_or(c->initial_pressure, 0.0f);
    c->initial_temperature = fin_or(c->initial_temperature, 0.0f);
    sanitize_face(&c->x_min); sanitize_face(&c->x_max);
    sanitize_face(&c->y_min); sanitize_face(&c->y_max);
    sanitize_face(&c->z_min); sanitize_face(&c->z_max);
}

/* ---- FP32 instance ---- */
#define R float
#define SFX(n) o32_##n
#define K(x) ((float)(x##f))
#define RSQRT_(x) sqrtf(x)
#define RFLOOR_(x) floorf(x)
#define RFABS_(x) fabsf(x)
#define RISFINITE_(x) isfinite(x)
#include "d3q27_oracle_impl.h"
#undef R
#undef SFX
#undef K
#undef RSQRT_
#undef RFLOOR_
#undef RFABS_
#undef RISFINITE_

/* ---- FP64 instance ---- */
#define R double
#define SFX(n) o64_##n
#define K(x) ((double)(x))
#define RSQRT_(x) sqrt(x)
#define RFLOOR_(x) floor(x)
#define RFABS_(x) fabs(x)
#define RISFINITE_(x) isfinite(x)
#include "d3q27_oracle_impl.h"
