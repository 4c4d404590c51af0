/*
 * aero_lbm C ABI -- MI355X (gfx950) implementation of the local-airflow solver.
 *
 * Drop-in boundary: every declaration below that carries a "ref:" citation has
 * the same name, argument meaning, return convention (1 = ok, 0 = failure) and
 * struct layout as the reference's native/include/aero_lbm_capi.h, so the
 * reference's JNI layer, ctypes tools and Java callers bind unchanged.
 * Declarations marked "extension" have no counterpart in the reference.
 *
 * Struct layouts are fixed by static_assert in the implementation
 * (x86-64: AeroLbmBenchmarkConfig 320 B, AeroLbmBoundaryFaceConfig 32 B,
 * AeroLbmTimingSnapshot 72 B, AeroLbmWorldDelta 48 B,
 * AeroLbmMesoscaleConfig 96 B, AeroLbmMesoscaleTransport 20 B).
 */
#ifndef AERO_LBM_CAPI_H
#define AERO_LBM_CAPI_H

#include <stdint.h>

#define AERO_LBM_CAPI_EXPORT __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* ref: native/include/aero_lbm_capi.h:16-26 */
enum {
    AERO_LBM_BENCHMARK_ABI_VERSION = 1,
    AERO_LBM_MESOSCALE_ABI_VERSION = 3,
    AERO_LBM_SIMULATION_ABI_VERSION = 1,
    AERO_LBM_MESOSCALE_FORCING_CHANNELS = 25,
    AERO_LBM_MESOSCALE_STATE_CHANNELS = 6,
    AERO_LBM_SIMULATION_FLOW_STATE_CHANNELS = 4,
    AERO_LBM_SIMULATION_PACKED_ATLAS_CHANNELS = 4,
    AERO_LBM_SIMULATION_PLAYER_PROBE_CHANNELS = 6,
    AERO_LBM_SIMULATION_BRICK_RUNTIME_STATUS_FIELDS = 8
};

/* ref: native/include/aero_lbm_capi.h:28-33; values 4-6 are extensions
 * (BF16, FP32 and 16-bit fixed-point in-place storage). */
enum {
    AERO_LBM_REALTIME_SOLVER_AUTO = 0,
    AERO_LBM_REALTIME_SOLVER_CLASSIC_D3Q27 = 1,
    AERO_LBM_REALTIME_SOLVER_COMPACT_EXPERIMENTAL = 2,
    AERO_LBM_REALTIME_SOLVER_D3Q27_FP16_INPLACE_EXPERIMENTAL = 3,
    AERO_LBM_REALTIME_SOLVER_D3Q27_BF16_INPLACE = 4,
    AERO_LBM_REALTIME_SOLVER_D3Q27_FP32_INPLACE = 5, /* extension: FP32 unshifted A-A */
    /* extension: 16-bit signed fixed point of f - w_q, range per direction class
     * (rest 0.25, axis 1.0, face diagonal 1.25, corner 1.5 times w_q) */
    AERO_LBM_REALTIME_SOLVER_D3Q27_S16_INPLACE = 6
};

/* ref: native/include/aero_lbm_capi.h:35-43 */
typedef enum AeroLbmBenchmarkPreset {
    AERO_LBM_BENCHMARK_PRESET_NONE = 0,
    AERO_LBM_BENCHMARK_PRESET_TAYLOR_GREEN_3D = 1,
    AERO_LBM_BENCHMARK_PRESET_LID_DRIVEN_CAVITY_2D = 2,
    AERO_LBM_BENCHMARK_PRESET_LID_DRIVEN_CAVITY_3D = 3,
    AERO_LBM_BENCHMARK_PRESET_CYLINDER_CROSSFLOW_2D = 4,
    AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_2D = 5,
    AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_3D = 6
} AeroLbmBenchmarkPreset;

/* ref: native/include/aero_lbm_capi.h:45-54 */
typedef enum AeroLbmHydrodynamicBoundaryKind {
    AERO_LBM_HYDRO_BOUNDARY_INHERIT_GAME = 0,
    AERO_LBM_HYDRO_BOUNDARY_PERIODIC = 1,
    AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK = 2,
    AERO_LBM_HYDRO_BOUNDARY_MOVING_WALL = 3,
    AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET = 4,
    AERO_LBM_HYDRO_BOUNDARY_PRESSURE_DIRICHLET = 5,
    AERO_LBM_HYDRO_BOUNDARY_CONVECTIVE_OUTFLOW = 6,
    AERO_LBM_HYDRO_BOUNDARY_SYMMETRY = 7
} AeroLbmHydrodynamicBoundaryKind;

/* ref: native/include/aero_lbm_capi.h:56-63 */
typedef enum AeroLbmThermalBoundaryKind {
    AERO_LBM_THERMAL_BOUNDARY_INHERIT_GAME = 0,
    AERO_LBM_THERMAL_BOUNDARY_DISABLED = 1,
    AERO_LBM_THERMAL_BOUNDARY_PERIODIC = 2,
    AERO_LBM_THERMAL_BOUNDARY_ADIABATIC = 3,
    AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET = 4,
    AERO_LBM_THERMAL_BOUNDARY_HEAT_FLUX_NEUMANN = 5
} AeroLbmThermalBoundaryKind;

/* ref: native/include/aero_lbm_capi.h:65-74 */
typedef enum AeroLbmBenchmarkFlags {
    AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING = 1u << 0,
    AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_NOISE = 1u << 1,
    AERO_LBM_BENCHMARK_FLAG_DISABLE_SPONGE = 1u << 2,
    AERO_LBM_BENCHMARK_FLAG_DISABLE_CONVECTIVE_OUTFLOW = 1u << 3,
    AERO_LBM_BENCHMARK_FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND = 1u << 4,
    AERO_LBM_BENCHMARK_FLAG_DISABLE_SGS = 1u << 5,
    AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE = 1u << 6,
    AERO_LBM_BENCHMARK_FLAG_DISABLE_BUOYANCY = 1u << 7
} AeroLbmBenchmarkFlags;

/* ref: native/include/aero_lbm_capi.h:76-83 */
typedef struct AeroLbmBoundaryFaceConfig {
    int hydrodynamic_kind;
    int thermal_kind;
    float velocity[3];
    float pressure;
    float temperature;
    float heat_flux;
} AeroLbmBoundaryFaceConfig;

/* ref: native/include/aero_lbm_capi.h:85-111 */
typedef struct AeroLbmBenchmarkConfig {
    uint32_t abi_version;
    uint32_t struct_size;
    uint32_t flags;
    int enabled;
    int preset;
    int reserved0;
    float reynolds_number;
    float rayleigh_number;
    float prandtl_number;
    float mach_number;
    float reference_density;
    float reference_temperature;
    float reference_length;
    float body_force[3];
    float gravity[3];
    float initial_velocity[3];
    float initial_pressure;
    float initial_temperature;
    AeroLbmBoundaryFaceConfig x_min;
    AeroLbmBoundaryFaceConfig x_max;
    AeroLbmBoundaryFaceConfig y_min;
    AeroLbmBoundaryFaceConfig y_max;
    AeroLbmBoundaryFaceConfig z_min;
    AeroLbmBoundaryFaceConfig z_max;
    uint32_t reserved[8];
} AeroLbmBenchmarkConfig;

/* ref: native/include/aero_lbm_capi.h:113-123 */
typedef struct AeroLbmTimingSnapshot {
    uint64_t ticks;
    double last_payload_copy_ms;
    double last_solver_ms;
    double last_readback_ms;
    double last_total_ms;
    double avg_payload_copy_ms;
    double avg_solver_ms;
    double avg_readback_ms;
    double avg_total_ms;
} AeroLbmTimingSnapshot;

/* ref: native/include/aero_lbm_capi.h:125-141 */
typedef struct AeroLbmMesoscaleConfig {
    uint32_t abi_version;
    uint32_t struct_size;
    int nx;
    int ny;
    int nz;
    float dx_m;
    float dt_s;
    float molecular_nu_m2_s;
    float prandtl_air;
    float turbulent_prandtl;
    float reference_density_kg_m3;
    float ambient_air_temperature_k;
    float deep_ground_temperature_k;
    float background_wind_m_s[3];
    uint32_t reserved[8];
} AeroLbmMesoscaleConfig;

/* ref: native/include/aero_lbm_capi.h:143-149 */
typedef struct AeroLbmMesoscaleTransport {
    float velocity_scale_m_s_per_lattice;
    float nu_molecular_lattice;
    float alpha_molecular_lattice;
    float tau_shear_molecular;
    float tau_thermal_molecular;
} AeroLbmMesoscaleTransport;

/* ref: native/include/aero_lbm_capi.h:151-164 */
typedef struct AeroLbmWorldDelta {
    int type;
    int x;
    int y;
    int z;
    int data0;
    int data1;
    int data2;
    int data3;
    float value0;
    float value1;
    float value2;
    float value3;
} AeroLbmWorldDelta;

/* ------------------------------------------------------------------------
 * Extension: physics and kernel overrides.
 *
 * The reference hard-codes the Smagorinsky-type constant (C = 0.025), the
 * Boussinesq coefficient (beta = 0.001635, acting along +y) and the benchmark
 * thermal diffusivity (0.035) (ref: native/src/aero_lbm_jni.cpp:131-167,
 * 2288-2295, 2346-2366). This struct lets a caller re-parametrise them. A NaN
 * float or a negative int keeps the reference value. Buoyancy always acts
 * against the gravity axis given here (default: axis 1 = y, sign -1, so the
 * force points to +y exactly as in the reference).
 * ---------------------------------------------------------------------- */
enum {
    AERO_LBM_KERNEL_FAST = 0,   /* chimera moment transforms, FMA contraction */
    AERO_LBM_KERNEL_EXACT = 1   /* operation order of the reference CPU path */
};

typedef struct AeroLbmSolverOverrides {
    uint32_t struct_size;
    float sgs_constant;          /* C in nu_t = C^2 |Pi_neq| / (2 rho nu0) */
    float sgs_nut_cap_ratio;     /* nu_t <= ratio * nu0 */
    float boussinesq_beta;
    float thermal_diffusivity;   /* lattice units, benchmark mode */
    int gravity_axis;            /* 0 = x, 1 = y, 2 = z */
    int gravity_sign;            /* -1: gravity points to -axis (buoyancy to +axis) */
    int kernel_variant;          /* AERO_LBM_KERNEL_FAST / AERO_LBM_KERNEL_EXACT */
    int reserved[8];
} AeroLbmSolverOverrides;

/* ---- core solver context API (ref: native/include/aero_lbm_capi.h:166-285) ---- */
AERO_LBM_CAPI_EXPORT int aero_lbm_init(int grid_size, int input_channels, int output_channels);
AERO_LBM_CAPI_EXPORT int aero_lbm_step(const float* packet, int grid_size, long long context_key, float* output_flow);
AERO_LBM_CAPI_EXPORT int aero_lbm_init_rect(int nx, int ny, int nz, int input_channels, int output_channels);
AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect(const float* packet, int nx, int ny, int nz, long long context_key, float* output_flow);
AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_scaled(const float* packet, int nx, int ny, int nz, long long context_key, float output_velocity_scale, float* output_flow);
AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_cached(int nx, int ny, int nz, long long context_key, float* output_flow);
AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_cached_scaled(int nx, int ny, int nz, long long context_key, float output_velocity_scale, float* output_flow);
/* ref: capi.h:173-182. Static-cell patch of an in-place (A-A) context: solid flags of
 * the listed cells (last value of a repeated cell wins), fan directions 0-6 (no fan
 * force on the in-place path: refused while fan forcing is on), cells that turn
 * fluid restart from the rest equilibrium; the link masks follow the new flags. */
AERO_LBM_CAPI_EXPORT int aero_lbm_patch_d3q27_f16_static_cells_rect(
    int nx, int ny, int nz, long long context_key, int patch_count,
    const int* cell_indices, const uint8_t* solid_values, const uint8_t* fan_dir_values);
/* ref: capi.h:195-205. One packet step whose payload takes, for each listed cell, 6
 * overlay values: vx, vy, vz, pressure, temperature, source. */
AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_with_sparse_overlays(
    const float* packet, int nx, int ny, int nz, long long context_key,
    const int* overlay_cells, const float* overlay_values, int overlay_count, float* output_flow);
/* ref: capi.h:206. Moves every field of a context by -d (new cell x takes old x + d);
 * cells entering from outside take the rest equilibrium. Exact FP32 contexts. */
AERO_LBM_CAPI_EXPORT int aero_lbm_shift_context(int grid_size, long long context_key, int dx, int dy, int dz);
AERO_LBM_CAPI_EXPORT int aero_lbm_has_context(long long context_key);
/* ref: capi.h:208. The reference's compact experimental solver is not part of this
 * library: always 0. */
AERO_LBM_CAPI_EXPORT int aero_lbm_context_compact_initialized(long long context_key);
AERO_LBM_CAPI_EXPORT int aero_lbm_context_realtime_cached_initialized(long long context_key);
/* ref: capi.h:210-228. Halo exchange between two cubic contexts whose second lies at
 * offset (-1..1 per axis) from the first: populations (and the temperature when it
 * was updated last step) of the boundary boxes copied both ways; halo = grid / 4 for
 * aero_lbm_exchange_halo. Exact FP32 contexts; in-place contexts are left alone, as
 * the reference's exchange leaves its in-place buffers. */
AERO_LBM_CAPI_EXPORT int aero_lbm_exchange_halo(
    int grid_size, long long first_context_key, long long second_context_key,
    int offset_x, int offset_y, int offset_z);
AERO_LBM_CAPI_EXPORT int aero_lbm_exchange_halo_layers_rect(
    int nx, int ny, int nz, int halo_layers, long long first_context_key, long long second_context_key,
    int offset_x, int offset_y, int offset_z);
AERO_LBM_CAPI_EXPORT int aero_lbm_get_temperature_state_rect(int nx, int ny, int nz, long long context_key, float* out_temperature);
AERO_LBM_CAPI_EXPORT int aero_lbm_sample_temperature_point_rect(
    int nx, int ny, int nz, long long context_key,
    int sample_x, int sample_y, int sample_z, float* out_temperature);
AERO_LBM_CAPI_EXPORT int aero_lbm_sample_flow_point_rect(
    int nx, int ny, int nz, long long context_key,
    int sample_x, int sample_y, int sample_z, float* out_flow);
AERO_LBM_CAPI_EXPORT int aero_lbm_extract_flow_atlas_rect(
    int nx, int ny, int nz, long long context_key,
    int stride, float* out_flow_atlas, int value_count);
AERO_LBM_CAPI_EXPORT int aero_lbm_copy_flow_temperature_subrect(
    int nx, int ny, int nz, long long context_key,
    int offset_x, int offset_y, int offset_z,
    int copy_nx, int copy_ny, int copy_nz,
    float* out_flow, float* out_temperature);
AERO_LBM_CAPI_EXPORT int aero_lbm_get_flow_state_rect(int nx, int ny, int nz, long long context_key, float* out_flow);
AERO_LBM_CAPI_EXPORT int aero_lbm_set_temperature_state_rect(int nx, int ny, int nz, long long context_key, const float* temperature);
AERO_LBM_CAPI_EXPORT int aero_lbm_get_last_force(long long context_key, float* out_fx, float* out_fy, float* out_fz);
AERO_LBM_CAPI_EXPORT void aero_lbm_release_context(long long context_key);
AERO_LBM_CAPI_EXPORT void aero_lbm_shutdown(void);
AERO_LBM_CAPI_EXPORT const char* aero_lbm_runtime_info(void);
AERO_LBM_CAPI_EXPORT const char* aero_lbm_last_error(void);
AERO_LBM_CAPI_EXPORT const char* aero_lbm_timing_info(void);
AERO_LBM_CAPI_EXPORT const char* aero_lbm_memory_info(void);
AERO_LBM_CAPI_EXPORT void aero_lbm_reset_timing(void);
AERO_LBM_CAPI_EXPORT void aero_lbm_set_realtime_solver_mode(int solver_mode);
AERO_LBM_CAPI_EXPORT int aero_lbm_finish(void);
AERO_LBM_CAPI_EXPORT int aero_lbm_get_timing_snapshot(AeroLbmTimingSnapshot* out_snapshot);

/* ---- extension: conservation diagnostics (no counterpart in the reference).
 * Total mass and momentum {m, jx, jy, jz} of the fluid cells' populations as
 * stored (lattice units), summed in FP64 on the device in a fixed order. ---- */
AERO_LBM_CAPI_EXPORT int aero_lbm_get_conservation_rect(int nx, int ny, int nz, long long context_key, double* out4);
/* No counterpart in the reference. The same sums without any device allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the partial sums live in a fixed 64 KiB device
 * array (2048 x 4 partials + 4 totals) that each GPU holds until the process ends.
 * Any context size fits (the partial kernel loops over cells). Calls that overlap on
 * one GPU share that array and race; every partial is rewritten by each call. */
AERO_LBM_CAPI_EXPORT int aero_lbm_get_conservation_rect_noalloc(int nx, int ny, int nz, long long context_key,
                                                               double* out4);

/* ---- boundary-reference shell of the in-place paths
 * (ref: native/include/aero_lbm_capi.h:183-194, native/src/aero_lbm_jni.cpp:9962-10030).
 * flow_state: cells x flow_channels floats (u_x, u_y, u_z, p first); face_mask bits
 * x-, x+, y-, y+, z-, z+; fluid cells within shell_layers of a flagged face relax
 * their density and velocity toward the (FP16-rounded) reference each step. */
AERO_LBM_CAPI_EXPORT int aero_lbm_set_d3q27_f16_boundary_shell_rect(
    int nx, int ny, int nz, long long context_key, const float* flow_state, int flow_channels,
    int face_mask, int shell_layers, float relaxation);
AERO_LBM_CAPI_EXPORT int aero_lbm_clear_d3q27_f16_boundary_shell_rect(long long context_key);

/* ---- native brick runtime (ref: native/include/aero_lbm_capi.h:303-696,
 * native/src/aero_lbm_simulation_bridge.cpp:4562-5535). One service (key 1), worlds
 * of cubic bricks keyed by the caller, bricks keyed by integer brick coordinates.
 * Flow state: 4 floats per cell (lattice u_x, u_y, u_z, p); packed atlas: 4 int16
 * per sampled cell (velocity in m/s over +-5.6, pressure over +-0.03). Counts and
 * copies return -1 on error; the rest return 1/0 with the reason in
 * aero_lbm_simulation_last_error(). World delta type 10 (client source patches) is
 * accepted and ignored. */
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_create_service(long long* out_service_key);
AERO_LBM_CAPI_EXPORT void aero_lbm_simulation_release_service(long long service_key);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_focus(long long service_key, int block_x, int block_y, int block_z,
                                                       int radius_blocks);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_submit_world_deltas(long long service_key,
                                                                 const AeroLbmWorldDelta* deltas, int delta_count);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_ensure_brick_world_runtime(long long service_key, long long world_key,
                                                                        int brick_size, float dx_meters,
                                                                        float dt_seconds);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_solver_mode(long long service_key, long long world_key,
                                                                         int solver_mode);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_active_hints(long long service_key, long long world_key,
                                                                          int brick_size, const int* brick_coords,
                                                                          int brick_count);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_exact_active_hints(long long service_key,
                                                                                long long world_key, int brick_size,
                                                                                const int* brick_coords,
                                                                                int brick_count);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_get_brick_world_runtime_status(long long service_key, long long world_key,
                                                                            int* out_status, int status_count);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_step_brick_world_runtime(long long service_key, long long world_key,
                                                                      int step_count);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_get_brick_world_resident_brick_count(long long service_key,
                                                                                  long long world_key);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_resident_bricks(long long service_key,
                                                                              long long world_key, int* out_coords,
                                                                              int brick_capacity);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_get_brick_world_active_brick_count(long long service_key,
                                                                                long long world_key);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_active_bricks(long long service_key, long long world_key,
                                                                            int* out_coords, int brick_capacity);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_static_brick(
    long long service_key, long long world_key, int brick_size, int brick_x, int brick_y, int brick_z,
    const uint8_t* obstacle, const uint8_t* surface_kind, const uint16_t* open_face_mask,
    const float* emitter_power_watts, const uint8_t* face_sky_exposure, const uint8_t* face_direct_exposure);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_static_brick_with_sources(
    long long service_key, long long world_key, int brick_size, int brick_x, int brick_y, int brick_z,
    const uint8_t* obstacle, const uint8_t* surface_kind, const uint16_t* open_face_mask,
    const float* emitter_power_watts, const uint8_t* source_fan_dir, const float* source_emitter_power_watts,
    const uint8_t* face_sky_exposure, const uint8_t* face_direct_exposure);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_queue_brick_world_static_brick_upload(
    long long service_key, long long world_key, int brick_size, int brick_x, int brick_y, int brick_z,
    const uint8_t* obstacle, const uint8_t* surface_kind, const uint16_t* open_face_mask,
    const float* emitter_power_watts, const uint8_t* face_sky_exposure, const uint8_t* face_direct_exposure);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_dynamic_brick(
    long long service_key, long long world_key, int brick_size, int brick_x, int brick_y, int brick_z,
    const float* flow_state, const float* air_temperature, const float* surface_temperature);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_boundary_reference_brick(
    long long service_key, long long world_key, int brick_size, int brick_x, int brick_y, int brick_z,
    const float* flow_state, const float* air_temperature, const float* surface_temperature);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_dynamic_brick(
    long long service_key, long long world_key, int brick_size, int brick_x, int brick_y, int brick_z,
    float* out_flow_state, float* out_air_temperature, float* out_surface_temperature);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_packed_flow_atlas(
    long long service_key, long long world_key, int brick_size, int brick_x, int brick_y, int brick_z,
    int sample_stride, int16_t* out_packed_flow, int packed_value_count);
/* ref: capi.h:671-682. A caller's packed flow atlas kept under a key of the service
 * and read back whole (poll: value_count must equal the stored count). */
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_packed_flow_atlas(
    long long service_key, long long atlas_key, const int16_t* atlas_values, int value_count);
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_poll_packed_flow_atlas(
    long long service_key, long long atlas_key, int16_t* out_atlas_values, int value_count);
/* ref: capi.h:695. Counts of the service's worlds, bricks and contexts as text. */
AERO_LBM_CAPI_EXPORT const char* aero_lbm_simulation_runtime_info(void);
AERO_LBM_CAPI_EXPORT const char* aero_lbm_simulation_last_error(void);

/* ---- batched stepping (extension: no counterpart in the reference, whose brick
 * runtime steps and publishes each brick with its own launches and readback,
 * native/src/aero_lbm_simulation_bridge.cpp:1004-1171, 5301-5410). The contexts
 * (in-place storages, same storage and step parity, each stepped at least once)
 * advance together: per step one rare-cell gather, one rare-cell update and one bulk
 * launch for all bricks, each brick keeping the parameters of its last individual
 * step; every steps_per_publish (even) steps the int16 atlas of every brick
 * (publish_stride, velocity x publish_scales[i] over +-5.6, pressure over +-0.03)
 * is written and copied to host memory. The cycle is captured in a HIP graph.
 * autotune_launch: time brick orders and store cache policies, keep the fastest.
 * Stepping a member outside the batch invalidates the batch. */
/* Brick worlds use the same launches: each world step collects the steps of the
 * cached in-place bricks (the reference's step_brick_compact_cached per brick,
 * native/src/aero_lbm_simulation_bridge.cpp:1678-1704) and runs them as one launch
 * set per storage and parity; their packed atlases are quantised on the device.
 * enabled = 0 steps each brick through its own cached-step call instead. */
AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_batching(long long service_key, long long world_key,
                                                                      int enabled);
AERO_LBM_CAPI_EXPORT int aero_lbm_batch_create(const long long* context_keys, int count, int steps_per_publish,
                                               int publish_stride, const float* publish_scales, int autotune_launch,
                                               long long* out_batch);
AERO_LBM_CAPI_EXPORT int aero_lbm_batch_run(long long batch, int cycles, int use_graph, float* out_ms);
AERO_LBM_CAPI_EXPORT int aero_lbm_batch_run_per_brick(long long batch, int cycles, float* out_ms);
AERO_LBM_CAPI_EXPORT int aero_lbm_batch_copy_atlas(long long batch, int index, int16_t* out, int value_count);
AERO_LBM_CAPI_EXPORT const char* aero_lbm_batch_info(long long batch);
AERO_LBM_CAPI_EXPORT void aero_lbm_batch_release(long long batch);

/* ---- adjoint of the brick step (extension: no counterpart in the reference).
 * Model: one n^3 D3Q27 cumulant brick, FP32, Boussinesq buoyancy (gravity -z),
 * per-(y, z) velocity inlet on x- (the inflow profile, 3 floats per (y, z),
 * index (y * n + z) * 3), convective outflow on x+, periodic y and z, halfway
 * bounce-back at solid cells, benchmark-mode temperature update with a heat
 * source trilinearly upsampled from a heat_grid^3 field (index (i * m + j) * m + k),
 * no SGS term. Loss: sum over probe cells of |u - target|^2, u the collision
 * velocity after `steps` steps. Forward stores checkpoints every
 * checkpoint_every steps; backward replays each segment and returns dL/dinflow
 * and dL/dheat. forward_fp64 runs the same model in FP64 (finite-difference
 * checks). */
typedef struct AeroLbmAdjointConfig {
    int n;                 /* brick edge */
    int heat_grid;         /* edge m of the heat-source parameter field */
    int steps;
    int checkpoint_every;
    float nu;              /* lattice viscosity (tau = 0.5 + 3 nu) */
    float beta;            /* Boussinesq coefficient */
    float alpha;           /* thermal diffusivity */
    float u0[3];           /* initial velocity */
    int probe_count;
} AeroLbmAdjointConfig;
AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_create(const AeroLbmAdjointConfig* config, const uint8_t* solid,
                                                 const int* probe_cells, const float* target_uvw,
                                                 long long* out_handle);
AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_forward(long long handle, const float* inflow, const float* heat,
                                                  double* out_loss, float* out_probe_u);
AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_backward(long long handle, float* grad_inflow, float* grad_heat);
AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_forward_fp64(long long handle, const double* inflow, const double* heat,
                                                       double* out_loss);
AERO_LBM_CAPI_EXPORT void aero_lbm_adjoint_release(long long handle);

/* ---- mesoscale D3Q19 step (ref: native/include/aero_lbm_capi.h:286-301) ----
 * forcing: cells x 25 floats, state: cells x 6 floats (ambient, deep ground,
 * surface, wind x, y, z), cell = (x*ny + y)*nz + z, y vertical. */
AERO_LBM_CAPI_EXPORT void aero_lbm_mesoscale_default_config(AeroLbmMesoscaleConfig* out_config);
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_derive_transport(
    const AeroLbmMesoscaleConfig* config,
    AeroLbmMesoscaleTransport* out_transport);
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_create_context(
    const AeroLbmMesoscaleConfig* config,
    long long* out_context_key);
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_step_context(
    long long context_key,
    const AeroLbmMesoscaleConfig* config,
    const float* forcing,
    float* out_state);
AERO_LBM_CAPI_EXPORT void aero_lbm_mesoscale_release_context(long long context_key);

/* ---- extension: mesoscale (no counterpart in the reference) ----
 * set_tau_floor: lower clamp of both relaxation times (the reference fixes 0.58;
 *   NaN restores it), for low-viscosity runs such as tau = 0.5015;
 * run_steps_timed: `steps` steps with the same forcing, no per-step readback,
 *   device time in *out_ms;
 * get_state: the 6-channel state after the last step. */
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_set_tau_floor(float tau_floor);
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_run_steps_timed(
    long long context_key, const AeroLbmMesoscaleConfig* config, const float* forcing, int steps, float* out_ms);
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_get_state(long long context_key, float* out_state);

/* ---- analysis codec (ref: native/src/aero_lbm_analysis_codec.hpp:11-29, the
 * functions behind nativeCompressFloatGrid3d / nativeDecompressFloatGrid3d,
 * native/src/aero_lbm_simulation_bridge.cpp:8880-8976). Streams are zfp 1.0.1's
 * (fixed-accuracy mode with the tolerance given, full header), byte for byte the
 * reference's; a grid is values[x + nx * (y + ny * z)]. Errors in
 * aero_lbm_last_error(): the reference's strings invalid_input, invalid_tolerance,
 * zfp_write_header_failed (an extent above 65536), zfp_read_header_failed,
 * zfp_header_mismatch, zfp_decompress_failed (a stream shorter than its blocks).
 *
 * max_compressed_size_3d: bytes a stream of that shape can need (zfp's
 *   zfp_stream_maximum_size); 0 for an invalid shape.
 * compress_float_grid_3d: host values in, host bytes out (capacity bytes at out);
 *   returns the stream's size in bytes, 0 on error. The encoder runs on the GPU.
 * decompress_float_grid_3d: host stream in, host grid out; 1 on success. Runs on
 *   the host (a zfp stream has no block index; blocks decode in sequence). */
AERO_LBM_CAPI_EXPORT long long aero_lbm_analysis_max_compressed_size_3d(int nx, int ny, int nz);
AERO_LBM_CAPI_EXPORT long long aero_lbm_analysis_compress_float_grid_3d(
    const float* values, int nx, int ny, int nz, double tolerance, uint8_t* out, long long capacity);
AERO_LBM_CAPI_EXPORT int aero_lbm_analysis_decompress_float_grid_3d(
    const uint8_t* compressed, long long compressed_size, int nx, int ny, int nz, float* out_values);
/* extension (no counterpart in the reference): `count` grids (at most 64) already in
 * device memory, coded in one set of launches. values[i]: device pointer of grid i;
 * dims[3i..3i+2]: its nx, ny, nz; out: device buffer (8-byte aligned) that receives
 * stream i at byte out_offsets[i] (a multiple of 8) with out_capacities[i] bytes
 * room; out_sizes[i]: the stream's size. Host arrays except values[i] and out. */
AERO_LBM_CAPI_EXPORT int aero_lbm_analysis_compress_float_grids_3d_device(
    int count, const float* const* values, const int* dims, const double* tolerances, uint8_t* out,
    const long long* out_offsets, const long long* out_capacities, long long* out_sizes);
/* extension: the same launches `repeats` times back to back on the library's stream;
 * *out_ms = device time per repeat (HIP events), for measurement */
AERO_LBM_CAPI_EXPORT int aero_lbm_analysis_compress_grids_timed(
    int count, const float* const* values, const int* dims, const double* tolerances, uint8_t* out,
    const long long* out_offsets, const long long* out_capacities, int repeats, float* out_ms);

/* ---- benchmark configuration (ref: native/include/aero_lbm_capi.h:698-703) ---- */
AERO_LBM_CAPI_EXPORT void aero_lbm_benchmark_default_config(AeroLbmBenchmarkConfig* out_config);
AERO_LBM_CAPI_EXPORT int aero_lbm_benchmark_default_preset_config(int preset, AeroLbmBenchmarkConfig* out_config);
AERO_LBM_CAPI_EXPORT int aero_lbm_benchmark_set_config(const AeroLbmBenchmarkConfig* config);
AERO_LBM_CAPI_EXPORT int aero_lbm_benchmark_get_config(AeroLbmBenchmarkConfig* out_config);
AERO_LBM_CAPI_EXPORT void aero_lbm_benchmark_reset_config(void);
AERO_LBM_CAPI_EXPORT const char* aero_lbm_benchmark_info(void);

/* ---- extension: overrides (no counterpart in the reference; for callers that
 * re-parametrise the SGS constant, beta or the gravity axis, or pick the exact
 * or the fast kernels) ---- */
AERO_LBM_CAPI_EXPORT void aero_lbm_default_solver_overrides(AeroLbmSolverOverrides* out_overrides);
AERO_LBM_CAPI_EXPORT int aero_lbm_set_solver_overrides(const AeroLbmSolverOverrides* overrides);
AERO_LBM_CAPI_EXPORT int aero_lbm_get_solver_overrides(AeroLbmSolverOverrides* out_overrides);

/* ---- extension: device-resident timing (no counterpart in the reference).
 * Runs `steps` cached steps of a context back to back on the device stream with
 * no readback and returns the elapsed device time (HIP events) in *out_ms. ---- */
AERO_LBM_CAPI_EXPORT int aero_lbm_run_cached_steps_timed(
    int nx, int ny, int nz, long long context_key, int steps, float* out_ms);

#ifdef __cplusplus
}
#endif

#endif /* AERO_LBM_CAPI_H */
