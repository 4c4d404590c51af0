// D3Q27 lattice and solver parameter block shared by host and device code.
//
// Velocity set (ref: native/include/aero_lbm_hydro_core.h:122-205):
// q = (ix*3 + iy)*3 + iz with c = i - 1 per axis, rest q = 13, opp(q) = 26 - q,
// weights are products of the 1-D weights {1/6, 2/3, 1/6}.
// Memory layout (ref: native/src/aero_lbm_jni.cpp:1180-1187): structure of
// arrays, q-major planes f[q * cells + cell], cell = (x * ny + y) * nz + z, so
// one wavefront of consecutive z cells reads one contiguous run per plane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aero {

constexpr int kQ = 27;

__host__ __device__ constexpr int lat_cx(int q) { return q / 9 - 1; }
__host__ __device__ constexpr int lat_cy(int q) { return (q / 3) % 3 - 1; }
__host__ __device__ constexpr int lat_cz(int q) { return q % 3 - 1; }
__host__ __device__ constexpr int lat_opp(int q) { return 26 - q; }
__host__ __device__ constexpr int lat_c2(int q) {
    return lat_cx(q) * lat_cx(q) + lat_cy(q) * lat_cy(q) + lat_cz(q) * lat_cz(q);
}

// Range of the 16-bit fixed-point population storage per direction class |c|^2
// (rest, axis, face diagonal, corner), in units of the class weight: a stored
// shifted value f - w_q covers [-R w_q, R w_q]. An equilibrium at density 1 stays
// inside it up to a speed of about 0.19 along a corner diagonal (Mach 0.33) and
// 0.26 along an axis.
__host__ __device__ constexpr float s16_class_range(int c2) {
    return c2 == 0 ? 0.25f : (c2 == 1 ? 1.0f : (c2 == 2 ? 1.25f : 1.5f));
}

// cell classification bits (one byte per cell)
enum : uint8_t {
    kCellSolid = 1u,     // obstacle
    kCellBoundary = 2u,  // fluid cell with a solid or out-of-domain link (per-link rules apply)
};

// Face order matches AeroLbmBenchmarkConfig: x_min, x_max, y_min, y_max, z_min, z_max.
struct FaceParams {
    int hydro_kind;
    int thermal_kind;
    float vel[3];
    float pressure;
    float temperature;
    float heat_flux;
    long long store_off;  // offset (floats) of this face's block in the A-A face store
};

// Everything a step kernel needs, resolved on the host from the benchmark
// configuration (ref: native/src/aero_lbm_jni.cpp:787-993) plus overrides.
struct StepParams {
    int nx, ny, nz;
    int cells;
    int plane;  // population plane stride of the fast kernels, in cells (cells + padding)
    FaceParams face[6];
    int hydro_periodic[3];    // axis wraps (both faces periodic on that side handled per side below)
    int hydro_wrap_lo[3];     // face kind of the min side is PERIODIC
    int hydro_wrap_hi[3];     // face kind of the max side is PERIODIC
    int thermal_wrap_lo[3];
    int thermal_wrap_hi[3];
    int sgs_on;
    int buoyancy_on;
    int sponge_on;
    int source_on;            // internal thermal source channel active
    int convective_disabled;  // DISABLE_CONVECTIVE_OUTFLOW
    int thermal_on;           // temperature field can be non-zero
    int shell_on;             // boundary-reference shell present (fast paths)
    int shell_layers;         // shell depth in cells
    float shell_relax;        // shell relaxation at the outermost layer
    int in_ch;
    float tau_shear_base, tau_normal_base, nu_base;
    float alpha, beta, sgs_c2, nut_cap_ratio;
    int g_axis, g_sign;
    float kmax_speed;
    // device copy of face[] for run-time indexed reads in kernels (a by-value
    // kernel argument indexed at run time would be copied to scratch memory)
    const FaceParams* faces_dev;
    // A-A face store: bit f set when face f needs previous-step values (symmetry or
    // convective rule); offsets (floats) of each face's [3 layers][9 q][A][B] block
    int face_store_faces;
    long long face_store_off[6];
    long long face_store_floats;
    // thermal periodic flags equal the hydrodynamic ones, so a cell whose 26
    // neighbours are fluid and in the domain can read its temperature stencil directly
    int thermal_wrap_matches;
    // x_min / x_max plane: its fluid cells take the face's rules in the bulk step
    // kernel (the face is not periodic and its hydrodynamic rule is not plain
    // bounce-back, or its thermal rule is Dirichlet / Neumann). A plane of x is a
    // set of whole wavefront rows, so the rules run as a wave-uniform branch.
    int xface_bulk[2];
    // y_min / y_max: a symmetry face (not periodic, no Dirichlet / Neumann thermal
    // rule) whose rows run the symmetry rule in the bulk step kernel; a row of
    // constant y is a whole wavefront, as for the x planes
    int yface_bulk[2];
    // temperature by BFECC advection (heated-cavity presets): three passes before
    // the step leave the new temperature in t_out, which the exact step kernel reads
    int thermal_bfecc;
    // cylinder crossflow preset: parabolic inlet on the x- velocity face and the
    // drag integral after each step (exact kernels)
    int cylinder;
    // runtime (non-benchmark) mode of the reference (jni.cpp:675-681): velocity and
    // temperature nudged toward the payload reference within 8 cells of every face,
    // thermal diffusivity alpha0 + nu_t / Pr_t, temperature every second step (dt 2)
    int runtime_mode;
    int thermal_update;   // this step updates the temperature (step counter % stride == 0)
    float thermal_dt;     // temperature time step: the update stride
    float nu_rt0;         // runtime molecular viscosity (caps the thermal turbulent viscosity)
    // fan forcing (jni.cpp:2258-2286): enabled and the payload has fan cells
    int fan_on;
    float fan_noise_amp;
    unsigned long long tick;  // the context's step counter while fan forcing runs (noise seed)
    // device copy of this whole struct: rarely taken kernel paths read their
    // parameters from it, so the bulk path keeps no scalar registers for them
    const StepParams* self_dev;
};

}  // namespace aero
