// Host-side launch wrappers for the HIP kernels (one translation unit per kernel
// family so each can carry its own floating-point flags).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

namespace aero {

// Per-context device state shared by all D3Q27 kernels.
struct CellFields {
    const uint8_t* flags;  // kCellSolid | kCellBoundary
    const float* ref_p;    // payload channel 8 (clamped)
    const float* ref_u;    // payload channels 5-7, three planes
    const float* ref_t;    // payload channel 10 (clamped)
    const float* src;      // payload channel 9 (clamped)
    const uint8_t* shell;         // boundary-reference shell layer per cell (0: outside)
    const uint16_t* shell_state;  // FP16 (u_x, u_y, u_z, p) targets of the shell, 4 per cell
    const float* fan;             // payload channels 1-4 (mask, velocity), four planes; null without fans
};

// ---- payload ingest and cell classification (d3q27_common.hip) ----
void launch_ingest(const StepParams& P, const float* payload, uint8_t* flags, float* ref_p, float* ref_u,
                   float* ref_t, float* src, float* fan, hipStream_t s);
void launch_classify(const StepParams& P, uint8_t* flags, hipStream_t s);

// ---- exact FP32 A-B path (d3q27_exact.hip): operation order of the reference CPU step ----
void launch_exact_init(const StepParams& P, const CellFields& C, float* f, float* t, float* u_state,
                       float initial_temperature, hipStream_t s);
void launch_exact_step(const StepParams& P, const CellFields& C, const float* fin, float* fout, const float* t_in,
                       float* t_out, float* u_state, int local_mode, hipStream_t s);
// BFECC temperature of one step (heated-cavity presets) into t_out; t_mid is a
// cell-sized scratch field
void launch_exact_bfecc(const StepParams& P, const CellFields& C, const float* t_in, float* t_out, float* t_mid,
                        const float* u_state, hipStream_t s);
// runtime mode, every step after the first: solid cells' temperature to 0 and the
// temperature within 8 cells of a face nudged toward the payload's (in place)
void launch_exact_t_reference(const StepParams& P, const CellFields& C, float* t, hipStream_t s);
// drag integral of the cylinder preset over post-collision populations: part holds
// 3 doubles per 256-cell block, then the float force (fx, fy, fz)
void launch_exact_force(const StepParams& P, const CellFields& C, const float* fpost, double* part, hipStream_t s);
void launch_exact_stream(const StepParams& P, const CellFields& C, const float* fin, float* fout, hipStream_t s);
void launch_exact_readout(const StepParams& P, const CellFields& C, const float* f, int post_collision, int stride,
                          int out_ch, float scale, float* out, hipStream_t s);
void launch_exact_sample(const StepParams& P, const CellFields& C, const float* f, int post_collision, int x, int y,
                         int z, float* out4, hipStream_t s);

}  // namespace aero
