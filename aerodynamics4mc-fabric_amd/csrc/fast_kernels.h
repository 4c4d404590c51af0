// Launchers of the fast D3Q27 kernels (d3q27_fast.hip).
#pragma once

#include "kernels.h"

namespace aero {

// population storage of the fast kernels: FP32, and 16-bit shifted (f - w_q) as
// IEEE half, bfloat16, or signed fixed point with a range per direction class
enum class FastStorage { F32, F16, BF16, S16 };

// per-cell link mask: bit q (0..26) = the pull source x - c_q is solid or out of the
// domain after periodic wrapping; plus the two flags below
constexpr uint32_t kMaskLinks = (1u << 27) - 1u;
constexpr uint32_t kMaskSolid = 1u << 27;  // the cell itself is solid
constexpr uint32_t kMaskFace = 1u << 28;   // a link leaves through a face whose rule is not plain bounce-back
constexpr uint32_t kMaskTherm = 1u << 29;  // a stencil neighbour is outside the domain at a Dirichlet/Neumann thermal face
// the cell is updated by the rare-cell kernel: a face rule other than plain
// bounce-back or a Dirichlet / Neumann thermal face applies (except through the
// x faces flagged StepParams::xface_bulk and the symmetry y faces flagged
// StepParams::yface_bulk, whose planes and rows the bulk kernel updates),
// the sponge touches it, or the thermal periodicity differs from the hydrodynamic one
constexpr uint32_t kMaskRare = 1u << 30;

// largest population array the fast step kernel addresses (one buffer resource,
// per-lane byte offsets below 2^31)
constexpr long long kFastMaxBufferBytes = (1ll << 31) - 1;

// plane stride (cells) of the fast kernels' population arrays for a brick of
// `cells` cells: padded so that the 27 planes do not start at addresses that are
// equal modulo a large power of two
int fast_plane_stride(int cells);

void launch_fast_mask(const StepParams& P, const uint8_t* flags, uint32_t* mask, hipStream_t s);
// ascending list of the fluid cells marked kMaskRare; *counter receives their
// number; scratch holds (cells + 255) / 256 ints
void launch_fast_rare_list(const StepParams& P, const uint32_t* mask, int* list, int* scratch, int* counter,
                           hipStream_t s);

// per-context buffers of the rare-cell pass
struct FastRare {
    const int* list;   // fluid cells marked kMaskRare
    int count;
    int cap;           // capacity of vals per value (>= count)
    float* vals;       // 28 x cap: incoming populations and new temperature of the listed cells
};
void launch_fast_init(const StepParams& P, const CellFields& C, FastStorage st, void* f, float* T, float* T_other, void* u_state,
                      float t_init, hipStream_t s);
// mode: 0 = A-B pull (local_first: collide the initial state in place),
//       1 = A-A even (local), 2 = A-A odd (pull + push)
void launch_fast_step(const StepParams& P, const CellFields& C, FastStorage st, const uint32_t* mask, const void* fin,
                      void* fout, const float* t_in, float* t_out, void* u_state, const float* store_in,
                      float* store_out, int mode, int local_first, const FastRare& rare, hipStream_t s);
// layout: 0 = A-B post-collision, 1 = A-A after an even step, 2 = A-A after an odd
// step, 3 = initial state (incoming values in own slots, no face rules)
void launch_fast_readout(const StepParams& P, const CellFields& C, FastStorage st, const uint32_t* mask, const void* f,
                         const float* store, int layout, int stride, int out_ch, float scale, float* out,
                         hipStream_t s);

// Static-cell patch of an A-A context after its flags and link mask were rebuilt.
// flip[i] bits: the populations of cells[i] restart (kPatchReset), its state differs
// from before the patch so its neighbours' links change (kPatchLinks), it ends solid
// (kPatchSolid); marks = 1 at every restarted cell. layout as in launch_fast_readout
// (1, 2 or 3).
constexpr uint8_t kPatchReset = 1, kPatchLinks = 2, kPatchSolid = 4;
void launch_fast_patch_reset(const StepParams& P, FastStorage st, const uint32_t* mask, const uint8_t* flags,
                             const uint8_t* marks, void* f, int layout, const int* cells, const uint8_t* flip,
                             int count, hipStream_t s);

// ---- many bricks per launch ----
// One brick of a batched step: its step parameters and buffers for that step.
// Blocks of a batched grid find their brick through a per-block owner table; a
// brick's blocks start at a multiple of 8 so its XCD-aware tile order is kept.
struct BatchBrick {
    StepParams P;
    CellFields C;
    const uint32_t* mask;
    void* f;                   // A-A population buffer
    const float* t_in;
    float* t_out;
    void* u_state;
    const float* store_in;
    float* store_out;
    const int* rare_list;
    int rare_count, rare_cap;
    float* rare_vals;
    int bulk_begin, gather_begin, update_begin, publish_begin;  // first block of each grid
    int publish_layout;        // A-A layout the publish reads (after this step)
    int publish_stride;
    float publish_scale;       // lattice -> m/s factor of the published velocity
    int16_t* publish_out;      // 4 int16 per sampled cell
};

// one batched step: every brick of `table` takes one A-A step of parity `mode`
// (1 even, 2 odd); the grids are sized by the owner tables
struct BatchGrids {
    const uint16_t* bulk_owner;
    const uint16_t* gather_owner;
    const uint16_t* update_owner;
    const uint16_t* publish_owner;
    int bulk_blocks, gather_blocks, update_blocks, publish_blocks;
    int yface = 1;  // some brick has a y face flagged yface_bulk: launch the kernels with the y rules
};
void launch_batch_step(FastStorage st, int mode, int store_policy, bool shell, const BatchBrick* table,
                       const BatchGrids& g, hipStream_t s);
// int16 flow atlas of every brick (velocity over +-5.6 m/s, pressure over +-0.03)
void launch_batch_publish(FastStorage st, const BatchBrick* table, const BatchGrids& g, hipStream_t s);
int fast_tile_blocks(const StepParams& P);
int fast_store_policy(size_t working_set_bytes);

}  // namespace aero
