// Internal runtime state shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aero_lbm_capi.h"
#include "kernels.h"

namespace aero {

enum class Storage { F32Exact, F32Fast, F16, BF16, S16, F32AA };

// 16-bit shifted populations (FP16, BF16, fixed point)
inline bool storage_16bit(Storage s) { return s == Storage::F16 || s == Storage::BF16 || s == Storage::S16; }

constexpr int kMaskSigLen = 22;

struct Context {
    int nx = 0, ny = 0, nz = 0;
    size_t cells = 0;
    Storage storage = Storage::F32Fast;
    bool initialized = false;      // populations seeded from the first payload
    bool post_collision = false;   // buffer holds post-collision populations
    bool has_payload = false;
    bool thermal_on = false;       // sticky: temperature can be non-zero
    bool temperature_nonzero = false;
    bool payload_src_nonzero = false;
    bool payload_t_nonzero = false;
    uint64_t steps = 0;
    // Bumped by every change that can outdate a batch's copy of this context's step
    // parameters or buffer pointers: an individual step, a parameter upload, a link
    // mask / rare-list rebuild, a temperature upload, a shell change. Batches record it
    // and refuse to launch when it moved.
    uint64_t generation = 0;
    int f_cur = 0;                 // A-B: index of the buffer holding the state
    int t_cur = 0;                 // temperature ping-pong index
    int aa_layout = 3;             // A-A: 1 after an even step, 2 after an odd step, 3 initial state
    int store_cur = 0;             // A-A: face store written by the last step
    uint32_t* mask = nullptr;      // fast paths: per-cell link mask
    // fast paths: cells the rare-cell kernel updates
    int* rare_list = nullptr;
    int* rare_counter = nullptr;
    int* rare_scratch = nullptr;
    int rare_count = 0;
    int rare_cap = 0;              // capacity of rare_vals per value
    float* rare_vals = nullptr;    // 28 x rare_cap values of the listed cells
    int mask_sig[kMaskSigLen] = {};  // parameters the link mask was built with
    bool mask_valid = false;
    // device buffers
    uint8_t* flags = nullptr;
    float* ref_p = nullptr;
    float* ref_u = nullptr;
    float* ref_t = nullptr;
    float* src = nullptr;
    float* fan = nullptr;          // payload fan mask and velocity (four planes), allocated when fans appear
    bool fan_active = false;       // the resident payload has fan cells and fan forcing is on
    float* T[2] = {nullptr, nullptr};
    float* u_state = nullptr;
    float* t_bfecc = nullptr;      // corrected field of the BFECC passes (allocated on first use)
    double* force_part = nullptr;  // per-block drag partials, then the float force (cylinder preset)
    bool force_valid = false;      // force_part ends with this context's last drag integral
    float* f32[2] = {nullptr, nullptr};  // FP32 A-B populations (27 planes)
    uint16_t* h16 = nullptr;             // 16-bit A-A populations (27 planes, shifted by w_q)
    float* face_store[2] = {nullptr, nullptr};  // A-A: near-face post-collision values per parity
    size_t face_store_floats = 0;
    float* payload_dev = nullptr;
    // boundary-reference shell of the fast paths
    uint8_t* shell = nullptr;
    uint16_t* shell_state = nullptr;
    bool shell_on = false;
    int shell_layers = 0;
    float shell_relax = 0.0f;
    float* out_dev = nullptr;
    float* sample_dev = nullptr;
    FaceParams* faces_dev = nullptr;   // device copy of the face table
    FaceParams faces_host[6] = {};     // last uploaded face table (source of the async copy)
    bool faces_valid = false;
    StepParams* params_dev = nullptr;  // device copy of the step parameters
    StepParams params_host = {};       // last uploaded parameters (source of the async copy)
    bool params_valid = false;
    int class_wrap[6] = {-1, -1, -1, -1, -1, -1};  // periodic flags the cell classes were built with
    // static-cell patches of A-A contexts: device copies of the last patch (cells, solid
    // values, flip kinds) and a per-cell marker of the flipped cells
    int* patch_cells = nullptr;
    uint8_t* patch_solid = nullptr;
    uint8_t* patch_flip = nullptr;
    int patch_cap = 0;
    uint8_t* patch_marks = nullptr;
    std::vector<float> payload_host;
    std::vector<void*> allocations;
    uint64_t timing_upload_bytes = 0;

    bool alloc(void** p, size_t bytes);
    ~Context();
};

// parameters the fast kernels' link mask depends on (besides the cell flags)
void mask_signature(const StepParams& P, int* out);

struct RuntimeConfig {
    bool initialized = false;
    int grid_size = 0, nx = 0, ny = 0, nz = 0, in_ch = 0, out_ch = 0;
};

struct Timing {
    uint64_t ticks = 0;
    double last_total = 0, last_solver = 0, last_readback = 0, last_copy = 0;
    double sum_total = 0, sum_solver = 0, sum_readback = 0, sum_copy = 0;
};

AeroLbmBenchmarkConfig default_benchmark_config();
AeroLbmSolverOverrides default_overrides();

struct RuntimeState {
    std::mutex mu;
    RuntimeConfig cfg;
    AeroLbmBenchmarkConfig bench = default_benchmark_config();
    AeroLbmSolverOverrides overrides = default_overrides();
    int solver_mode = AERO_LBM_REALTIME_SOLVER_AUTO;
    std::string last_error;
    std::string runtime_info;
    std::string device_name, device_arch;
    int cu_count = 0;
    bool device_ready = false;
    hipStream_t stream = nullptr;
    std::unordered_map<long long, std::unique_ptr<Context>> contexts;
    Timing timing;
};

RuntimeState& rt();
void set_error(const std::string& e);
void clear_error();
bool hip_ok(hipError_t e, const char* what);
bool ensure_device();
bool resolve_params(int nx, int ny, int nz, int in_ch, StepParams& P);
// resolve_params + upload of the face table into the context's device copy
bool prepare_params(Context& c, StepParams& P);
bool sync_params(Context& c, StepParams& P);
void sanitize_benchmark_config(AeroLbmBenchmarkConfig& c);
Context* find_context(long long key);
Context* get_context(long long key, int nx, int ny, int nz, bool create);
CellFields cell_fields(const Context& c);
bool step_context(Context& c, const float* packet, float* out, int out_ch, float scale, bool packet_on_device = false);
// aero_lbm_step_rect_scaled for a packet in device memory (fan, source and
// temperature channels zero); packet_dev == nullptr steps the resident payload
int step_rect_device_packet(const float* packet_dev, int nx, int ny, int nz, long long key, float scale, float* out);
// context `key` exists with these dimensions and holds an ingested payload
bool context_resident(long long key, int nx, int ny, int nz);
void step_switches(const Context& c, StepParams& P);
// the sticky thermal switch of a context under parameters P
void update_thermal_activity(Context& c, const StepParams& P);

// storage-specific operations (storage.hip)
bool storage_alloc(Context& c);
bool storage_step(Context& c, const StepParams& P);
bool storage_readout(Context& c, const StepParams& P, int stride, int out_ch, float scale, float* out_dev);
bool storage_sample(Context& c, const StepParams& P, int x, int y, int z, float* out_dev4);
size_t context_device_bytes(const Context& c);
bool storage_rebuild_masks(Context& c, const StepParams& P);
// buffers a step with parameters P needs besides the populations (A-A face stores)
bool storage_prepare_step(Context& c, const StepParams& P);

// ---- batched stepping of resident in-place contexts (batch.hip) ----
// The context can join a batched step: FP16 / BF16 / S16 / FP32 A-A storage, stepped
// at least once (A-A layout 1 or 2), resident payload, benchmark-mode parameters.
bool context_batchable(long long key);
// The host half of a cached step of `key` under the current configuration: the
// parameters are resolved and uploaded as aero_lbm_step_rect_cached would, without
// the launch. Returns the context, or nullptr with the error set.
Context* prepare_cached_step(long long key, int nx, int ny, int nz);
// One step of every listed context (prepared by prepare_cached_step), in shared
// launches: contexts at the same storage and parity share one rare-gather, one
// rare-update and one bulk launch.
bool batch_step_contexts(const std::vector<Context*>& ctxs);
// int16 packed flow atlas of one in-place context, quantised on the device (velocity
// x vscale over +-5.6 m/s, pressure over +-0.03), copied to out (4 values per sample)
bool publish_context_atlas(long long key, int stride, float vscale, int16_t* out, int value_count);
// the same for many contexts in one launch and one readback: out holds each
// context's atlas at offsets[i] (4 values per sample); all share one storage
bool publish_contexts_atlas(const std::vector<long long>& keys, int stride, float vscale, std::vector<int16_t>& out,
                            std::vector<size_t>& offsets);

}  // namespace aero
