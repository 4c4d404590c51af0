// Context operations of the core C ABI besides stepping and reading: the static-cell
// patch of in-place contexts, steps with sparse overlays of the payload, the context
// shift and the halo exchange between two contexts (ref: native/include/
// aero_lbm_capi.h:173-221; native/src/aero_lbm_jni.cpp:9905-10131).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "fast_kernels.h"
#include "runtime_internal.h"

namespace aero {
namespace {

constexpr int kMaxPlanes = 32;
struct Fringe {
    float v[kMaxPlanes];  // value of each plane where the shifted source lies outside
};

__device__ __forceinline__ void cell_xyz(int cell, int ny, int nz, int* x, int* y, int* z) {
    const int yz = ny * nz;
    *x = cell / yz;
    *y = (cell - *x * yz) / nz;
    *z = cell - *x * yz - *y * nz;
}

// out[p][x] = in[p][x + d] (float planes of stride `stride`), fringe value outside
__global__ void __launch_bounds__(256) shift_planes_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                           int nx, int ny, int nz, int planes, size_t stride, int dx,
                                                           int dy, int dz, Fringe fringe) {
    const size_t cells = (size_t)nx * ny * nz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * planes) return;
    const int p = (int)(i / cells);
    const int cell = (int)(i - (size_t)p * cells);
    int x, y, z;
    cell_xyz(cell, ny, nz, &x, &y, &z);
    const int ox = x + dx, oy = y + dy, oz = z + dz;
    const bool inside = ox >= 0 && oy >= 0 && oz >= 0 && ox < nx && oy < ny && oz < nz;
    out[p * stride + cell] = inside ? in[p * stride + ((size_t)ox * ny + oy) * nz + oz] : fringe.v[p];
}

__global__ void __launch_bounds__(256) shift_bytes_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                          int nx, int ny, int nz, int dx, int dy, int dz,
                                                          uint8_t fill) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nx * ny * nz) return;
    int x, y, z;
    cell_xyz(cell, ny, nz, &x, &y, &z);
    const int ox = x + dx, oy = y + dy, oz = z + dz;
    const bool inside = ox >= 0 && oy >= 0 && oz >= 0 && ox < nx && oy < ny && oz < nz;
    out[cell] = inside ? in[((size_t)ox * ny + oy) * nz + oz] : fill;
}

// box copy between two n^3 grids: planes of stride `stride`, source box at s0, target at t0
struct Box {
    int s0[3], t0[3], size[3];
};
__global__ void __launch_bounds__(256) box_copy_kernel(const float* __restrict__ src, float* __restrict__ dst, int n,
                                                       int planes, size_t stride, Box b) {
    const size_t per = (size_t)b.size[0] * b.size[1] * b.size[2];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per * planes) return;
    const int p = (int)(i / per);
    const int r = (int)(i - (size_t)p * per);
    const int k = r % b.size[2], j = (r / b.size[2]) % b.size[1], l = r / (b.size[2] * b.size[1]);
    const size_t from = ((size_t)(b.s0[0] + l) * n + (b.s0[1] + j)) * n + (b.s0[2] + k);
    const size_t to = ((size_t)(b.t0[0] + l) * n + (b.t0[1] + j)) * n + (b.t0[2] + k);
    dst[p * stride + to] = src[p * stride + from];
}

// patch of the solid flags. solid[i]: bit 0 the cell's last patched value, bit 1 some
// entry set it solid, bit 2 some entry set it fluid. flip[i] (kPatch* bits): the
// populations restart (an entry changed the state, as the reference resets a cell
// each time an entry changes it), the links of the neighbours change (the state
// before and after differ), the cell ends solid
__global__ void __launch_bounds__(256) patch_flags_kernel(uint8_t* __restrict__ flags, uint8_t* __restrict__ marks,
                                                          const int* __restrict__ cells,
                                                          const uint8_t* __restrict__ solid,
                                                          uint8_t* __restrict__ flip, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int cell = cells[i];
    const bool was = flags[cell] & kCellSolid, now = solid[i] & 1;
    const bool changed = was ? (solid[i] & 4) != 0 : (solid[i] & 2) != 0;
    flags[cell] = (uint8_t)((flags[cell] & ~kCellSolid) | (now ? kCellSolid : 0));
    flip[i] = (uint8_t)((changed ? kPatchReset : 0) | (was != now ? kPatchLinks : 0) | (now ? kPatchSolid : 0));
    marks[cell] = changed ? 1 : 0;
}

__global__ void __launch_bounds__(256) clear_marks_kernel(uint8_t* __restrict__ marks, const int* __restrict__ cells,
                                                          int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) marks[cells[i]] = 0;
}

dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

bool in_place(Storage s) { return storage_16bit(s) || s == Storage::F32AA; }

FastStorage fast_storage_of(Storage s) {
    switch (s) {
        case Storage::F16: return FastStorage::F16;
        case Storage::BF16: return FastStorage::BF16;
        case Storage::S16: return FastStorage::S16;
        default: return FastStorage::F32;
    }
}

void* in_place_buffer(Context& c) { return c.storage == Storage::F32AA ? (void*)c.f32[0] : (void*)c.h16; }

bool fail(const std::string& msg) {
    set_error(msg);
    return false;
}

// the exact path keeps a step's post-collision populations until the next step
// streams them; the operations below act on the streamed populations (the
// reference's pre-collision f), so a pending stream runs first
void settle_exact(Context& c, const StepParams& P) {
    if (c.storage != Storage::F32Exact || !c.initialized || !c.post_collision) return;
    launch_exact_stream(P, cell_fields(c), c.f32[c.f_cur], c.f32[1 - c.f_cur], rt().stream);
    c.f_cur ^= 1;
    c.post_collision = false;
}

bool grow_patch_buffers(Context& c, int count) {
    if (count <= c.patch_cap) return true;
    if (!hip_ok(hipStreamSynchronize(rt().stream), "sync")) return false;
    void* old[3] = {c.patch_cells, c.patch_solid, c.patch_flip};
    for (void* p : old) {
        if (!p) continue;
        (void)hipFree(p);
        for (auto& a : c.allocations)
            if (a == p) a = nullptr;
    }
    c.patch_cells = nullptr;
    c.patch_solid = c.patch_flip = nullptr;
    c.patch_cap = 0;
    const int cap = (count + 1023) & ~1023;
    if (!c.alloc((void**)&c.patch_cells, cap * sizeof(int)) || !c.alloc((void**)&c.patch_solid, cap) ||
        !c.alloc((void**)&c.patch_flip, cap))
        return false;
    c.patch_cap = cap;
    return true;
}

// ref: native/src/aero_lbm_jni.cpp:6156-6415 (patch_d3q27_f16_static_cells) and
// 4725-4778 (the patch and reset kernels)
bool patch_static_cells(Context& c, int count, const int* cells, const uint8_t* solid, const uint8_t* fan_dir) {
    // one entry per cell, in order of first appearance, with its last value and the
    // values it took on the way (ref: 6190-6216)
    std::vector<int> pc;
    std::vector<uint8_t> ps;
    std::unordered_map<int, size_t> at;
    pc.reserve(count);
    ps.reserve(count);
    bool fan = false;
    for (int i = 0; i < count; ++i) {
        const int cell = cells[i];
        if (cell < 0 || (size_t)cell >= c.cells) continue;
        const bool s = solid[i] != 0;
        auto ins = at.emplace(cell, pc.size());
        if (ins.second) {
            pc.push_back(cell);
            ps.push_back(0);
        }
        uint8_t& v = ps[ins.first->second];
        v = (uint8_t)((v & ~1) | (s ? 1 : 0) | (s ? 2 : 4));
        fan |= !s && fan_dir[i] != 0;
    }
    if (pc.empty()) return true;
    StepParams P;
    if (!prepare_params(c, P)) return false;
    // the fan direction codes drive a fan force the cumulant in-place step does not
    // have (fan forcing runs on the exact kernels); with fan forcing off they are moot
    if (fan && P.fan_on)
        return fail("patch_d3q27_f16_static_cells: fan cells need fan forcing, which runs on the exact FP32 kernels; "
                    "set AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING or step the context with a payload");
    const int n = (int)pc.size();
    hipStream_t st = rt().stream;
    if (!grow_patch_buffers(c, n)) return false;
    if (!c.patch_marks) {
        if (!c.alloc((void**)&c.patch_marks, c.cells) || !hip_ok(hipMemsetAsync(c.patch_marks, 0, c.cells, st), "memset"))
            return false;
    }
    if (!hip_ok(hipMemcpyAsync(c.patch_cells, pc.data(), n * sizeof(int), hipMemcpyHostToDevice, st), "patch upload") ||
        !hip_ok(hipMemcpyAsync(c.patch_solid, ps.data(), n, hipMemcpyHostToDevice, st), "patch upload"))
        return false;
    hipLaunchKernelGGL(patch_flags_kernel, grid_for(n), dim3(256), 0, st, c.flags, c.patch_marks, c.patch_cells,
                       c.patch_solid, c.patch_flip, n);
    // cell classes and link masks of the new flags, then the populations
    launch_classify(P, c.flags, st);
    if (!storage_rebuild_masks(c, P)) return false;
    launch_fast_patch_reset(P, fast_storage_of(c.storage), c.mask, c.flags, c.patch_marks, in_place_buffer(c),
                            c.aa_layout, c.patch_cells, c.patch_flip, n, st);
    hipLaunchKernelGGL(clear_marks_kernel, grid_for(n), dim3(256), 0, st, c.patch_marks, c.patch_cells, n);
    // wait before the host arrays go out of scope
    return hip_ok(hipGetLastError(), "patch") && hip_ok(hipStreamSynchronize(st), "patch");
}

// ref: native/src/aero_lbm_jni.cpp:8432-8476 (shift_context_cpu_state): every field
// moves by -d (new(x) = old(x + d)); populations entering from outside take the
// rest equilibrium, scalar fields zero (density one)
bool shift_context(Context& c, int dx, int dy, int dz) {
    StepParams P;
    if (!prepare_params(c, P)) return false;
    settle_exact(c, P);
    hipStream_t st = rt().stream;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const size_t N = c.cells;
    Fringe eq{}, zero{};
    // the rest equilibrium f_q = w_q, the product of the 1-D weights as the exact kernels form it
    const float w1[3] = {1.0f / 6.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    for (int q = 0; q < 27; ++q) eq.v[q] = w1[lat_cx(q) + 1] * w1[lat_cy(q) + 1] * w1[lat_cz(q) + 1];
    // populations: into the other A-B buffer
    hipLaunchKernelGGL(shift_planes_kernel, grid_for(N * 27), dim3(256), 0, st, c.f32[c.f_cur], c.f32[1 - c.f_cur],
                       nx, ny, nz, 27, N, dx, dy, dz, eq);
    c.f_cur ^= 1;
    // temperature: into the other buffer, which then takes the same values
    hipLaunchKernelGGL(shift_planes_kernel, grid_for(N), dim3(256), 0, st, c.T[c.t_cur], c.T[1 - c.t_cur], nx, ny, nz,
                       1, N, dx, dy, dz, zero);
    c.t_cur ^= 1;
    if (!hip_ok(hipMemcpyAsync(c.T[1 - c.t_cur], c.T[c.t_cur], N * 4, hipMemcpyDeviceToDevice, st), "shift")) return false;
    // the other per-cell fields through one scratch plane
    float* tmp = nullptr;
    if (!hip_ok(hipMalloc(&tmp, N * 4), "hipMalloc")) return false;
    struct Field {
        float* p;
        int planes;
    };
    const Field fields[] = {{c.u_state, 3}, {c.ref_p, 1}, {c.ref_u, 3}, {c.ref_t, 1}, {c.src, 1}, {c.fan, 4}};
    bool ok = true;
    for (const Field& f : fields) {
        for (int k = 0; ok && f.p && k < f.planes; ++k) {
            float* plane = f.p + (size_t)k * N;
            hipLaunchKernelGGL(shift_planes_kernel, grid_for(N), dim3(256), 0, st, plane, tmp, nx, ny, nz, 1, N, dx, dy,
                               dz, zero);
            ok = hip_ok(hipMemcpyAsync(plane, tmp, N * 4, hipMemcpyDeviceToDevice, st), "shift");
        }
    }
    uint8_t* btmp = reinterpret_cast<uint8_t*>(tmp);
    hipLaunchKernelGGL(shift_bytes_kernel, grid_for(N), dim3(256), 0, st, c.flags, btmp, nx, ny, nz, dx, dy, dz,
                       (uint8_t)0);
    ok = ok && hip_ok(hipMemcpyAsync(c.flags, btmp, N, hipMemcpyDeviceToDevice, st), "shift");
    launch_classify(P, c.flags, st);
    ok = ok && hip_ok(hipGetLastError(), "shift") && hip_ok(hipStreamSynchronize(st), "shift");
    (void)hipFree(tmp);
    // a payload handed over again must be ingested again (ref: the CPU path ingests
    // the payload on every packet step)
    c.payload_host.clear();
    c.force_valid = false;
    return ok;
}

// one axis of the halo boxes (ref: native/src/aero_lbm_jni.cpp:8201-8263): the box
// of `first` copied into `second` and the box of `second` copied into `first`
bool halo_axis(int offset, int n, int halo, int core, int* first_src, int* second_dst, int* second_src,
               int* first_dst, int* size) {
    switch (offset) {
        case -1: *first_src = halo, *second_dst = n - halo, *second_src = core, *first_dst = 0, *size = halo; return true;
        case 0: *first_src = halo, *second_dst = halo, *second_src = halo, *first_dst = halo, *size = core; return true;
        case 1: *first_src = core, *second_dst = 0, *second_src = halo, *first_dst = n - halo, *size = halo; return true;
        default: return false;
    }
}

// the temperature was updated by the context's last step (ref: jni.cpp:2198-2207)
bool temperature_updated_last(const Context& c, const StepParams& P) {
    const uint64_t stride = P.runtime_mode ? 2u : 1u;
    return stride <= 1 || c.steps == 0 || ((c.steps - 1) % stride) == 0;
}

// ref: native/src/aero_lbm_jni.cpp:7728-7900 (GPU) and 8266-8426 (CPU)
bool exchange_halo(Context& a, Context& b, int ox, int oy, int oz, int halo_layers, const char* who) {
    if (!a.initialized || !b.initialized) return true;  // nothing stepped yet
    // the reference's exchange acts on its classic populations and leaves the
    // in-place buffers alone
    if (in_place(a.storage) || in_place(b.storage)) return true;
    if (a.storage != Storage::F32Exact || b.storage != Storage::F32Exact)
        return fail(std::string(who) + ": the halo exchange runs on the exact FP32 kernels (AERO_LBM_KERNEL_EXACT)");
    const int n = a.nx;
    const int halo = std::max(1, std::min(halo_layers, n / 2 - 1));
    const int core = n - 2 * halo;
    if (core <= 0) return fail(std::string(who) + ": invalid halo/core layout");
    Box ab{}, ba{};
    const int off[3] = {ox, oy, oz};
    for (int k = 0; k < 3; ++k)
        if (!halo_axis(off[k], n, halo, core, &ab.s0[k], &ab.t0[k], &ba.s0[k], &ba.t0[k], &ab.size[k]))
            return fail(std::string(who) + ": halo slab bounds invalid");
    std::copy(ab.size, ab.size + 3, ba.size);
    StepParams Pa, Pb;
    if (!prepare_params(a, Pa) || !prepare_params(b, Pb)) return false;
    settle_exact(a, Pa);
    settle_exact(b, Pb);
    hipStream_t st = rt().stream;
    const size_t N = a.cells, per = (size_t)ab.size[0] * ab.size[1] * ab.size[2];
    hipLaunchKernelGGL(box_copy_kernel, grid_for(per * 27), dim3(256), 0, st, a.f32[a.f_cur], b.f32[b.f_cur], n, 27, N,
                       ab);
    hipLaunchKernelGGL(box_copy_kernel, grid_for(per * 27), dim3(256), 0, st, b.f32[b.f_cur], a.f32[a.f_cur], n, 27, N,
                       ba);
    if (temperature_updated_last(a, Pa) || temperature_updated_last(b, Pb)) {
        hipLaunchKernelGGL(box_copy_kernel, grid_for(per), dim3(256), 0, st, a.T[a.t_cur], b.T[b.t_cur], n, 1, N, ab);
        hipLaunchKernelGGL(box_copy_kernel, grid_for(per), dim3(256), 0, st, b.T[b.t_cur], a.T[a.t_cur], n, 1, N, ba);
        for (Context* c : {&a, &b})
            if (!hip_ok(hipMemcpyAsync(c->T[1 - c->t_cur], c->T[c->t_cur], N * 4, hipMemcpyDeviceToDevice, st), "halo"))
                return false;
    }
    return hip_ok(hipGetLastError(), "halo exchange");
}

std::string missing_pair(const char* who, long long k1, long long k2, const Context* a, const Context* b) {
    return std::string(who) + ": missing context(s) first=" + std::to_string(k1) + (a ? "" : " [missing]") +
           ", second=" + std::to_string(k2) + (b ? "" : " [missing]");
}

int halo_entry(int nx, int ny, int nz, int layers, long long k1, long long k2, int ox, int oy, int oz,
               const char* who) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!rt().cfg.initialized) return fail(std::string(who) + ": runtime not initialized");
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(std::string(who) + ": invalid dimensions");
    if (layers <= 0) return fail(std::string(who) + ": halo_layers must be positive");
    if (ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1)
        return fail(std::string(who) + ": invalid offset (" + std::to_string(ox) + "," + std::to_string(oy) + "," +
                    std::to_string(oz) + ")");
    if (ox == 0 && oy == 0 && oz == 0) return fail(std::string(who) + ": zero offset is invalid");
    if (nx != ny || ny != nz)
        return fail(std::string(who) + ": only cubic contexts are supported, got " + std::to_string(nx) + "x" +
                    std::to_string(ny) + "x" + std::to_string(nz));
    if (layers * 2 >= nx) return fail(std::string(who) + ": invalid halo layout");
    Context* a = find_context(k1);
    Context* b = find_context(k2);
    if (!a || !b) return fail(missing_pair(who, k1, k2, a, b));
    for (const Context* c : {a, b})
        if (c->nx != nx || c->ny != ny || c->nz != nz) return fail(std::string(who) + ": context shape mismatch");
    return exchange_halo(*a, *b, ox, oy, oz, layers, who) ? 1 : 0;
}

}  // namespace
}  // namespace aero

using namespace aero;

extern "C" {

// ref: native/include/aero_lbm_capi.h:173-182, native/src/aero_lbm_jni.cpp:9905-9960
AERO_LBM_CAPI_EXPORT int aero_lbm_patch_d3q27_f16_static_cells_rect(int nx, int ny, int nz, long long context_key,
                                                                    int patch_count, const int* cell_indices,
                                                                    const uint8_t* solid_values,
                                                                    const uint8_t* fan_dir_values) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    const RuntimeConfig& g = rt().cfg;
    if (!g.initialized) return fail("patch_d3q27_f16_static_cells: runtime is not initialized") ? 1 : 0;
    if (patch_count <= 0) return 1;
    if (!cell_indices || !solid_values || !fan_dir_values)
        return fail("patch_d3q27_f16_static_cells: missing patch arrays") ? 1 : 0;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx != g.nx || ny != g.ny || nz != g.nz)
        return fail("patch_d3q27_f16_static_cells: invalid dimensions") ? 1 : 0;
    Context* c = find_context(context_key);
    if (!c) return fail("patch_d3q27_f16_static_cells: context not found") ? 1 : 0;
    if (c->nx != nx || c->ny != ny || c->nz != nz)
        return fail("patch_d3q27_f16_static_cells: context shape mismatch") ? 1 : 0;
    if (!in_place(c->storage) || !c->initialized)
        return fail("patch_d3q27_f16_static_cells: the context holds no stepped in-place populations (step it in "
                    "an in-place solver mode first)")
                   ? 1
                   : 0;
    return patch_static_cells(*c, patch_count, cell_indices, solid_values, fan_dir_values) ? 1 : 0;
}

// ref: native/include/aero_lbm_capi.h:195-205, native/src/aero_lbm_jni.cpp:8686-8730 and
// 6995-7044: the overlay values replace the payload's velocity, pressure, source and
// temperature channels of the listed cells for this step
AERO_LBM_CAPI_EXPORT int aero_lbm_step_rect_with_sparse_overlays(const float* packet, int nx, int ny, int nz,
                                                                 long long context_key, const int* overlay_cells,
                                                                 const float* overlay_values, int overlay_count,
                                                                 float* output_flow) {
    int ch = 0;
    {
        std::lock_guard<std::mutex> lk(rt().mu);
        clear_error();
        if (!rt().cfg.initialized) return fail("native_step_raw_dims_sparse: runtime not initialized") ? 1 : 0;
        if (!packet) return fail("native_step_raw_dims_sparse: missing packet") ? 1 : 0;
        if (nx <= 0 || ny <= 0 || nz <= 0) return fail("native_step_raw_dims_sparse: invalid dimensions") ? 1 : 0;
        if (overlay_count < 0) return fail("native_step_raw_dims_sparse: invalid overlay count") ? 1 : 0;
        if (overlay_count > 0 && (!overlay_cells || !overlay_values))
            return fail("native_step_raw_dims_sparse: missing overlay data") ? 1 : 0;
        ch = rt().cfg.in_ch;
    }
    if (overlay_count == 0) return aero_lbm_step_rect(packet, nx, ny, nz, context_key, output_flow);
    const size_t cells = (size_t)nx * ny * nz;
    std::vector<float> merged(packet, packet + cells * ch);
    for (int i = 0; i < overlay_count; ++i) {
        const int cell = overlay_cells[i];
        if (cell < 0 || (size_t)cell >= cells) continue;
        const float* v = overlay_values + (size_t)i * 6;  // vx, vy, vz, pressure, temperature, source
        float* d = merged.data() + (size_t)cell * ch;
        std::copy(v, v + 4, d + 5);
        if (ch > 9) d[9] = v[5];
        if (ch > 10) d[10] = v[4];
    }
    return aero_lbm_step_rect(merged.data(), nx, ny, nz, context_key, output_flow);
}

// ref: native/include/aero_lbm_capi.h:206, native/src/aero_lbm_jni.cpp:9603-9616
AERO_LBM_CAPI_EXPORT int aero_lbm_shift_context(int grid_size, long long context_key, int dx, int dy, int dz) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    const RuntimeConfig& g = rt().cfg;
    if (!g.initialized || grid_size != g.grid_size || grid_size <= 0)
        return fail("shift_context: runtime not initialized for this cubic grid size") ? 1 : 0;
    Context* c = find_context(context_key);
    if (!c) return fail("shift_context: context not found") ? 1 : 0;
    if (c->nx != g.nx || c->ny != g.ny || c->nz != g.nz) return fail("shift_context: context shape mismatch") ? 1 : 0;
    if (dx == 0 && dy == 0 && dz == 0) return 1;
    if (std::abs(dx) >= c->nx || std::abs(dy) >= c->ny || std::abs(dz) >= c->nz)
        return fail("shift_context: shift exceeds the grid") ? 1 : 0;
    if (!c->has_payload) return 1;  // no state yet
    if (c->storage != Storage::F32Exact)
        return fail("shift_context: runs on the exact FP32 kernels (classic solver mode, AERO_LBM_KERNEL_EXACT)") ? 1
                                                                                                               : 0;
    return shift_context(*c, dx, dy, dz) ? 1 : 0;
}

// ref: native/include/aero_lbm_capi.h:208, native/src/aero_lbm_jni.cpp:10071-10079. The
// compact experimental solver (a reduced-lattice path of the reference) is not
// part of this library, so no context is ever compact-initialized.
AERO_LBM_CAPI_EXPORT int aero_lbm_context_compact_initialized(long long context_key) {
    (void)context_key;
    return 0;
}

// ref: native/include/aero_lbm_capi.h:210-217, native/src/aero_lbm_jni.cpp:9623-9691: the
// halo width is a quarter of the grid
AERO_LBM_CAPI_EXPORT int aero_lbm_exchange_halo(int grid_size, long long first_context_key,
                                                long long second_context_key, int offset_x, int offset_y,
                                                int offset_z) {
    {
        std::lock_guard<std::mutex> lk(rt().mu);
        if (!rt().cfg.initialized || grid_size != rt().cfg.grid_size || grid_size <= 0) {
            clear_error();
            return fail("native_exchange_halo: grid size mismatch (requested=" + std::to_string(grid_size) +
                        ", configured=" + std::to_string(rt().cfg.grid_size) + ")")
                       ? 1
                       : 0;
        }
    }
    return halo_entry(grid_size, grid_size, grid_size, std::max(1, grid_size / 4), first_context_key,
                      second_context_key, offset_x, offset_y, offset_z, "native_exchange_halo");
}

// ref: native/include/aero_lbm_capi.h:218-228, native/src/aero_lbm_jni.cpp:9693-9788
AERO_LBM_CAPI_EXPORT int aero_lbm_exchange_halo_layers_rect(int nx, int ny, int nz, int halo_layers,
                                                            long long first_context_key,
                                                            long long second_context_key, int offset_x, int offset_y,
                                                            int offset_z) {
    return halo_entry(nx, ny, nz, halo_layers, first_context_key, second_context_key, offset_x, offset_y, offset_z,
                      "native_exchange_halo_rect_layers");
}

}  // extern "C"
