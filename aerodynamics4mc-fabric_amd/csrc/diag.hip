// Conservation diagnostics: total mass and momentum of the fluid cells'
// populations as they are stored, summed in FP64 on the device. Each thread
// accumulates cells in FP64, a wavefront reduces with lane shuffles, a
// workgroup through LDS, and one final workgroup adds the per-workgroup
// partials in a fixed order, so the result does not depend on scheduling.
#include <algorithm>
#include <mutex>
#include <string>

#include "fast_kernels.h"
#include "runtime_internal.h"

namespace aero {
namespace diag {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;

__device__ __forceinline__ float wgt(int q) {
    const int n = (lat_cx(q) != 0) + (lat_cy(q) != 0) + (lat_cz(q) != 0);
    return n == 0 ? 8.0f / 27.0f : (n == 1 ? 2.0f / 27.0f : (n == 2 ? 1.0f / 54.0f : 1.0f / 216.0f));
}

// storage: 0 FP32, 1 FP16 shifted, 2 BF16 shifted, 3 16-bit fixed point shifted
// (step s16_class_range(|c|^2) * w / 32767); flip: slot s holds direction opp(s)
__device__ __forceinline__ float slot_value(const void* f, int storage, size_t off, int q) {
    if (storage == 0) return ((const float*)f)[off];
    const uint16_t b = ((const uint16_t*)f)[off];
    if (storage == 1) return (float)__builtin_bit_cast(_Float16, b);
    if (storage == 3) return (float)(int16_t)b * (s16_class_range(lat_c2(q)) * wgt(q) / 32767.0f);
    return __uint_as_float((uint32_t)b << 16);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void __launch_bounds__(kBlock) partial_kernel(const void* f, int storage, int flip, size_t plane, int cells,
                                                         const uint8_t* __restrict__ flags, double* __restrict__ partial) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int cell = blockIdx.x * kBlock + threadIdx.x; cell < cells; cell += gridDim.x * kBlock) {
        if (flags[cell] & kCellSolid) continue;
#pragma unroll
        for (int s = 0; s < 27; ++s) {
            const int q = flip ? lat_opp(s) : s;  // direction stored in slot s
            float v = slot_value(f, storage, (size_t)s * plane + cell, q);
            if (storage != 0) v += wgt(q);
            const double d = (double)v;
            acc[0] += d;
            acc[1] += (double)lat_cx(q) * d;
            acc[2] += (double)lat_cy(q) * d;
            acc[3] += (double)lat_cz(q) * d;
        }
    }
    __shared__ double lds[kBlock / 64][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double w = wave_sum(acc[k]);
        if (lane == 0) lds[wave][k] = w;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w][threadIdx.x];
        partial[blockIdx.x * 4 + threadIdx.x] = t;
    }
}

__global__ void __launch_bounds__(64) final_kernel(const double* __restrict__ partial, int blocks,
                                                   double* __restrict__ out) {
    // one wavefront; lane l adds partials l, l+64, ... in order, then a fixed shuffle tree
    for (int k = 0; k < 4; ++k) {
        double t = 0.0;
        for (int b = threadIdx.x; b < blocks; b += 64) t += partial[b * 4 + k];
        t = wave_sum(t);
        if (threadIdx.x == 0) out[k] = t;
    }
}

}  // namespace diag

// Fixed-size device scratch of the allocation-free entry point: the per-workgroup
// partials (at most kMaxBlocks x 4) followed by the four totals. One copy per
// device, held until the process ends. In host code `scratch` names a host-side
// stand-in; its device address comes from hipGetSymbolAddress (scratch_ptr).
constexpr int kScratchDoubles = diag::kMaxBlocks * 4 + 4;
static __device__ __align__(256) double scratch[kScratchDoubles];

// device address of `scratch` on the current device, looked up once per device
static double* scratch_ptr() {
    constexpr int kMaxDevices = 64;
    static double* ptrs[kMaxDevices] = {};
    static std::mutex mu;
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("get_conservation_noalloc: device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return nullptr;
        ptrs[dev] = (double*)p;
    }
    return ptrs[dev];
}

struct SlotView {
    const void* f = nullptr;
    int storage = 0, flip = 0;
    size_t plane = 0;
};

static SlotView slot_view(const Context& c) {
    SlotView v;
    v.plane = c.cells;
    switch (c.storage) {
        case Storage::F32Exact:
            v.f = c.f32[c.f_cur];
            break;
        case Storage::F32Fast:
            v.f = c.f32[c.f_cur];
            v.plane = (size_t)fast_plane_stride((int)c.cells);
            break;
        case Storage::F16:
        case Storage::BF16:
        case Storage::S16:
        case Storage::F32AA:
            v.f = c.storage == Storage::F32AA ? (const void*)c.f32[0] : (const void*)c.h16;
            v.storage = c.storage == Storage::F16 ? 1 : (c.storage == Storage::BF16 ? 2 : (c.storage == Storage::S16 ? 3 : 0));
            v.plane = (size_t)fast_plane_stride((int)c.cells);
            v.flip = c.aa_layout == 1 ? 1 : 0;  // after an even step, slot opp(q) holds direction q
            break;
    }
    return v;
}

static int conservation_blocks(const Context& c) {
    return (int)std::min<size_t>(diag::kMaxBlocks, (c.cells + diag::kBlock - 1) / diag::kBlock);
}

// partial sums into `partial` (blocks x 4), totals into `out_dev` (4), copy to out4
static bool conservation_into(Context& c, double* partial, double* out_dev, double* out4) {
    hipStream_t st = rt().stream;
    const SlotView v = slot_view(c);
    const int blocks = conservation_blocks(c);
    hipLaunchKernelGGL(diag::partial_kernel, dim3(blocks), dim3(diag::kBlock), 0, st, v.f, v.storage, v.flip, v.plane,
                       (int)c.cells, c.flags, partial);
    hipLaunchKernelGGL(diag::final_kernel, dim3(1), dim3(64), 0, st, partial, blocks, out_dev);
    return hip_ok(hipGetLastError(), "conservation") &&
           hip_ok(hipMemcpyAsync(out4, out_dev, 4 * sizeof(double), hipMemcpyDeviceToHost, st), "readback");
}

// mass and momentum of a context's populations; out4 receives {mass, jx, jy, jz}
bool storage_conservation(Context& c, double* out4) {
    hipStream_t st = rt().stream;
    const int blocks = conservation_blocks(c);
    double* partial = nullptr;
    double* out_dev = nullptr;
    if (!hip_ok(hipMallocAsync((void**)&partial, (size_t)blocks * 4 * sizeof(double), st), "hipMallocAsync") ||
        !hip_ok(hipMallocAsync((void**)&out_dev, 4 * sizeof(double), st), "hipMallocAsync"))
        return false;
    const bool ok = conservation_into(c, partial, out_dev, out4);
    (void)hipFreeAsync(partial, st);
    (void)hipFreeAsync(out_dev, st);
    return ok && hip_ok(hipStreamSynchronize(st), "sync");
}

// the same sums through the fixed device scratch: no allocation
static bool storage_conservation_noalloc(Context& c, double* out4) {
    const int blocks = conservation_blocks(c);
    if (blocks * 4 + 4 > kScratchDoubles) {
        set_error("get_conservation_noalloc: " + std::to_string(blocks) + " partial sums exceed the scratch of " +
                  std::to_string(diag::kMaxBlocks));
        return false;
    }
    double* base = scratch_ptr();
    if (!base) return false;
    return conservation_into(c, base, base + diag::kMaxBlocks * 4, out4) &&
           hip_ok(hipStreamSynchronize(rt().stream), "sync");
}

}  // namespace aero

using namespace aero;

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_get_conservation_rect(int nx, int ny, int nz, long long key,
                                                                   double* out4) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!out4) {
        set_error("get_conservation: null output");
        return 0;
    }
    Context* c = get_context(key, nx, ny, nz, false);
    if (!c || !c->initialized) {
        set_error("get_conservation: no initialised context for this key");
        return 0;
    }
    return storage_conservation(*c, out4) ? 1 : 0;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_get_conservation_rect_noalloc(int nx, int ny, int nz, long long key,
                                                                           double* out4) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!out4) {
        set_error("get_conservation_noalloc: null output");
        return 0;
    }
    Context* c = find_context(key);
    if (!c || !c->initialized || c->nx != nx || c->ny != ny || c->nz != nz) {
        set_error("get_conservation_noalloc: no initialised context of these dimensions for this key");
        return 0;
    }
    return storage_conservation_noalloc(*c, out4) ? 1 : 0;
}
