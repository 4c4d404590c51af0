// Payload ingest and cell classification ("flag build").
//
// ingest: the reference re-reads its 11-channel AoS payload on every CPU step
// (ref: native/src/aero_lbm_jni.cpp:1671-1694); here the payload is uploaded
// only when it changes and unpacked once into structure-of-arrays fields.
// classify: a fluid cell gets kCellBoundary when any of its 27 links pulls from
// a solid cell or from outside the domain after periodic wrapping -- exactly the
// links for which the reference's stream_and_bounce takes a non-plain branch
// (ref: native/src/aero_lbm_jni.cpp:2415-2452). Bulk cells then skip all
// per-link checks.
#include "kernels.h"

namespace aero {

__device__ __forceinline__ float clampf_ref(float v, float lo, float hi) {
    const float m = (lo < v) ? v : lo;
    return (m < hi) ? m : hi;
}
__device__ __forceinline__ float finite_or0(float v) { return __builtin_isfinite(v) ? v : 0.0f; }

__global__ void __launch_bounds__(256) ingest_kernel(int cells, int ch, const float* __restrict__ payload,
                                                     uint8_t* __restrict__ flags, float* __restrict__ ref_p,
                                                     float* __restrict__ ref_u, float* __restrict__ ref_t,
                                                     float* __restrict__ src, float* __restrict__ fan) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const float* p = payload + (size_t)cell * ch;
    flags[cell] = p[0] > 0.5f ? kCellSolid : 0;
    ref_u[cell] = finite_or0(p[5]);
    ref_u[(size_t)cells + cell] = finite_or0(p[6]);
    ref_u[2 * (size_t)cells + cell] = finite_or0(p[7]);
    ref_p[cell] = clampf_ref(p[8], -0.03f, 0.03f);
    ref_t[cell] = clampf_ref(ch > 10 ? p[10] : 0.0f, -1.0f, 1.0f);
    src[cell] = clampf_ref(ch > 9 ? p[9] : 0.0f, -0.006f, 0.006f);
    if (fan) {
        fan[cell] = clampf_ref(p[1], 0.0f, 1.0f);
        fan[(size_t)cells + cell] = finite_or0(p[2]);
        fan[2 * (size_t)cells + cell] = finite_or0(p[3]);
        fan[3 * (size_t)cells + cell] = finite_or0(p[4]);
    }
}

__device__ __forceinline__ bool wrap_coord(int& c, int n, int wrap_lo, int wrap_hi) {
    if (c >= 0 && c < n) return true;
    if (c < 0 && wrap_lo) { c += n; return true; }
    if (c >= n && wrap_hi) { c -= n; return true; }
    return false;
}

__global__ void __launch_bounds__(256) classify_kernel(StepParams P, uint8_t* __restrict__ flags) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    const uint8_t f = flags[cell];
    if (f & kCellSolid) return;
    const int yz = P.ny * P.nz;
    const int x = cell / yz;
    const int rem = cell - x * yz;
    const int y = rem / P.nz;
    const int z = rem - y * P.nz;
    bool boundary = false;
    for (int q = 0; q < kQ && !boundary; ++q) {
        int sx = x - lat_cx(q), sy = y - lat_cy(q), sz = z - lat_cz(q);
        const bool in = wrap_coord(sx, P.nx, P.hydro_wrap_lo[0], P.hydro_wrap_hi[0]) &
                        wrap_coord(sy, P.ny, P.hydro_wrap_lo[1], P.hydro_wrap_hi[1]) &
                        wrap_coord(sz, P.nz, P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]);
        if (!in) boundary = true;
        else if (flags[(sx * P.ny + sy) * P.nz + sz] & kCellSolid) boundary = true;
    }
    // bit 1 is written by this kernel only; bit 0 is read by neighbours, so keep it
    flags[cell] = boundary ? (uint8_t)(f | kCellBoundary) : (uint8_t)(f & ~kCellBoundary);
}

static inline dim3 grid_for(int n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_ingest(const StepParams& P, const float* payload, uint8_t* flags, float* ref_p, float* ref_u,
                   float* ref_t, float* src, float* fan, hipStream_t s) {
    hipLaunchKernelGGL(ingest_kernel, grid_for(P.cells), dim3(256), 0, s, P.cells, P.in_ch, payload, flags, ref_p,
                       ref_u, ref_t, src, fan);
}

void launch_classify(const StepParams& P, uint8_t* flags, hipStream_t s) {
    hipLaunchKernelGGL(classify_kernel, grid_for(P.cells), dim3(256), 0, s, P, flags);
}

}  // namespace aero
