// Boundary-reference shell of the in-place paths: the per-cell layer index and the
// FP16 reference state that the step kernels blend toward (the blend itself is
// shell_blend in fast_core.h).
#include <cmath>

#include "runtime_internal.h"

namespace aero {
namespace shell {

// float -> IEEE binary16 bits, rounding to nearest with ties away from zero and
// flushing below the subnormal range to a signed zero: the rounding the reference
// applies to its shell targets (native/src/aero_lbm_jni.cpp:1205-1235). Inputs are
// finite (the build kernel replaces non-finite values first).
__device__ __forceinline__ uint16_t half_bits(float v) {
    const uint32_t b = __float_as_uint(v);
    const uint32_t sign = (b >> 16) & 0x8000u;
    uint32_t man = b & 0x007FFFFFu;
    const int e = (int)((b >> 23) & 0xFFu) - 112;  // rebias 127 -> 15
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x00800000u;
        const int shift = 14 - e;
        uint32_t h = man >> shift;
        h += (man >> (shift - 1)) & 1u;
        return (uint16_t)(sign | h);
    }
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
    h += (man >> 12) & 1u;
    return (uint16_t)h;
}

__device__ __forceinline__ float finite_or0(float v) { return isfinite(v) ? v : 0.0f; }

// One thread per cell (ref: native/src/aero_lbm_jni.cpp:6471-6522): the layer is
// the distance + 1 to the nearest exposed face within `layers`; targets are the
// reference velocity capped at the lattice speed limit and the pressure clamped.
__global__ void __launch_bounds__(256) build_kernel(int nx, int ny, int nz, const float* __restrict__ flow, int ch,
                                                    int face_mask, int layers, float kmax,
                                                    const uint8_t* __restrict__ flags, uint8_t* __restrict__ layer_out,
                                                    uint16_t* __restrict__ state) {
    const int cells = nx * ny * nz;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    const int yz = ny * nz;
    const int x = cell / yz, rem = cell - x * yz;
    const int y = rem / nz, z = rem - y * nz;
    int best = layers + 1;
    if ((face_mask & 1) && x < layers) best = min(best, x + 1);
    if ((face_mask & 2) && x >= nx - layers) best = min(best, nx - x);
    if ((face_mask & 4) && y < layers) best = min(best, y + 1);
    if ((face_mask & 8) && y >= ny - layers) best = min(best, ny - y);
    if ((face_mask & 16) && z < layers) best = min(best, z + 1);
    if ((face_mask & 32) && z >= nz - layers) best = min(best, nz - z);
    uint2 packed = make_uint2(0u, 0u);
    if ((flags[cell] & kCellSolid) || best > layers) {
        layer_out[cell] = 0;
    } else {
        const float* f = flow + (size_t)cell * ch;
        float vx = finite_or0(f[0]), vy = finite_or0(f[1]), vz = finite_or0(f[2]);
        const float p = fminf(fmaxf(finite_or0(f[3]), -0.03f), 0.03f);
        const float s2 = vx * vx + vy * vy + vz * vz;
        if (!isfinite(s2) || s2 > kmax * kmax) {
            if (!isfinite(s2) || s2 <= 0.0f) {
                vx = vy = vz = 0.0f;
            } else {
                const float sc = kmax / sqrtf(s2);
                vx *= sc;
                vy *= sc;
                vz *= sc;
            }
        }
        layer_out[cell] = (uint8_t)best;
        packed.x = (uint32_t)half_bits(vx) | ((uint32_t)half_bits(vy) << 16);
        packed.y = (uint32_t)half_bits(vz) | ((uint32_t)half_bits(p) << 16);
    }
    reinterpret_cast<uint2*>(state)[cell] = packed;
}

}  // namespace shell

// the speed limit of the in-place paths: sound speed x 0.6 (jni.cpp:85, 150)
constexpr float kShellMaxSpeed = 0.57735026919f * 0.60f;

static bool shell_supported(const Context& c) { return c.storage != Storage::F32Exact; }

bool context_set_shell(Context& c, const float* flow_state, int flow_channels, int face_mask, int shell_layers,
                       float relaxation) {
    const float relax = std::isfinite(relaxation) ? std::fmin(std::fmax(relaxation, 0.0f), 0.5f) : 0.0f;
    c.generation++;
    if (face_mask == 0 || shell_layers <= 0 || relax <= 0.0f) {
        c.shell_on = false;
        return true;
    }
    if (!c.has_payload) {
        set_error("set_d3q27_f16_boundary_shell: the context has no payload (solid flags) yet");
        return false;
    }
    const int layers = std::max(1, std::min({shell_layers, c.nx, c.ny, c.nz, 255}));
    hipStream_t st = rt().stream;
    if (!c.shell) {
        if (!c.alloc((void**)&c.shell, c.cells) || !c.alloc((void**)&c.shell_state, c.cells * 8)) return false;
    }
    // the reference flow state travels through the payload staging buffer
    const size_t bytes = c.cells * (size_t)flow_channels * sizeof(float);
    float* staging = nullptr;
    if (!hip_ok(hipMallocAsync((void**)&staging, bytes, st), "hipMallocAsync") ||
        !hip_ok(hipMemcpyAsync(staging, flow_state, bytes, hipMemcpyHostToDevice, st), "upload shell reference"))
        return false;
    const int cells = (int)c.cells;
    hipLaunchKernelGGL(shell::build_kernel, dim3((cells + 255) / 256), dim3(256), 0, st, c.nx, c.ny, c.nz, staging,
                       flow_channels, face_mask, layers, kShellMaxSpeed, c.flags, c.shell, c.shell_state);
    (void)hipFreeAsync(staging, st);
    if (!hip_ok(hipGetLastError(), "shell build") || !hip_ok(hipStreamSynchronize(st), "sync")) return false;
    c.shell_on = true;
    c.shell_layers = layers;
    c.shell_relax = relax;
    return true;
}

}  // namespace aero

using namespace aero;

// ref: native/src/aero_lbm_jni.cpp:9962-10010 (declaration aero_lbm_capi.h:183-194)
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_set_d3q27_f16_boundary_shell_rect(int nx, int ny, int nz,
                                                                              long long context_key,
                                                                              const float* flow_state,
                                                                              int flow_channels, int face_mask,
                                                                              int shell_layers, float relaxation) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        set_error("set_d3q27_f16_boundary_shell: invalid dimensions");
        return 0;
    }
    if (!flow_state || flow_channels < 4) {
        set_error("set_d3q27_f16_boundary_shell: missing flow state");
        return 0;
    }
    Context* c = find_context(context_key);
    if (!c) {
        set_error("set_d3q27_f16_boundary_shell: context not found");
        return 0;
    }
    if (c->nx != nx || c->ny != ny || c->nz != nz) {
        set_error("set_d3q27_f16_boundary_shell: context shape mismatch");
        return 0;
    }
    if (!shell_supported(*c)) {
        set_error("set_d3q27_f16_boundary_shell: the context uses the exact FP32 path, which has no shell");
        return 0;
    }
    return context_set_shell(*c, flow_state, flow_channels, face_mask, shell_layers, relaxation) ? 1 : 0;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_clear_d3q27_f16_boundary_shell_rect(long long context_key) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    Context* c = find_context(context_key);
    if (!c) {
        set_error("clear_d3q27_f16_boundary_shell: context not found");
        return 0;
    }
    c->shell_on = false;
    c->generation++;
    return 1;
}
