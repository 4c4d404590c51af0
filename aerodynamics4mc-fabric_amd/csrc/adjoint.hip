// Adjoint of the brick step (model in adjoint_core.h): forward run with
// checkpoints every k steps, and a hand-written reverse pass that replays each
// segment from its checkpoint and walks it backwards. Per step, three kernels:
//   collide adjoint  one thread per cell: the 27 incoming values and the new
//                    temperature are recomputed; the collision depends on them
//                    through 10 raw moments and T, so the hand-written reverse
//                    mode of the post-collision map, then the transposed moment
//                    map, carry the adjoint of the cell's outputs back to them;
//   stream adjoint   one thread per (cell, direction): gathers the adjoints of
//                    every read of the previous post-collision value (pull,
//                    bounce-back, convective outflow);
//   thermal adjoint  one thread per cell: scatters the temperature adjoint onto
//                    the stencil of the semi-Lagrangian update (atomics), and
//                    gives the velocity and heat-source adjoints.
// The loss is the squared error of the collision velocity at probe cells after
// the last step; the parameters are the x- inflow profile (3 floats per (y, z))
// and a coarse heat-source field trilinearly upsampled to the brick.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aero_lbm_capi.h"
#include "adjoint_core.h"
#include "runtime_internal.h"

namespace aero {
namespace adj {

constexpr int kState = 31;  // floats per cell: 27 populations, temperature, 3 velocity

template <class R>
__device__ __forceinline__ void cell_xyz(const Params& P, int c, int& x, int& y, int& z) {
    const int yz = P.ny * P.nz;
    x = c / yz;
    const int r = c - x * yz;
    y = r / P.nz;
    z = r - y * P.nz;
}

// state s: F (27 planes), T, U (3 planes), each plane `cells` long
template <class R>
__global__ void __launch_bounds__(256) init_kernel(Params P, const uint8_t* __restrict__ solid, R* __restrict__ s) {
    const int N = P.nx * P.ny * P.nz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const bool sol = solid[c] != 0;
    const R ux = sol ? R(0) : R(P.u0[0]), uy = sol ? R(0) : R(P.u0[1]), uz = sol ? R(0) : R(P.u0[2]);
    for (int q = 0; q < 27; ++q) s[(size_t)q * N + c] = feq<R>(q, R(1), ux, uy, uz);
    s[(size_t)27 * N + c] = R(0);
    s[(size_t)28 * N + c] = ux;
    s[(size_t)29 * N + c] = uy;
    s[(size_t)30 * N + c] = uz;
}

// heat source of every cell from the m^3 parameter field (trilinear, clamped)
template <class R>
__global__ void __launch_bounds__(256) source_kernel(Params P, int m, const R* __restrict__ heat, R* __restrict__ src,
                                                     uint8_t* __restrict__ src_free) {
    const int N = P.nx * P.ny * P.nz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int xyz[3];
    cell_xyz<R>(P, c, xyz[0], xyz[1], xyz[2]);
    const int n[3] = {P.nx, P.ny, P.nz};
    int lo[3];
    float t[3];
    for (int a = 0; a < 3; ++a) {
        float p;
        source_weights(n[a], m, xyz[a], &p);
        lo[a] = min((int)floorf(p), m - 2 < 0 ? 0 : m - 2);
        t[a] = p - (float)lo[a];
    }
    R v = R(0);
    for (int k = 0; k < 8; ++k) {
        const int i = min(lo[0] + (k >> 2), m - 1), j = min(lo[1] + ((k >> 1) & 1), m - 1), l = min(lo[2] + (k & 1), m - 1);
        const float w = ((k >> 2) ? t[0] : 1.0f - t[0]) * (((k >> 1) & 1) ? t[1] : 1.0f - t[1]) * ((k & 1) ? t[2] : 1.0f - t[2]);
        v = v + R(w) * heat[(i * m + j) * m + l];
    }
    if (src_free) src_free[c] = (fabs((double)val(v)) < 0.006) ? 1 : 0;
    src[c] = rclamp(v, -0.006f, 0.006f);
}

// per-cell mask of the solid flags of the 27 cells at offsets -1..1 (bit
// (dx + 1) * 9 + (dy + 1) * 3 + dz + 1; y and z periodic, outside in x -> 0): the
// update of a cell then tests its stencil's solid links from one register
__global__ void __launch_bounds__(256) nbr_kernel(Params P, const uint8_t* __restrict__ solid,
                                                  uint32_t* __restrict__ nb) {
    const int N = P.nx * P.ny * P.nz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int x, y, z;
    cell_xyz<float>(P, c, x, y, z);
    uint32_t m = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        const int xs = x + dx;
        if (xs < 0 || xs >= P.nx) continue;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
                if (solid[cidx(P, xs, wrap(y + dy, P.ny), wrap(z + dz, P.nz))])
                    m |= 1u << ((dx + 1) * 9 + (dy + 1) * 3 + dz + 1);
    }
    nb[c] = m;
}

// first: the initial state, whose populations are collided where they are (no
// stream before the first collision, as the reference's first step)
template <class R>
__global__ void __launch_bounds__(256) forward_kernel(Params P, const uint8_t* __restrict__ solid,
                                                      const uint32_t* __restrict__ nbm, const R* __restrict__ uin,
                                                      const R* __restrict__ src,
                                                      const R* __restrict__ si, R* __restrict__ so, int first) {
    const int N = P.nx * P.ny * P.nz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (solid[c]) {
        for (int q = 0; q < 28; ++q) so[(size_t)q * N + c] = si[(size_t)q * N + c];
        for (int k = 0; k < 3; ++k) so[(size_t)(28 + k) * N + c] = R(0);
        return;
    }
    int x, y, z;
    cell_xyz<R>(P, c, x, y, z);
    const R* T = si + (size_t)27 * N;
    const SolidNbrs sq{nbm[c], solid};
    int idx[14];
    R w[8], pre;
    const R Tn = temperature(P, T, sq, src[c], x, y, z, si[(size_t)28 * N + c], si[(size_t)29 * N + c],
                             si[(size_t)30 * N + c], idx, w, &pre);
    R g[27], u[3];
    for (int q = 0; q < 27; ++q) g[q] = first ? si[(size_t)q * N + c] : incoming(P, si, sq, uin, x, y, z, q);
    collide(P, g, Tn, u);
    for (int q = 0; q < 27; ++q) so[(size_t)q * N + c] = g[q];
    so[(size_t)27 * N + c] = Tn;
    for (int k = 0; k < 3; ++k) so[(size_t)(28 + k) * N + c] = u[k];
}

// ---- reverse pass (FP32) ----
// adjoints of the step's outputs (a1: F, T, U of state n+1) -> per-cell adjoint of
// the 27 incoming values (gbar) and of the new temperature (tbar); inflow adjoint
__global__ void __launch_bounds__(128) collide_adjoint_kernel(Params P, const uint8_t* __restrict__ solid,
                                                              const uint32_t* __restrict__ nbm,
                                                              const float* __restrict__ uin,
                                                              const float* __restrict__ src,
                                                              const float* __restrict__ s,
                                                              const float* __restrict__ a1,
                                                              float* __restrict__ gbar, float* __restrict__ tbar,
                                                              float* __restrict__ uin_bar, int first) {
    const int N = P.nx * P.ny * P.nz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    if (solid[c]) {
        tbar[c] = a1[(size_t)27 * N + c];
        return;
    }
    int x, y, z;
    cell_xyz<float>(P, c, x, y, z);
    const SolidNbrs sq{nbm[c], solid};
    int idx[14];
    float w[8], pre;
    const float Tn = temperature(P, s + (size_t)27 * N, sq, src[c], x, y, z, s[(size_t)28 * N + c],
                                 s[(size_t)29 * N + c], s[(size_t)30 * N + c], idx, w, &pre);
    float g[27];
    for (int q = 0; q < 27; ++q) g[q] = first ? s[(size_t)q * N + c] : incoming(P, s, sq, uin, x, y, z, q);
    float fbar[27], ubar[3];
    for (int q = 0; q < 27; ++q) fbar[q] = a1[(size_t)q * N + c];
    for (int k = 0; k < 3; ++k) ubar[k] = a1[(size_t)(28 + k) * N + c];
    float tb = a1[(size_t)27 * N + c];  // T^{n+1} is also read by the next step's update
    // the collision depends on g only through its kLow raw moments of order <= 2:
    // the reverse mode of the post-collision map gives their adjoints (and T's),
    // then the transposed moment map carries them back onto the 27 populations
    float m[kLow], mbar[kLow], tcol;
    low_moments(g, m);
    collide_low_adjoint<float>(P, m, Tn, fbar, ubar, mbar, &tcol);
    tb += tcol;
    for (int q = 0; q < 27; ++q) {
        const float cx = (float)cxq(q), cy = (float)cyq(q), cz = (float)czq(q);
        const float v = mbar[0] + cx * mbar[1] + cy * mbar[2] + cz * mbar[3] + cx * cx * mbar[4] +
                        cy * cy * mbar[5] + cz * cz * mbar[6] + cx * cy * mbar[7] + cx * cz * mbar[8] +
                        cy * cz * mbar[9];
        gbar[(size_t)q * N + c] = v;
    }
    tbar[c] = tb;
    if (x == 0 && !first) {
        // links from the x- face read feq(rho_in, u_in(y, z)): d feq / d u
        const float* u = uin + 3 * (y * P.nz + z);
        float ub[3] = {0.0f, 0.0f, 0.0f};
        for (int q = 0; q < 27; ++q) {
            if (cxq(q) != 1) continue;
            const float gb = gbar[(size_t)q * N + c];
            const float cq[3] = {(float)cxq(q), (float)cyq(q), (float)czq(q)};
            const float cu = cq[0] * u[0] + cq[1] * u[1] + cq[2] * u[2];
            const float wr = wq<float>(q) * P.rho_in;
            for (int a = 0; a < 3; ++a) ub[a] += gb * wr * (3.0f * cq[a] + 9.0f * cu * cq[a] - 3.0f * u[a]);
        }
        for (int a = 0; a < 3; ++a) uin_bar[3 * (y * P.nz + z) + a] += ub[a];
    }
}

// adjoint of the previous post-collision value F^n(c, q): every read of it
__global__ void __launch_bounds__(256) stream_adjoint_kernel(Params P, const uint8_t* __restrict__ solid,
                                                             const float* __restrict__ a1,
                                                             const float* __restrict__ gbar, float* __restrict__ a0) {
    const int N = P.nx * P.ny * P.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 27 * N) return;
    const int q = t / N, c = t - q * N;
    if (solid[c]) {
        a0[(size_t)q * N + c] = a1[(size_t)q * N + c];  // solid cells keep their populations
        return;
    }
    int x, y, z;
    cell_xyz<float>(P, c, x, y, z);
    const int cx = cxq(q);
    float acc = 0.0f;
    // pulled by the neighbour x + c_q
    const int dx = x + cx, dy = wrap(y + cyq(q), P.ny), dz = wrap(z + czq(q), P.nz);
    if (dx >= 0 && dx < P.nx) {
        const int d = cidx(P, dx, dy, dz);
        if (!solid[d]) acc += gbar[(size_t)q * N + d];
        else acc += gbar[(size_t)oppq(q) * N + c];  // the cell's own link opp(q) bounces off d
    } else if (dx >= P.nx) {
        // link opp(q) of this x+ face cell takes the convective rule; it bounces when x - 1 is solid
        if (solid[cidx(P, x - 1, y, z)]) acc += gbar[(size_t)oppq(q) * N + c];
    }
    // convective reads by the x+ face cell of the same (y, z) for links with c_x = -1
    if (cx == -1 && (x == P.nx - 2 || x == P.nx - 3)) {
        const int d = cidx(P, P.nx - 1, y, z);
        if (!solid[d]) {
            const bool s1 = solid[cidx(P, P.nx - 2, y, z)] != 0;
            const bool s2 = solid[cidx(P, P.nx - 3, y, z)] != 0;
            const float gb = gbar[(size_t)q * N + d];
            if (x == P.nx - 2 && !s1) acc += (s2 ? 1.0f : 1.15f) * gb;
            if (x == P.nx - 3 && !s1 && !s2) acc += -0.15f * gb;
        }
    }
    a0[(size_t)q * N + c] = acc;
}

// temperature and velocity adjoints of state n from tbar (the adjoint of T^{n+1})
__global__ void __launch_bounds__(256) thermal_adjoint_kernel(Params P, const uint8_t* __restrict__ solid,
                                                              const uint32_t* __restrict__ nbm,
                                                              const float* __restrict__ src,
                                                              const uint8_t* __restrict__ src_free,
                                                              const float* __restrict__ s,
                                                              const float* __restrict__ tbar, float* __restrict__ a0,
                                                              float* __restrict__ src_bar) {
    const int N = P.nx * P.ny * P.nz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float* Tb = a0 + (size_t)27 * N;
    if (solid[c]) {
        atomicAdd(&Tb[c], tbar[c]);  // the solid cell keeps its temperature
        for (int k = 0; k < 3; ++k) a0[(size_t)(28 + k) * N + c] = 0.0f;
        return;
    }
    int x, y, z;
    cell_xyz<float>(P, c, x, y, z);
    const float* T = s + (size_t)27 * N;
    const float ux = s[(size_t)28 * N + c], uy = s[(size_t)29 * N + c], uz = s[(size_t)30 * N + c];
    int idx[14];
    float w[8], pre;
    (void)temperature(P, T, SolidNbrs{nbm[c], solid}, src[c], x, y, z, ux, uy, uz, idx, w, &pre);
    const float tb = (pre > 1.0f || pre < -1.0f) ? 0.0f : tbar[c];
    float self = -6.0f * P.alpha * tb;
    for (int k = 0; k < 8; ++k) {
        if (idx[k] >= 0) atomicAdd(&Tb[idx[k]], tb * w[k]);
        else self += tb * w[k];
    }
    for (int k = 0; k < 6; ++k) {
        if (idx[8 + k] >= 0) atomicAdd(&Tb[idx[8 + k]], tb * P.alpha);
        else self += tb * P.alpha;
    }
    atomicAdd(&Tb[c], self);
    // velocity: the departure point x - u moves the trilinear weights
    for (int a = 0; a < 3; ++a) {
        const Dual<float> du[3] = {Dual<float>(ux, a == 0 ? 1.0f : 0.0f), Dual<float>(uy, a == 1 ? 1.0f : 0.0f),
                                   Dual<float>(uz, a == 2 ? 1.0f : 0.0f)};
        // the stencil reads the stored temperatures (no tangent)
        const int x0 = (int)floorf((float)x - ux), y0 = (int)floorf((float)y - uy), z0 = (int)floorf((float)z - uz);
        const Dual<float> fx = Dual<float>((float)x) - du[0] - Dual<float>((float)x0);
        const Dual<float> fy = Dual<float>((float)y) - du[1] - Dual<float>((float)y0);
        const Dual<float> fz = Dual<float>((float)z) - du[2] - Dual<float>((float)z0);
        Dual<float> adv(0.0f);
        for (int k = 0; k < 8; ++k) {
            const Dual<float> wx = (k >> 2) ? fx : Dual<float>(1.0f) - fx;
            const Dual<float> wy = ((k >> 1) & 1) ? fy : Dual<float>(1.0f) - fy;
            const Dual<float> wz = (k & 1) ? fz : Dual<float>(1.0f) - fz;
            const float tv = idx[k] >= 0 ? T[idx[k]] : T[c];
            adv = adv + wx * wy * wz * tv;
        }
        a0[(size_t)(28 + a) * N + c] = tb * adv.d;
    }
    if (src_free[c]) src_bar[c] += tb;
}

__global__ void __launch_bounds__(256) zero_kernel(float* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

struct Model {
    Params P{};
    int m = 0, steps = 0, every = 1, nprobe = 0, N = 0;
    std::vector<int> probes;
    std::vector<float> target;
    uint8_t* solid = nullptr;
    uint32_t* nbmask = nullptr;  // neighbour solid mask of each cell (nbr_kernel)
    uint8_t* src_free = nullptr;
    float *uin = nullptr, *heat = nullptr, *src = nullptr;
    float *ckpt = nullptr, *seg = nullptr;
    float *a0 = nullptr, *a1 = nullptr, *gbar = nullptr, *tbar = nullptr, *src_bar = nullptr, *uin_bar = nullptr;
    double *uin64 = nullptr, *heat64 = nullptr, *src64 = nullptr, *s64[2] = {nullptr, nullptr};
    std::vector<float> final_u;  // probe velocities after the last forward
    bool forward_done = false;
    std::vector<void*> allocs;
    bool alloc(void** p, size_t bytes) {
        if (!hip_ok(hipMalloc(p, std::max<size_t>(bytes, 4)), "hipMalloc")) return false;
        allocs.push_back(*p);
        return true;
    }
    ~Model() {
        for (void* p : allocs) (void)hipFree(p);
    }
};

std::unordered_map<long long, std::unique_ptr<Model>> g_models;
long long g_next_model = 1;

dim3 blocks(size_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

float* state(Model& M, float* base, int i) { return base + (size_t)i * kState * M.N; }

bool run_forward(Model& M) {
    hipStream_t st = rt().stream;
    const int N = M.N;
    hipLaunchKernelGGL(source_kernel<float>, blocks(N), dim3(256), 0, st, M.P, M.m, M.heat, M.src, M.src_free);
    hipLaunchKernelGGL(init_kernel<float>, blocks(N), dim3(256), 0, st, M.P, M.solid, state(M, M.ckpt, 0));
    // seg[0..1] ping-pong; checkpoint every M.every steps
    float* cur = state(M, M.ckpt, 0);
    for (int n = 0; n < M.steps; ++n) {
        float* nxt = ((n + 1) % M.every == 0) ? state(M, M.ckpt, (n + 1) / M.every) : state(M, M.seg, (n & 1));
        if (nxt == cur) nxt = state(M, M.seg, 1 - (n & 1));
        hipLaunchKernelGGL(forward_kernel<float>, blocks(N), dim3(256), 0, st, M.P, M.solid, M.nbmask, M.uin, M.src,
                           cur, nxt, n == 0 ? 1 : 0);
        cur = nxt;
    }
    // probe velocities of the final state
    std::vector<float> u(3 * (size_t)N);
    if (!hip_ok(hipMemcpyAsync(u.data(), cur + (size_t)28 * N, u.size() * 4, hipMemcpyDeviceToHost, st), "readback") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return false;
    M.final_u.resize(3 * (size_t)M.nprobe);
    for (int i = 0; i < M.nprobe; ++i)
        for (int k = 0; k < 3; ++k) M.final_u[3 * i + k] = u[(size_t)k * N + M.probes[i]];
    M.forward_done = true;
    return true;
}

double loss_of(const Model& M, const std::vector<double>& u) {
    double L = 0.0;
    for (size_t i = 0; i < u.size(); ++i) {
        const double d = u[i] - (double)M.target[i];
        L += d * d;
    }
    return L;
}

bool run_backward(Model& M) {
    hipStream_t st = rt().stream;
    const int N = M.N;
    const size_t sN = (size_t)kState * N;
    hipLaunchKernelGGL(zero_kernel, blocks(sN), dim3(256), 0, st, M.a1, sN);
    hipLaunchKernelGGL(zero_kernel, blocks(N), dim3(256), 0, st, M.src_bar, (size_t)N);
    const size_t nin = 3 * (size_t)M.P.ny * M.P.nz;
    hipLaunchKernelGGL(zero_kernel, blocks(nin), dim3(256), 0, st, M.uin_bar, nin);
    // seed: dL/dU at the probes of the final state
    std::vector<float> seed(3 * (size_t)N, 0.0f);
    for (int i = 0; i < M.nprobe; ++i)
        for (int k = 0; k < 3; ++k)
            seed[(size_t)k * N + M.probes[i]] += 2.0f * (M.final_u[3 * i + k] - M.target[3 * i + k]);
    if (!hip_ok(hipMemcpyAsync(M.a1 + (size_t)28 * N, seed.data(), seed.size() * 4, hipMemcpyHostToDevice, st),
                "seed"))
        return false;
    const int segments = (M.steps + M.every - 1) / M.every;
    for (int sgi = segments - 1; sgi >= 0; --sgi) {
        const int start = sgi * M.every, len = std::min(M.every, M.steps - start);
        // replay the segment: seg[j] = state start + j
        if (!hip_ok(hipMemcpyAsync(state(M, M.seg, 0), state(M, M.ckpt, sgi), sN * 4, hipMemcpyDeviceToDevice, st),
                    "checkpoint"))
            return false;
        for (int j = 0; j < len - 1; ++j)
            hipLaunchKernelGGL(forward_kernel<float>, blocks(N), dim3(256), 0, st, M.P, M.solid, M.nbmask, M.uin, M.src,
                               state(M, M.seg, j), state(M, M.seg, j + 1), start + j == 0 ? 1 : 0);
        for (int j = len - 1; j >= 0; --j) {
            const float* s = state(M, M.seg, j);
            const int first = start + j == 0 ? 1 : 0;
            hipLaunchKernelGGL(collide_adjoint_kernel, blocks(N, 128), dim3(128), 0, st, M.P, M.solid, M.nbmask, M.uin, M.src, s,
                               M.a1, M.gbar, M.tbar, M.uin_bar, first);
            // the initial populations depend on no parameter: their adjoint is not needed
            if (!first)
                hipLaunchKernelGGL(stream_adjoint_kernel, blocks(27 * (size_t)N), dim3(256), 0, st, M.P, M.solid, M.a1,
                                   M.gbar, M.a0);
            hipLaunchKernelGGL(zero_kernel, blocks(4 * (size_t)N), dim3(256), 0, st, M.a0 + (size_t)27 * N,
                               4 * (size_t)N);
            hipLaunchKernelGGL(thermal_adjoint_kernel, blocks(N), dim3(256), 0, st, M.P, M.solid, M.nbmask, M.src, M.src_free, s,
                               M.tbar, M.a0, M.src_bar);
            std::swap(M.a0, M.a1);
        }
    }
    return hip_ok(hipGetLastError(), "adjoint") && hip_ok(hipStreamSynchronize(st), "sync");
}

}  // namespace adj
}  // namespace aero

using namespace aero;
using namespace aero::adj;

// ---------------------------------------------------------------- C ABI (extension: no counterpart in the reference)

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_create(const AeroLbmAdjointConfig* cfg, const uint8_t* solid,
                                                            const int* probe_cells, const float* target_uvw,
                                                            long long* out_handle) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!cfg || !solid || !probe_cells || !target_uvw || !out_handle || cfg->n < 4 || cfg->heat_grid < 2 ||
        cfg->steps < 1 || cfg->checkpoint_every < 1 || cfg->probe_count < 1 || !(cfg->nu > 0.0f)) {
        set_error("adjoint_create: invalid arguments");
        return 0;
    }
    if (!ensure_device()) return 0;
    auto M = std::make_unique<Model>();
    const int n = cfg->n;
    M->N = n * n * n;
    M->P.nx = M->P.ny = M->P.nz = n;
    M->P.tau_shear = M->P.tau_normal = 0.5 + 3.0 * (double)cfg->nu;
    M->P.beta = cfg->beta;
    M->P.g_axis = 2;
    M->P.g_sign = -1;
    M->P.alpha = cfg->alpha;
    M->P.kmax = 0.57735026919f * 0.60f;
    M->P.rho_in = 1.0f;
    for (int k = 0; k < 3; ++k) M->P.u0[k] = cfg->u0[k];
    M->m = cfg->heat_grid;
    M->steps = cfg->steps;
    M->every = cfg->checkpoint_every;
    M->nprobe = cfg->probe_count;
    M->probes.assign(probe_cells, probe_cells + cfg->probe_count);
    for (int p : M->probes)
        if (p < 0 || p >= M->N) {
            set_error("adjoint_create: probe cell out of range");
            return 0;
        }
    M->target.assign(target_uvw, target_uvw + 3 * (size_t)cfg->probe_count);
    const size_t N = M->N, sN = (size_t)kState * N;
    const int nck = M->steps / M->every + 1;
    bool ok = M->alloc((void**)&M->solid, N) && M->alloc((void**)&M->nbmask, N * 4) &&
              M->alloc((void**)&M->src_free, N) &&
              M->alloc((void**)&M->uin, 3 * (size_t)n * n * 4) && M->alloc((void**)&M->heat, (size_t)M->m * M->m * M->m * 4) &&
              M->alloc((void**)&M->src, N * 4) && M->alloc((void**)&M->ckpt, (size_t)nck * sN * 4) &&
              M->alloc((void**)&M->seg, (size_t)std::max(2, M->every) * sN * 4) && M->alloc((void**)&M->a0, sN * 4) &&
              M->alloc((void**)&M->a1, sN * 4) && M->alloc((void**)&M->gbar, 27 * N * 4) &&
              M->alloc((void**)&M->tbar, N * 4) && M->alloc((void**)&M->src_bar, N * 4) &&
              M->alloc((void**)&M->uin_bar, 3 * (size_t)n * n * 4);
    if (!ok) return 0;
    if (!hip_ok(hipMemcpy(M->solid, solid, N, hipMemcpyHostToDevice), "upload solid")) return 0;
    hipLaunchKernelGGL(nbr_kernel, blocks(N), dim3(256), 0, rt().stream, M->P, M->solid, M->nbmask);
    if (!hip_ok(hipGetLastError(), "nbr_kernel") || !hip_ok(hipStreamSynchronize(rt().stream), "sync")) return 0;
    const long long h = g_next_model++;
    g_models[h] = std::move(M);
    *out_handle = h;
    return 1;
}

static Model* find_model(long long h) {
    auto it = g_models.find(h);
    if (it == g_models.end()) {
        set_error("adjoint: unknown handle");
        return nullptr;
    }
    return it->second.get();
}

// FP32 forward with checkpoints; out_loss = sum over probes of |u - target|^2
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_forward(long long h, const float* inflow, const float* heat,
                                                             double* out_loss, float* out_probe_u) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    Model* M = find_model(h);
    if (!M || !inflow || !heat) return 0;
    const size_t nin = 3 * (size_t)M->P.ny * M->P.nz, nh = (size_t)M->m * M->m * M->m;
    if (!hip_ok(hipMemcpy(M->uin, inflow, nin * 4, hipMemcpyHostToDevice), "upload inflow") ||
        !hip_ok(hipMemcpy(M->heat, heat, nh * 4, hipMemcpyHostToDevice), "upload heat") || !run_forward(*M))
        return 0;
    std::vector<double> u(M->final_u.begin(), M->final_u.end());
    if (out_loss) *out_loss = loss_of(*M, u);
    if (out_probe_u) std::copy(M->final_u.begin(), M->final_u.end(), out_probe_u);
    return 1;
}

// gradient of the last forward's loss: grad_inflow (3 per (y, z)), grad_heat (m^3)
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_backward(long long h, float* grad_inflow, float* grad_heat) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    Model* M = find_model(h);
    if (!M || !grad_inflow || !grad_heat) return 0;
    if (!M->forward_done) {
        set_error("adjoint_backward: run aero_lbm_adjoint_forward first");
        return 0;
    }
    if (!run_backward(*M)) return 0;
    const int N = M->N;
    const size_t nin = 3 * (size_t)M->P.ny * M->P.nz;
    std::vector<float> sb(N);
    if (!hip_ok(hipMemcpy(grad_inflow, M->uin_bar, nin * 4, hipMemcpyDeviceToHost), "readback") ||
        !hip_ok(hipMemcpy(sb.data(), M->src_bar, (size_t)N * 4, hipMemcpyDeviceToHost), "readback"))
        return 0;
    // heat-source field: transpose of the trilinear upsampling (source_kernel)
    const int m = M->m, n = M->P.nx;
    std::vector<double> acc((size_t)m * m * m, 0.0);
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
            for (int z = 0; z < n; ++z) {
                const float g = sb[(size_t)(x * n + y) * n + z];
                if (g == 0.0f) continue;
                const int xyz[3] = {x, y, z};
                int lo[3];
                float t[3];
                for (int a = 0; a < 3; ++a) {
                    const float p = std::min(std::max(((float)xyz[a] + 0.5f) * (float)m / (float)n - 0.5f, 0.0f),
                                             (float)(m - 1));
                    lo[a] = std::min((int)std::floor(p), std::max(0, m - 2));
                    t[a] = p - (float)lo[a];
                }
                for (int k = 0; k < 8; ++k) {
                    const int i = std::min(lo[0] + (k >> 2), m - 1), j = std::min(lo[1] + ((k >> 1) & 1), m - 1),
                              l = std::min(lo[2] + (k & 1), m - 1);
                    const float w = ((k >> 2) ? t[0] : 1.0f - t[0]) * (((k >> 1) & 1) ? t[1] : 1.0f - t[1]) *
                                    ((k & 1) ? t[2] : 1.0f - t[2]);
                    acc[(size_t)(i * m + j) * m + l] += (double)g * w;
                }
            }
    for (size_t i = 0; i < acc.size(); ++i) grad_heat[i] = (float)acc[i];
    return 1;
}

// FP64 forward of the same model (no checkpoints), for finite-difference checks
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_adjoint_forward_fp64(long long h, const double* inflow, const double* heat,
                                                                  double* out_loss) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    Model* M = find_model(h);
    if (!M || !inflow || !heat || !out_loss) return 0;
    const size_t N = M->N, sN = (size_t)kState * N, nin = 3 * (size_t)M->P.ny * M->P.nz,
                 nh = (size_t)M->m * M->m * M->m;
    if (!M->s64[0]) {
        if (!M->alloc((void**)&M->uin64, nin * 8) || !M->alloc((void**)&M->heat64, nh * 8) ||
            !M->alloc((void**)&M->src64, N * 8) || !M->alloc((void**)&M->s64[0], sN * 8) ||
            !M->alloc((void**)&M->s64[1], sN * 8))
            return 0;
    }
    hipStream_t st = rt().stream;
    if (!hip_ok(hipMemcpy(M->uin64, inflow, nin * 8, hipMemcpyHostToDevice), "upload") ||
        !hip_ok(hipMemcpy(M->heat64, heat, nh * 8, hipMemcpyHostToDevice), "upload"))
        return 0;
    hipLaunchKernelGGL(source_kernel<double>, blocks(N), dim3(256), 0, st, M->P, M->m, M->heat64, M->src64,
                       (uint8_t*)nullptr);
    hipLaunchKernelGGL(init_kernel<double>, blocks(N), dim3(256), 0, st, M->P, M->solid, M->s64[0]);
    for (int s = 0; s < M->steps; ++s)
        hipLaunchKernelGGL(forward_kernel<double>, blocks(N), dim3(256), 0, st, M->P, M->solid, M->nbmask, M->uin64, M->src64,
                           M->s64[s & 1], M->s64[1 - (s & 1)], s == 0 ? 1 : 0);
    std::vector<double> u(3 * N);
    const double* fin = M->s64[M->steps & 1];
    if (!hip_ok(hipMemcpyAsync(u.data(), fin + 28 * N, u.size() * 8, hipMemcpyDeviceToHost, st), "readback") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return 0;
    std::vector<double> pu(3 * (size_t)M->nprobe);
    for (int i = 0; i < M->nprobe; ++i)
        for (int k = 0; k < 3; ++k) pu[3 * i + k] = u[(size_t)k * N + M->probes[i]];
    *out_loss = loss_of(*M, pu);
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT void aero_lbm_adjoint_release(long long h) {
    std::lock_guard<std::mutex> lk(rt().mu);
    g_models.erase(h);
}
