// D3Q19 mesoscale step on the GPU: the reference's CPU step
// (ref: native/src/aero_lbm_mesoscale.cpp:548-738, targets 393-530, seeding
// 298-333) in its own operation order, so results match the CPU step bit for
// bit (this file is compiled with -ffp-contract=off), behind the reference's
// mesoscale C ABI (aero_lbm_mesoscale_*, mesoscale.cpp:1397-1509).
//
// Layout: populations f and g as 19 FP32 planes each (structure of arrays,
// q-major, cell = (x*ny + y)*nz + z, y vertical). The 25-channel forcing is
// uploaded when it changes and reduced once into per-cell coefficient planes
// (each term computed exactly as the reference computes it), so a step reads
// no strided forcing records.
//
// One step = two launches:
//   collide: stream the previous post-collision state into the cell (pull, with
//            the reference's wall, floor, open-boundary and sponge rules), form
//            the cell targets, two-relaxation-time collision of f and g, write
//            the new post-collision planes and the next surface / deep fields;
//   macro  : stream again to form the macroscopic fields the reference's third
//            pass computes (ambient temperature for the next step's targets and,
//            when the caller asks for it, the 6-channel state).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <vector>

#include "../../include/aero_lbm_capi.h"
#include "runtime_internal.h"

namespace aero {
namespace meso {

constexpr int kQ = 19;
constexpr int kForcing = AERO_LBM_MESOSCALE_FORCING_CHANNELS;  // 25
constexpr int kState = AERO_LBM_MESOSCALE_STATE_CHANNELS;      // 6

// D3Q19 directions in the reference's order (mesoscale.cpp:146-176)
__host__ __device__ constexpr int cx(int q) {
    return q == 1 || q == 7 || q == 9 || q == 11 || q == 13 ? 1 : (q == 2 || q == 8 || q == 10 || q == 12 || q == 14 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int q) {
    return q == 3 || q == 7 || q == 10 || q == 15 || q == 17 ? 1 : (q == 4 || q == 8 || q == 9 || q == 16 || q == 18 ? -1 : 0);
}
__host__ __device__ constexpr int cz(int q) {
    return q == 5 || q == 11 || q == 14 || q == 15 || q == 18 ? 1 : (q == 6 || q == 12 || q == 13 || q == 16 || q == 17 ? -1 : 0);
}
__host__ __device__ constexpr int opp(int q) { return q == 0 ? 0 : (q & 1 ? q + 1 : q - 1); }
__device__ __forceinline__ float wq(int q) { return q == 0 ? 1.0f / 3.0f : (q < 7 ? 1.0f / 18.0f : 1.0f / 36.0f); }

// per-cell coefficient planes derived from the forcing
enum Stat {
    kAmbT,   // ambient target (+ nested delta); also the open-boundary temperature
    kDeepT,  // deep-ground target
    kSurfT,  // surface target (+ nested delta); also the solid wall temperature
    kDrag,
    kDamp,   // 1 - 0.12 humidity
    kKh,     // 0.018 + 0.004 humidity
    kT4, kT5, kT6, kT7,  // forcing terms of the thermal source
    kN4, kN5,            // forcing terms of the next surface temperature
    kSlopeX, kSlopeZ,
    kBx, kBy, kBz,       // background wind in lattice units (targets and open boundaries)
    kStats
};

struct Params {
    int nx, ny, nz, cells;
    float vden;  // max(1e-6, velocity scale)
    float amb_relax, deep_relax, surf_relax, wind_relax;
    float oh_e, oh_o, ot_e, ot_o;
    float vscale, max_speed;
    // coefficient planes holding one value in every cell (bit k for plane k): the
    // kernels take that value from here instead of reading the plane
    uint32_t uniform;
    float uval[kStats];
};

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float clamp_speed(float u) { return clampf(u, -0.35f, 0.35f); }

__device__ __forceinline__ float hfeq(int q, float rho, float ux, float uy, float uz) {
    const float cu = 3.0f * ((float)cx(q) * ux + (float)cy(q) * uy + (float)cz(q) * uz);
    const float uu = ux * ux + uy * uy + uz * uz;
    return wq(q) * rho * (1.0f + cu + 0.5f * cu * cu - 1.5f * uu);
}
__device__ __forceinline__ float tfeq(int q, float t, float ux, float uy, float uz) {
    const float cu = 3.0f * ((float)cx(q) * ux + (float)cy(q) * uy + (float)cz(q) * uz);
    return wq(q) * t * (1.0f + cu);
}

__device__ __forceinline__ int idx(const Params& P, int x, int y, int z) { return (x * P.ny + y) * P.nz + z; }

// coefficient plane k at cell i (k is a compile-time constant at every call, so the
// uniform test is a wave-uniform branch on a kernel parameter)
__device__ __forceinline__ float sv(const Params& P, const float* __restrict__ S, int k, size_t N, size_t i) {
    return ((P.uniform >> k) & 1u) ? P.uval[k] : S[k * N + i];
}

// ---------------------------------------------------------------- forcing reduction
// the 17 coefficient plane values of cell i (ref: mesoscale.cpp:372-391 for the slope)
__device__ __forceinline__ void plane_values(const Params& P, const float* __restrict__ F, int i, float v[kStats]) {
    const int z = i % P.nz, y = (i / P.nz) % P.ny, x = i / (P.nz * P.ny);
    const float* p = F + (size_t)i * kForcing;
    v[kAmbT] = p[2] + p[22];
    v[kDeepT] = p[3];
    v[kSurfT] = p[4] + p[23];
    const float rough = smax(0.0f, p[5]);
    v[kDrag] = clampf(rough * 0.02f, 0.0f, 0.25f);
    const float hum = clampf(p[9] + p[11] + p[17], 0.0f, 1.0f);
    v[kDamp] = 1.0f - 0.12f * hum;
    v[kKh] = 0.018f + 0.004f * hum;
    v[kT4] = 0.08f * p[10] * (0.5f + 0.5f * hum);
    v[kT5] = 0.10f * p[16] * (0.45f + 0.55f * hum);
    v[kT6] = 0.06f * p[18];
    v[kT7] = 0.06f * p[19];
    v[kN4] = 0.04f * p[10];
    v[kN5] = 0.08f * p[16];
    // terrain slope from the clamped neighbours' heights
    const int xp = min(P.nx - 1, x + 1), xm = max(0, x - 1), zp = min(P.nz - 1, z + 1), zm = max(0, z - 1);
    v[kSlopeX] = 0.5f * (F[(size_t)idx(P, xp, y, z) * kForcing] - F[(size_t)idx(P, xm, y, z) * kForcing]);
    v[kSlopeZ] = 0.5f * (F[(size_t)idx(P, x, y, zp) * kForcing] - F[(size_t)idx(P, x, y, zm) * kForcing]);
    const float bx = p[6] + p[20] + p[12] + p[14];
    const float bz = p[7] + p[21] + p[13] + p[15];
    const float by = p[19] + p[18];
    v[kBx] = clamp_speed(bx / P.vden);
    v[kBy] = clamp_speed(by / P.vden);
    v[kBz] = clamp_speed(bz / P.vden);
}

// uni[kStats] (as bits) cleared before prep_kernel ORs into it
__global__ void uniform_clear_kernel(float* __restrict__ uni) { reinterpret_cast<uint32_t*>(uni)[kStats] = 0u; }

// Coefficient planes and solid flags from the forcing. Each thread also derives
// cell 0's values (the same arithmetic on the same inputs, so bitwise equal where
// the planes are uniform) and flags the planes where its cell differs: uni[k] =
// cell 0's value of plane k, bit k of uni[kStats] = plane k is not uniform. One
// atomic per wavefront, skipped once the bits it would set are visible.
__global__ void __launch_bounds__(256) prep_kernel(Params P, const float* __restrict__ F, float* __restrict__ S,
                                                   uint8_t* __restrict__ solid, float* __restrict__ uni) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = (size_t)P.cells;
    uint32_t differs = 0;
    if (i < P.cells) {
        solid[i] = F[(size_t)i * kForcing + 24] > 0.5f ? 1 : 0;
        float v[kStats], r[kStats];
        plane_values(P, F, i, v);
        plane_values(P, F, 0, r);
#pragma unroll
        for (int k = 0; k < kStats; ++k) {
            S[k * N + i] = v[k];
            if (__float_as_uint(v[k]) != __float_as_uint(r[k])) differs |= 1u << k;
            if (i == 0) uni[k] = v[k];
        }
    }
    uint32_t wave = 0;
#pragma unroll
    for (int k = 0; k < kStats; ++k)
        if (__ballot((differs >> k) & 1u)) wave |= 1u << k;
    if (wave && (threadIdx.x & 63) == 0) {
        uint32_t* flags = reinterpret_cast<uint32_t*>(uni) + kStats;
        const uint32_t seen = __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (wave & ~seen) atomicOr(flags, wave);
    }
}

// per-cell link mask: bit q (1..18) when direction q's source cell lies in the
// domain and is solid, bit 31 when the cell itself is solid; rebuilt with the
// coefficient planes, so a pull tests one bit instead of loading a flag byte
constexpr uint32_t kSelfSolid = 1u << 31;

__global__ void __launch_bounds__(256) mask_kernel(Params P, const uint8_t* __restrict__ solid,
                                                   uint32_t* __restrict__ links) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.cells) return;
    const int z = i % P.nz, y = (i / P.nz) % P.ny, x = i / (P.nz * P.ny);
    uint32_t m = solid[i] ? kSelfSolid : 0u;
#pragma unroll
    for (int q = 1; q < kQ; ++q) {
        const int sx = x - cx(q), sy = y - cy(q), sz = z - cz(q);
        if (sx >= 0 && sx < P.nx && sy >= 0 && sy < P.ny && sz >= 0 && sz < P.nz && solid[idx(P, sx, sy, sz)])
            m |= 1u << q;
    }
    links[i] = m;
}

// ---------------------------------------------------------------- seeding (mesoscale.cpp:298-333)
__global__ void __launch_bounds__(256) seed_kernel(Params P, const float* __restrict__ F, float* __restrict__ f,
                                                   float* __restrict__ g, float* __restrict__ ambient,
                                                   float* __restrict__ surface, float* __restrict__ deep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.cells) return;
    const size_t N = (size_t)P.cells;
    const float* p = F + (size_t)i * kForcing;
    const float amb_t = p[2] + p[22];
    const float surf_t = p[4] + p[23];
    const bool solid = p[24] > 0.5f;
    const float bx = p[6] + p[20];
    const float bz = p[7] + p[21];
    const float by = p[19] + p[18];
    const float ux = clamp_speed((solid ? 0.0f : bx) / P.vden);
    const float uy = clamp_speed((solid ? 0.0f : by) / P.vden);
    const float uz = clamp_speed((solid ? 0.0f : bz) / P.vden);
    const float amb = solid ? surf_t : amb_t;
    ambient[i] = amb;
    deep[i] = p[3];
    surface[i] = surf_t;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        f[q * N + i] = hfeq(q, 1.0f, ux, uy, uz);
        g[q * N + i] = tfeq(q, amb, ux, uy, uz);
    }
}

// ---------------------------------------------------------------- streaming rules
// Value of direction q arriving at fluid cell (x, y, z) from the previous
// post-collision state (mesoscale.cpp:643-690).
struct Pulled {
    float f, g;
};

__device__ __forceinline__ Pulled pull(const Params& P, const float* __restrict__ pf, const float* __restrict__ pg,
                                       const float* __restrict__ S, uint32_t links, int i, int x, int y, int z, int q,
                                       float ea, float bgx, float bgy, float bgz, float btemp) {
    const size_t N = (size_t)P.cells;
    const int sx = x - cx(q), sy = y - cy(q), sz = z - cz(q);
    Pulled v;
    bool bounced = false;
    if (sx >= 0 && sx < P.nx && sy >= 0 && sy < P.ny && sz >= 0 && sz < P.nz) {
        const int src = idx(P, sx, sy, sz);
        if ((links >> q) & 1u) {
            v.f = pf[opp(q) * N + i];
            v.g = tfeq(q, sv(P, S, kSurfT, N, src), 0.0f, 0.0f, 0.0f);
            bounced = true;
        } else {
            v.f = pf[q * N + src];
            v.g = pg[q * N + src];
        }
    } else if (sy < 0) {
        v.f = pf[opp(q) * N + i];
        v.g = tfeq(q, sv(P, S, kSurfT, N, i), 0.0f, 0.0f, 0.0f);
        bounced = true;
    } else {
        v.f = hfeq(q, 1.0f, bgx, bgy, bgz);
        v.g = tfeq(q, btemp, bgx, bgy, bgz);
    }
    if (ea > 0.0f && !bounced) {
        const float fb = hfeq(q, 1.0f, bgx, bgy, bgz);
        const float gb = tfeq(q, btemp, bgx, bgy, bgz);
        v.f = (1.0f - ea) * v.f + ea * fb;
        v.g = (1.0f - ea) * v.g + ea * gb;
    }
    return v;
}

// the same rules for a cell at least two cells from the side and top faces and
// above the floor: every source lies in the domain and no blending applies, so a
// pull is the neighbour's value or, from a solid neighbour, the bounce
__device__ __forceinline__ Pulled pull_interior(const Params& P, const float* __restrict__ pf, const float* __restrict__ pg,
                                                const float* __restrict__ S, uint32_t links, size_t N, int i, int q,
                                                int sxy, int sy1) {
    const int src = i - (cx(q) * sxy + cy(q) * sy1 + cz(q));
    Pulled v;
    if ((links >> q) & 1u) {
        v.f = pf[opp(q) * N + i];
        v.g = tfeq(q, sv(P, S, kSurfT, N, src), 0.0f, 0.0f, 0.0f);
    } else {
        v.f = pf[q * N + src];
        v.g = pg[q * N + src];
    }
    return v;
}

__device__ __forceinline__ bool interior(const Params& P, int x, int y, int z) {
    const int edge = min(min(x, z), min(P.nx - 1 - x, P.nz - 1 - z));
    return min(edge, P.ny - 1 - y) >= 2 && y >= 1;
}

__device__ __forceinline__ float edge_alpha(const Params& P, int x, int y, int z) {
    const int edge = min(min(x, z), min(P.nx - 1 - x, P.nz - 1 - z));
    const int top = P.ny - 1 - y;
    const int d = min(edge, top);
    return d <= 0 ? 1.0f : (d < 2 ? 0.5f : 0.0f);
}

// horizontal and vertical Laplacians with clamped neighbours (mesoscale.cpp:335-370)
__device__ __forceinline__ float hlap(const Params& P, const float* __restrict__ v, int x, int y, int z) {
    const float c = v[idx(P, x, y, z)];
    const float xp = v[idx(P, min(P.nx - 1, x + 1), y, z)];
    const float xm = v[idx(P, max(0, x - 1), y, z)];
    const float zp = v[idx(P, x, y, min(P.nz - 1, z + 1))];
    const float zm = v[idx(P, x, y, max(0, z - 1))];
    return xp + xm + zp + zm - 4.0f * c;
}
__device__ __forceinline__ float vlap(const Params& P, const float* __restrict__ v, int x, int y, int z) {
    if (P.ny <= 1) return 0.0f;
    const float c = v[idx(P, x, y, z)];
    const float yp = v[idx(P, x, min(P.ny - 1, y + 1), z)];
    const float ym = v[idx(P, x, max(0, y - 1), z)];
    return yp + ym - 2.0f * c;
}

// ---------------------------------------------------------------- collide
__global__ void __launch_bounds__(256) collide_kernel(Params P, const float* __restrict__ pf_in,
                                                      const float* __restrict__ pg_in, float* __restrict__ pf_out,
                                                      float* __restrict__ pg_out, const float* __restrict__ ambient,
                                                      const float* __restrict__ surf_in, const float* __restrict__ deep_in,
                                                      float* __restrict__ surf_out, float* __restrict__ deep_out,
                                                      const float* __restrict__ S, const uint32_t* __restrict__ links,
                                                      int local_first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.cells) return;
    const size_t N = (size_t)P.cells;
    const int z = i % P.nz, y = (i / P.nz) % P.ny, x = i / (P.nz * P.ny);
    const uint32_t lk = links[i];
    if (lk & kSelfSolid) {
        // solid targets (mesoscale.cpp:448-457); their populations are never read
        surf_out[i] = sv(P, S, kSurfT, N, i);
        deep_out[i] = sv(P, S, kDeepT, N, i);
        return;
    }
    const float bgx = sv(P, S, kBx, N, i), bgy = sv(P, S, kBy, N, i), bgz = sv(P, S, kBz, N, i), amb_t = sv(P, S, kAmbT, N, i);
    float f[kQ], g[kQ];
    if (local_first) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            f[q] = pf_in[q * N + i];
            g[q] = pg_in[q * N + i];
        }
    } else if (__all(interior(P, x, y, z))) {
        const int sy1 = P.nz, sxy = P.ny * P.nz;
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const Pulled v = pull_interior(P, pf_in, pg_in, S, lk, N, i, q, sxy, sy1);
            f[q] = v.f;
            g[q] = v.g;
        }
    } else {
        const float ea = edge_alpha(P, x, y, z);
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const Pulled v = pull(P, pf_in, pg_in, S, lk, i, x, y, z, q, ea, bgx, bgy, bgz, amb_t);
            f[q] = v.f;
            g[q] = v.g;
        }
    }
    // ---- targets (mesoscale.cpp:393-530)
    float rho = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f, amb = 0.0f;
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        rho += f[q];
        ux += f[q] * (float)cx(q);
        uy += f[q] * (float)cy(q);
        uz += f[q] * (float)cz(q);
        amb += g[q];
    }
    rho = smax(1.0e-6f, rho);
    ux /= rho;
    uy /= rho;
    uz /= rho;
    const float deep_t = sv(P, S, kDeepT, N, i), surf_t = sv(P, S, kSurfT, N, i);
    const float sx = sv(P, S, kSlopeX, N, i), sz = sv(P, S, kSlopeZ, N, i);
    const float s_hl = hlap(P, surf_in, x, y, z);
    const float s_vl = vlap(P, surf_in, x, y, z);
    const float d_vl = vlap(P, deep_in, x, y, z);
    const float a_hl = hlap(P, ambient, x, y, z);
    const float a_vl = vlap(P, ambient, x, y, z);
    const float deep_c = deep_in[i], surf_c = surf_in[i];
    const float ndeep = deep_c + P.deep_relax * (deep_t - deep_c) + 0.01f * d_vl;
    const float nsurf = surf_c + P.surf_relax * sv(P, S, kDamp, N, i) * (surf_t - surf_c) + 0.10f * (ndeep - surf_c) +
                        sv(P, S, kN4, N, i) + sv(P, S, kN5, N, i) + 0.05f * s_hl + 0.025f * s_vl;
    const float push = 0.0008f * (nsurf - amb);
    const float nsw = 1.0f - (float)y / smax(1.0f, (float)(P.ny - 1));
    const float lift = 0.0015f * (bgx * sx + bgz * sz) * nsw;
    const float blift = 0.0009f * (nsurf - amb) * nsw;
    const float steer_x = -0.005f * sx * bgz;
    const float steer_z = 0.005f * sz * bgx;
    const float drag = sv(P, S, kDrag, N, i);
    const float tux = clamp_speed(bgx + steer_x + push * sx - drag * ux);
    const float tuz = clamp_speed(bgz + steer_z + push * sz - drag * uz);
    const float tuy = clamp_speed(bgy + lift + blift - drag * uy);
    const float rux = clamp_speed((1.0f - P.wind_relax) * ux + P.wind_relax * tux);
    const float ruy = clamp_speed((1.0f - P.wind_relax) * uy + P.wind_relax * tuy);
    const float ruz = clamp_speed((1.0f - P.wind_relax) * uz + P.wind_relax * tuz);
    const float src = P.amb_relax * (amb_t - amb) + sv(P, S, kKh, N, i) * (nsurf - amb) + 0.01f * (ndeep - amb) +
                      sv(P, S, kT4, N, i) + sv(P, S, kT5, N, i) + sv(P, S, kT6, N, i) + sv(P, S, kT7, N, i) + 0.01f * a_hl +
                      0.02f * a_vl;
    const float teq = amb + src;
    surf_out[i] = nsurf;
    deep_out[i] = ndeep;
    // ---- two-relaxation-time collision (mesoscale.cpp:585-628)
    {
        const float fe = hfeq(0, rho, rux, ruy, ruz);
        const float ge = tfeq(0, teq, rux, ruy, ruz);
        pf_out[i] = f[0] - P.oh_e * (f[0] - fe);
        pg_out[i] = g[0] - P.ot_e * (g[0] - ge);
    }
#pragma unroll
    for (int q = 1; q < kQ; q += 2) {
        const int qo = q + 1;
        const float fe = hfeq(q, rho, rux, ruy, ruz), feo = hfeq(qo, rho, rux, ruy, ruz);
        const float ge = tfeq(q, teq, rux, ruy, ruz), geo = tfeq(qo, teq, rux, ruy, ruz);
        const float f_ev = 0.5f * (f[q] + f[qo]);
        const float f_od = 0.5f * (f[q] - f[qo]);
        const float fe_ev = 0.5f * (fe + feo);
        const float fe_od = 0.5f * (fe - feo);
        const float f_evp = f_ev - P.oh_e * (f_ev - fe_ev);
        const float f_odp = f_od - P.oh_o * (f_od - fe_od);
        pf_out[q * N + i] = f_evp + f_odp;
        pf_out[qo * N + i] = f_evp - f_odp;
        const float g_ev = 0.5f * (g[q] + g[qo]);
        const float g_od = 0.5f * (g[q] - g[qo]);
        const float ge_ev = 0.5f * (ge + geo);
        const float ge_od = 0.5f * (ge - geo);
        const float g_evp = g_ev - P.ot_e * (g_ev - ge_ev);
        const float g_odp = g_od - P.ot_o * (g_od - ge_od);
        pg_out[q * N + i] = g_evp + g_odp;
        pg_out[qo * N + i] = g_evp - g_odp;
    }
}

// ---------------------------------------------------------------- macroscopic pass (mesoscale.cpp:697-734)
__global__ void __launch_bounds__(256) macro_kernel(Params P, const float* __restrict__ pf, const float* __restrict__ pg,
                                                    float* __restrict__ ambient, const float* __restrict__ surface,
                                                    const float* __restrict__ deep, const float* __restrict__ S,
                                                    const uint32_t* __restrict__ links, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.cells) return;
    const size_t N = (size_t)P.cells;
    float amb, wx = 0.0f, wy = 0.0f, wz = 0.0f;
    const uint32_t lk = links[i];
    if (lk & kSelfSolid) {
        amb = sv(P, S, kSurfT, N, i);
    } else {
        const int z = i % P.nz, y = (i / P.nz) % P.ny, x = i / (P.nz * P.ny);
        const float ea = edge_alpha(P, x, y, z);
        // the background wind and temperature enter a pull only on the open faces
        // and in the blending band, both within one cell of the side and top faces
        const bool band = min(min(min(x, z), min(P.nx - 1 - x, P.nz - 1 - z)), P.ny - 1 - y) <= 1;
        float bgx = 0.0f, bgy = 0.0f, bgz = 0.0f, amb_t = 0.0f;
        if (band) {
            bgx = sv(P, S, kBx, N, i);
            bgy = sv(P, S, kBy, N, i);
            bgz = sv(P, S, kBz, N, i);
            amb_t = sv(P, S, kAmbT, N, i);
        }
        float rho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
        amb = 0.0f;
        if (!out && __all(interior(P, x, y, z))) {
            const int sy1 = P.nz, sxy = P.ny * P.nz;
#pragma unroll
            for (int q = 0; q < kQ; ++q) amb += pull_interior(P, pf, pg, S, lk, N, i, q, sxy, sy1).g;
        } else if (out) {
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const Pulled v = pull(P, pf, pg, S, lk, i, x, y, z, q, ea, bgx, bgy, bgz, amb_t);
                rho += v.f;
                jx += v.f * (float)cx(q);
                jy += v.f * (float)cy(q);
                jz += v.f * (float)cz(q);
                amb += v.g;
            }
            rho = smax(1.0e-6f, rho);
            const float ux = jx / rho, uy = jy / rho, uz = jz / rho;
            wx = clampf(ux * P.vscale, -P.max_speed, P.max_speed);
            wy = clampf(uy * P.vscale, -P.max_speed, P.max_speed);
            wz = clampf(uz * P.vscale, -P.max_speed, P.max_speed);
        } else {
#pragma unroll
            for (int q = 0; q < kQ; ++q) amb += pull(P, pf, pg, S, lk, i, x, y, z, q, ea, bgx, bgy, bgz, amb_t).g;
        }
    }
    ambient[i] = amb;
    if (out) {
        float* o = out + (size_t)i * kState;
        o[0] = amb;
        o[1] = deep[i];
        o[2] = surface[i];
        o[3] = wx;
        o[4] = wy;
        o[5] = wz;
    }
}

// ---------------------------------------------------------------- materialised stream
// The reference streams step n with step n's forcing. The collide kernel pulls
// at step n + 1, so before new forcing replaces the coefficient planes this
// kernel applies step n's stream with the old planes and stores the streamed
// state (solid cells: the reference's solid-cell values, mesoscale.cpp:633-641);
// the next collide then reads the cell's own values.
__global__ void __launch_bounds__(256) stream_kernel(Params P, const float* __restrict__ pf, const float* __restrict__ pg,
                                                     float* __restrict__ f_out, float* __restrict__ g_out,
                                                     const float* __restrict__ S, const uint32_t* __restrict__ links) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.cells) return;
    const size_t N = (size_t)P.cells;
    const uint32_t lk = links[i];
    if (lk & kSelfSolid) {
        const float wt = sv(P, S, kSurfT, N, i);
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            f_out[q * N + i] = hfeq(q, 1.0f, 0.0f, 0.0f, 0.0f);
            g_out[q * N + i] = tfeq(q, wt, 0.0f, 0.0f, 0.0f);
        }
        return;
    }
    const int z = i % P.nz, y = (i / P.nz) % P.ny, x = i / (P.nz * P.ny);
    const float ea = edge_alpha(P, x, y, z);
    const float bgx = sv(P, S, kBx, N, i), bgy = sv(P, S, kBy, N, i), bgz = sv(P, S, kBz, N, i), amb_t = sv(P, S, kAmbT, N, i);
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const Pulled v = pull(P, pf, pg, S, lk, i, x, y, z, q, ea, bgx, bgy, bgz, amb_t);
        f_out[q * N + i] = v.f;
        g_out[q * N + i] = v.g;
    }
}

// ---------------------------------------------------------------- host runtime
struct Context {
    AeroLbmMesoscaleConfig cfg{};
    bool initialized = false;
    bool local_first = false;
    int cur = 0;  // post-collision planes in use
    int scur = 0; // surface / deep buffers in use
    size_t cells = 0;
    float prep_vscale = NAN;  // velocity scale the coefficient planes were derived with
    std::vector<float> forcing_host;
    float* pf[2] = {nullptr, nullptr};
    float* pg[2] = {nullptr, nullptr};
    float* ambient = nullptr;
    float* surface[2] = {nullptr, nullptr};
    float* deep[2] = {nullptr, nullptr};
    float* stat = nullptr;
    float* uni_dev = nullptr;  // uniform-plane values and flags (prep_kernel)
    uint32_t uniform = 0;      // planes holding one value, and that value
    float uval[kStats] = {};
    uint8_t* solid = nullptr;
    uint32_t* links = nullptr;
    float* forcing_dev = nullptr;
    float* out_dev = nullptr;
    std::vector<void*> allocations;

    ~Context() { release(); }
    void release() {
        for (void* p : allocations) (void)hipFree(p);
        allocations.clear();
    }
    bool alloc(void** p, size_t bytes) {
        if (!hip_ok(hipMalloc(p, bytes), "hipMalloc")) return false;
        allocations.push_back(*p);
        return true;
    }
    bool allocate() {
        release();
        const size_t N = cells;
        bool ok = true;
        for (int k = 0; k < 2; ++k) {
            ok = ok && alloc((void**)&pf[k], kQ * N * 4) && alloc((void**)&pg[k], kQ * N * 4) &&
                 alloc((void**)&surface[k], N * 4) && alloc((void**)&deep[k], N * 4);
        }
        ok = ok && alloc((void**)&ambient, N * 4) && alloc((void**)&stat, kStats * N * 4) &&
             alloc((void**)&uni_dev, (kStats + 1) * 4) &&
             alloc((void**)&solid, N) && alloc((void**)&links, N * 4) && alloc((void**)&forcing_dev, kForcing * N * 4) &&
             alloc((void**)&out_dev, kState * N * 4);
        forcing_host.clear();
        prep_vscale = NAN;
        uniform = 0;
        initialized = false;
        return ok;
    }
};

struct State {
    std::unordered_map<long long, std::unique_ptr<Context>> contexts;
    long long next_key = 1;
    float tau_floor = 0.58f;  // lower clamp of both relaxation times (mesoscale.cpp:559-566)
};

static State& ms() {
    static State s;
    return s;
}

// hydro_core.cpp:7-32, thermal_core.cpp:5-14 (and their header helpers)
static bool transport_valid(const AeroLbmMesoscaleConfig& c) {
    return c.nx > 0 && c.ny > 0 && c.nz > 0 && std::isfinite(c.dx_m) && std::isfinite(c.dt_s) &&
           std::isfinite(c.molecular_nu_m2_s) && std::isfinite(c.prandtl_air) && std::isfinite(c.turbulent_prandtl) &&
           c.dx_m > 0.0f && c.dt_s > 0.0f && c.molecular_nu_m2_s > 0.0f && c.prandtl_air > 0.0f &&
           c.turbulent_prandtl > 0.0f;
}

static AeroLbmMesoscaleTransport derive(const AeroLbmMesoscaleConfig& c) {
    AeroLbmMesoscaleTransport t{};
    if (!transport_valid(c)) return t;
    t.velocity_scale_m_s_per_lattice = c.dx_m / c.dt_s;
    t.nu_molecular_lattice = c.molecular_nu_m2_s * c.dt_s / (c.dx_m * c.dx_m);
    t.tau_shear_molecular = 0.5f + 3.0f * t.nu_molecular_lattice;
    const float alpha = c.molecular_nu_m2_s / c.prandtl_air;
    t.alpha_molecular_lattice = alpha * c.dt_s / (c.dx_m * c.dx_m);
    t.tau_thermal_molecular = 0.5f + t.alpha_molecular_lattice / 0.25f;
    return t;
}

static bool config_valid(const AeroLbmMesoscaleConfig* c) {
    return c && c->abi_version == AERO_LBM_MESOSCALE_ABI_VERSION && c->struct_size >= sizeof(AeroLbmMesoscaleConfig) &&
           transport_valid(*c) && (long long)c->nx * c->ny * c->nz <= (1ll << 30);
}

static float odd_tau(float even) {
    const float d = std::max(0.03f, even - 0.5f);
    return std::clamp(0.5f + 0.25f / d, 0.55f, 2.5f);
}

static Params make_params(const Context& c) {
    const AeroLbmMesoscaleTransport t = derive(c.cfg);
    Params P{};
    P.nx = c.cfg.nx;
    P.ny = c.cfg.ny;
    P.nz = c.cfg.nz;
    P.cells = (int)c.cells;
    P.vscale = t.velocity_scale_m_s_per_lattice;
    P.vden = std::max(1.0e-6f, P.vscale);
    const float dt = std::max(1.0e-3f, c.cfg.dt_s);
    P.amb_relax = std::clamp(dt / 600.0f, 0.0f, 1.0f);
    P.deep_relax = std::clamp(dt / 2400.0f, 0.0f, 1.0f);
    P.surf_relax = std::clamp(dt / 300.0f, 0.0f, 1.0f);
    P.wind_relax = std::clamp(dt / 60.0f, 0.0f, 1.0f);
    const float floor = ms().tau_floor;
    const float tau_h = std::clamp(t.tau_shear_molecular, floor, 2.5f);
    const float tau_t = std::clamp(t.tau_thermal_molecular, floor, 2.5f);
    P.oh_e = 1.0f / tau_h;
    P.oh_o = 1.0f / odd_tau(tau_h);
    P.ot_e = 1.0f / tau_t;
    P.ot_o = 1.0f / odd_tau(tau_t);
    P.max_speed = 0.35f * P.vscale;
    P.uniform = c.uniform;
    std::memcpy(P.uval, c.uval, sizeof(P.uval));
    return P;
}

static dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// upload the forcing (when it changed) and derive the coefficient planes; seed on first use
static bool prepare(Context& c, Params& P, const float* forcing) {
    hipStream_t st = rt().stream;
    const size_t n = c.cells * kForcing;
    const bool changed = c.forcing_host.size() != n || std::memcmp(c.forcing_host.data(), forcing, n * 4) != 0;
    const bool replan = changed || !(c.prep_vscale == P.vscale);
    if (replan && c.initialized && !c.local_first) {
        // stream the pending post-collision state with the coefficients it belongs to
        hipLaunchKernelGGL(stream_kernel, grid1(c.cells), dim3(256), 0, st, P, c.pf[c.cur], c.pg[c.cur],
                           c.pf[1 - c.cur], c.pg[1 - c.cur], c.stat, c.links);
        c.cur ^= 1;
        c.local_first = true;
    }
    if (changed) {
        // the previous upload may still read forcing_host
        if (!hip_ok(hipStreamSynchronize(st), "sync")) return false;
        c.forcing_host.assign(forcing, forcing + n);
        if (!hip_ok(hipMemcpyAsync(c.forcing_dev, c.forcing_host.data(), n * 4, hipMemcpyHostToDevice, st),
                    "upload forcing"))
            return false;
    }
    if (replan) {
        hipLaunchKernelGGL(uniform_clear_kernel, dim3(1), dim3(1), 0, st, c.uni_dev);
        hipLaunchKernelGGL(prep_kernel, grid1(c.cells), dim3(256), 0, st, P, c.forcing_dev, c.stat, c.solid, c.uni_dev);
        hipLaunchKernelGGL(mask_kernel, grid1(c.cells), dim3(256), 0, st, P, c.solid, c.links);
        c.prep_vscale = P.vscale;
        float u[kStats + 1];
        if (!hip_ok(hipMemcpyAsync(u, c.uni_dev, sizeof(u), hipMemcpyDeviceToHost, st), "uniform planes") ||
            !hip_ok(hipStreamSynchronize(st), "sync"))
            return false;
        uint32_t differs;
        std::memcpy(&differs, &u[kStats], 4);
        c.uniform = ~differs & ((1u << kStats) - 1u);
        std::memcpy(c.uval, u, sizeof(c.uval));
        P.uniform = c.uniform;
        std::memcpy(P.uval, c.uval, sizeof(P.uval));
    }
    if (!c.initialized) {
        hipLaunchKernelGGL(seed_kernel, grid1(c.cells), dim3(256), 0, st, P, c.forcing_dev, c.pf[c.cur], c.pg[c.cur],
                           c.ambient, c.surface[c.scur], c.deep[c.scur]);
        c.initialized = true;
        c.local_first = true;
    }
    return hip_ok(hipGetLastError(), "mesoscale prepare");
}

static void launch_step(Context& c, const Params& P, bool with_output) {
    hipStream_t st = rt().stream;
    const int n = c.cur, s = c.scur;
    hipLaunchKernelGGL(collide_kernel, grid1(c.cells), dim3(256), 0, st, P, c.pf[n], c.pg[n], c.pf[1 - n],
                       c.pg[1 - n], c.ambient, c.surface[s], c.deep[s], c.surface[1 - s], c.deep[1 - s], c.stat,
                       c.links, c.local_first ? 1 : 0);
    c.cur ^= 1;
    c.scur ^= 1;
    c.local_first = false;
    hipLaunchKernelGGL(macro_kernel, grid1(c.cells), dim3(256), 0, st, P, c.pf[c.cur], c.pg[c.cur], c.ambient,
                       c.surface[c.scur], c.deep[c.scur], c.stat, c.links, with_output ? c.out_dev : nullptr);
}

static Context* bind(long long key, const AeroLbmMesoscaleConfig* config) {
    auto it = ms().contexts.find(key);
    if (it == ms().contexts.end()) {
        set_error("mesoscale: unknown context key");
        return nullptr;
    }
    Context& c = *it->second;
    const bool shape = c.cfg.nx != config->nx || c.cfg.ny != config->ny || c.cfg.nz != config->nz || !c.pf[0];
    c.cfg = *config;
    if (shape) {
        if (!ensure_device()) return nullptr;
        c.cells = (size_t)config->nx * config->ny * config->nz;
        if (!c.allocate()) return nullptr;
        c.cur = c.scur = 0;
    }
    return &c;
}

}  // namespace meso
}  // namespace aero

using namespace aero;

extern "C" {

AERO_LBM_CAPI_EXPORT void aero_lbm_mesoscale_default_config(AeroLbmMesoscaleConfig* out) {
    // mesoscale.cpp:1397-1420
    if (!out) return;
    *out = AeroLbmMesoscaleConfig{};
    out->abi_version = AERO_LBM_MESOSCALE_ABI_VERSION;
    out->struct_size = sizeof(AeroLbmMesoscaleConfig);
    out->nx = 32;
    out->ny = 4;
    out->nz = 32;
    out->dx_m = 64.0f;
    out->dt_s = 5.0f;
    out->molecular_nu_m2_s = 1.5e-5f;
    out->prandtl_air = 0.71f;
    out->turbulent_prandtl = 0.85f;
    out->reference_density_kg_m3 = 1.225f;
    out->ambient_air_temperature_k = 288.15f;
    out->deep_ground_temperature_k = 289.65f;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_derive_transport(const AeroLbmMesoscaleConfig* config,
                                                             AeroLbmMesoscaleTransport* out) {
    if (!meso::config_valid(config) || !out) return 0;
    *out = meso::derive(*config);
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_create_context(const AeroLbmMesoscaleConfig* config,
                                                           long long* out_key) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!meso::config_valid(config) || !out_key) {
        set_error("mesoscale: invalid configuration");
        return 0;
    }
    auto c = std::make_unique<meso::Context>();
    c->cfg = *config;
    c->cfg.nx = 0;  // storage is allocated on the first step (bind)
    const long long key = meso::ms().next_key++;
    meso::ms().contexts.emplace(key, std::move(c));
    *out_key = key;
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_step_context(long long key, const AeroLbmMesoscaleConfig* config,
                                                         const float* forcing, float* out_state) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!meso::config_valid(config) || !forcing || !out_state) {
        set_error("mesoscale: invalid step arguments");
        return 0;
    }
    meso::Context* c = meso::bind(key, config);
    if (!c) return 0;
    meso::Params P = meso::make_params(*c);
    if (!meso::prepare(*c, P, forcing)) return 0;
    meso::launch_step(*c, P, true);
    hipStream_t st = rt().stream;
    return (hip_ok(hipGetLastError(), "mesoscale step") &&
            hip_ok(hipMemcpyAsync(out_state, c->out_dev, c->cells * meso::kState * 4, hipMemcpyDeviceToHost, st),
                   "mesoscale readback") &&
            hip_ok(hipStreamSynchronize(st), "sync"))
               ? 1
               : 0;
}

AERO_LBM_CAPI_EXPORT void aero_lbm_mesoscale_release_context(long long key) {
    std::lock_guard<std::mutex> lk(rt().mu);
    auto& m = meso::ms().contexts;
    auto it = m.find(key);
    if (it == m.end()) return;
    if (rt().stream) (void)hipStreamSynchronize(rt().stream);
    m.erase(it);
}

// extension: lower clamp of the two relaxation times (the reference fixes 0.58);
// NaN restores 0.58
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_set_tau_floor(float tau_floor) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (std::isnan(tau_floor)) {
        meso::ms().tau_floor = 0.58f;
        return 1;
    }
    if (!(tau_floor > 0.5f && tau_floor <= 2.5f)) {
        set_error("mesoscale: tau floor must lie in (0.5, 2.5]");
        return 0;
    }
    meso::ms().tau_floor = tau_floor;
    return 1;
}

// extension: `steps` steps with the same forcing on the device, no per-step
// readback; elapsed device time (HIP events on the library stream) in *out_ms
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_run_steps_timed(long long key, const AeroLbmMesoscaleConfig* config,
                                                            const float* forcing, int steps, float* out_ms) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!meso::config_valid(config) || !forcing || steps <= 0 || !out_ms) {
        set_error("mesoscale: invalid timed-run arguments");
        return 0;
    }
    meso::Context* c = meso::bind(key, config);
    if (!c) return 0;
    meso::Params P = meso::make_params(*c);
    if (!meso::prepare(*c, P, forcing)) return 0;
    hipStream_t st = rt().stream;
    hipEvent_t e0, e1;
    if (!hip_ok(hipEventCreate(&e0), "hipEventCreate")) return 0;
    if (!hip_ok(hipEventCreate(&e1), "hipEventCreate")) {
        (void)hipEventDestroy(e0);
        return 0;
    }
    bool ok = hip_ok(hipEventRecord(e0, st), "hipEventRecord");
    for (int s = 0; ok && s < steps; ++s) {
        meso::launch_step(*c, P, false);
        ok = hip_ok(hipGetLastError(), "mesoscale step");
    }
    ok = ok && hip_ok(hipEventRecord(e1, st), "hipEventRecord") && hip_ok(hipEventSynchronize(e1), "sync");
    float ms = 0.0f;
    if (ok) ok = hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *out_ms = ms;
    return ok ? 1 : 0;
}

// extension: the 6-channel state of the last step without stepping
AERO_LBM_CAPI_EXPORT int aero_lbm_mesoscale_get_state(long long key, float* out_state) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    auto& m = meso::ms().contexts;
    auto it = m.find(key);
    if (it == m.end() || !out_state || !it->second->initialized) {
        set_error("mesoscale: no stepped context for this key");
        return 0;
    }
    meso::Context& c = *it->second;
    hipStream_t st = rt().stream;
    const meso::Params P = meso::make_params(c);
    hipLaunchKernelGGL(meso::macro_kernel, meso::grid1(c.cells), dim3(256), 0, st, P, c.pf[c.cur], c.pg[c.cur],
                       c.ambient, c.surface[c.scur], c.deep[c.scur], c.stat, c.links, c.out_dev);
    return (hip_ok(hipGetLastError(), "mesoscale state") &&
            hip_ok(hipMemcpyAsync(out_state, c.out_dev, c.cells * meso::kState * 4, hipMemcpyDeviceToHost, st),
                   "mesoscale readback") &&
            hip_ok(hipStreamSynchronize(st), "sync"))
               ? 1
               : 0;
}

}  // extern "C"
