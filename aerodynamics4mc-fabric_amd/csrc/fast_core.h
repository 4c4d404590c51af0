// Device-side building blocks of the fast D3Q27 kernels: storage traits (FP32,
// FP16 and BF16 with the equilibrium shift g = f - w), neighbour offsets with
// periodic wrap, the reference's face rules, the temperature update and the
// factorised cumulant collision. Included by d3q27_fast.hip only.
#pragma once

#include "fast_kernels.h"

namespace aero {
namespace fast {

// ---------------------------------------------------------------- buffer resources
// The step kernel reaches populations through one buffer resource over the
// whole population array. The byte offset of an access is split into a
// wave-uniform scalar part (population plane + neighbour row) and a per-lane
// vector part (z inside the row), and the hardware drops an access whose vector
// offset is beyond the resource's range (a load returns 0, a store writes
// nothing). The array must stay below 2 GiB so that kDropByte is out of range.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle
constexpr int kDropByte = INT32_MIN;    // vector byte offset outside every resource

__device__ __forceinline__ Rsrc make_rsrc(const void* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes, kRsrcWord3);
}

// cache policy bits of a buffer access: 0 = default, 2 = non-temporal (streaming)
constexpr int kPolDefault = 0;
constexpr int kPolStream = 2;

// ---------------------------------------------------------------- storage traits
// Every access names the population slot q it touches. Only the 16-bit fixed-point
// format uses it (its step depends on the direction class of q, and a slot and
// its opposite belong to the same class, so a value may sit in either).
struct StF32 {
    using T = float;
    static constexpr bool kShifted = false;
    static __device__ __forceinline__ float ld(const T* p, int) { return __builtin_nontemporal_load(p); }
    static __device__ __forceinline__ void st(T* p, float v, int) { __builtin_nontemporal_store(v, p); }
    template <int Pol>
    static __device__ __forceinline__ float bld(Rsrc r, int off, int soff, int) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, soff, Pol));
    }
    template <int Pol>
    static __device__ __forceinline__ void bst(Rsrc r, int off, int soff, float v, int) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, soff, Pol);
    }
};
struct StF16 {
    using T = uint16_t;
    static constexpr bool kShifted = true;
    static __device__ __forceinline__ float ld(const T* p, int) {
        const uint16_t b = __builtin_nontemporal_load(p);
        return (float)__builtin_bit_cast(_Float16, b);
    }
    static __device__ __forceinline__ void st(T* p, float v, int) {
        __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, (_Float16)v), p);
    }
    template <int Pol>
    static __device__ __forceinline__ float bld(Rsrc r, int off, int soff, int) {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, off, soff, Pol);
        return (float)__builtin_bit_cast(_Float16, b);
    }
    template <int Pol>
    static __device__ __forceinline__ void bst(Rsrc r, int off, int soff, float v, int) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)v), r, off, soff, Pol);
    }
};
struct StBF16 {
    using T = uint16_t;
    static constexpr bool kShifted = true;
    static __device__ __forceinline__ uint16_t pack(float v) {
        uint32_t u = __float_as_uint(v);
        u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
        return (uint16_t)(u >> 16);
    }
    static __device__ __forceinline__ float ld(const T* p, int) {
        const uint32_t b = __builtin_nontemporal_load(p);
        return __uint_as_float(b << 16);
    }
    static __device__ __forceinline__ void st(T* p, float v, int) { __builtin_nontemporal_store(pack(v), p); }
    template <int Pol>
    static __device__ __forceinline__ float bld(Rsrc r, int off, int soff, int) {
        const uint32_t b = __builtin_amdgcn_raw_buffer_load_b16(r, off, soff, Pol);
        return __uint_as_float(b << 16);
    }
    template <int Pol>
    static __device__ __forceinline__ void bst(Rsrc r, int off, int soff, float v, int) {
        __builtin_amdgcn_raw_buffer_store_b16(pack(v), r, off, soff, Pol);
    }
};
// 16-bit signed fixed point of the shifted value g = f - w_q: g = k * R_c * w_q / 32767
// with k in [-32767, 32767] and a range R_c per direction class c = |c_q|^2
// (s16_range). The step is uniform inside the range, so a small deviation from
// w_q keeps an absolute error of R_c w_q / 65534 instead of FP16's relative
// 2^-11; values beyond the range saturate. Encoding is one v_cvt_pknorm_i16_f32
// (round to nearest of clamp(x, -1, 1) * 32767) of g / (R_c w_q).
__device__ constexpr float s16_range(int q) { return s16_class_range(lat_c2(q)); }
__device__ constexpr float s16_w(int q) {
    return lat_c2(q) == 0 ? 8.0f / 27.0f : (lat_c2(q) == 1 ? 2.0f / 27.0f : (lat_c2(q) == 2 ? 1.0f / 54.0f : 1.0f / 216.0f));
}
struct StS16 {
    using T = uint16_t;
    static constexpr bool kShifted = true;
    static __device__ __forceinline__ float step(int q) { return s16_range(q) * s16_w(q) / 32767.0f; }
    static __device__ __forceinline__ float inv_range(int q) { return 1.0f / (s16_range(q) * s16_w(q)); }
    static __device__ __forceinline__ uint16_t pack(float v, int q) {
        const auto k = __builtin_amdgcn_cvt_pknorm_i16(v * inv_range(q), 0.0f);
        return (uint16_t)k.x;
    }
    static __device__ __forceinline__ float unpack(uint32_t b, int q) {
        return (float)(int)(int16_t)(uint16_t)b * step(q);
    }
    static __device__ __forceinline__ float ld(const T* p, int q) { return unpack(__builtin_nontemporal_load(p), q); }
    static __device__ __forceinline__ void st(T* p, float v, int q) { __builtin_nontemporal_store(pack(v, q), p); }
    template <int Pol>
    static __device__ __forceinline__ float bld(Rsrc r, int off, int soff, int q) {
        return unpack(__builtin_amdgcn_raw_buffer_load_b16(r, off, soff, Pol), q);
    }
    template <int Pol>
    static __device__ __forceinline__ void bst(Rsrc r, int off, int soff, float v, int q) {
        __builtin_amdgcn_raw_buffer_store_b16(pack(v, q), r, off, soff, Pol);
    }
};

// velocity carried to the next step's temperature advection; bld / bst address
// it as a buffer (byte offsets: per-lane off, wave-uniform soff)
struct UF32 {
    using T = float;
    static __device__ __forceinline__ float ld(const T* p) { return *p; }
    static __device__ __forceinline__ void st(T* p, float v) { *p = v; }
    static __device__ __forceinline__ float bld(Rsrc r, int off, int soff) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
    }
    static __device__ __forceinline__ void bst(Rsrc r, int off, int soff, float v) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, soff, 0);
    }
};
struct UF16 {
    using T = uint16_t;
    static __device__ __forceinline__ float ld(const T* p) { return (float)__builtin_bit_cast(_Float16, *p); }
    static __device__ __forceinline__ void st(T* p, float v) { *p = __builtin_bit_cast(uint16_t, (_Float16)v); }
    static __device__ __forceinline__ float bld(Rsrc r, int off, int soff) {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, off, soff, 0);
        return (float)__builtin_bit_cast(_Float16, b);
    }
    static __device__ __forceinline__ void bst(Rsrc r, int off, int soff, float v) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)v), r, off, soff, 0);
    }
};

// one float of a buffer at byte offset off
__device__ __forceinline__ float bld_f32(Rsrc r, int off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void bst_f32(Rsrc r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, 0);
}
// all ones where bit b of v is set, else zero
__device__ __forceinline__ int bit_mask(uint32_t v, int b) { return (int)(v << (31 - b)) >> 31; }

// Per-lane pick of b where bit `bit` of v is set, else a: one bitfield extract to a
// 0 / all-ones mask and one three-input bitwise operation. The extract is opaque to
// the compiler, which would otherwise rewrite the pair as a compare, a select and a
// wait on VCC. `bit` must be a constant once the caller's loop is unrolled.
__device__ __forceinline__ int pick_by_bit(uint32_t v, int bit, int a, int b) {
    int m;
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"(v), "i"(bit));
    return (m & b) | (~m & a);
}

__device__ constexpr float wq(int q) {
    return (lat_cx(q) * lat_cx(q) + lat_cy(q) * lat_cy(q) + lat_cz(q) * lat_cz(q)) == 0   ? 8.0f / 27.0f
           : (lat_cx(q) * lat_cx(q) + lat_cy(q) * lat_cy(q) + lat_cz(q) * lat_cz(q)) == 1 ? 2.0f / 27.0f
           : (lat_cx(q) * lat_cx(q) + lat_cy(q) * lat_cy(q) + lat_cz(q) * lat_cz(q)) == 2 ? 1.0f / 54.0f
                                                                                         : 1.0f / 216.0f;
}

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(hi, fmaxf(lo, v)); }
__device__ __forceinline__ bool finite(float v) { return __builtin_isfinite(v); }

// equilibrium (ref: native/src/aero_lbm_hydro_core.cpp:364-368); `shifted` returns feq - w
template <bool Shifted>
__device__ __forceinline__ float feq_s(int q, float rho, float ux, float uy, float uz) {
    const float cu = 3.0f * ((float)lat_cx(q) * ux + (float)lat_cy(q) * uy + (float)lat_cz(q) * uz);
    const float uu = ux * ux + uy * uy + uz * uz;
    const float poly = cu + 0.5f * cu * cu - 1.5f * uu;
    if (Shifted) return wq(q) * ((rho - 1.0f) + rho * poly);
    return wq(q) * rho * (1.0f + poly);
}

// ---------------------------------------------------------------- neighbour offsets
// Offsets (in cells) to the neighbour at +-1 along each axis with periodic wrap;
// an out-of-domain neighbour on a non-periodic face maps to the cell itself so
// every computed index stays in bounds (those links are flagged special).
struct Nbr {
    int xm, xp, ym, yp, zm, zp;
    __device__ __forceinline__ int ox(int d) const { return d < 0 ? xm : (d > 0 ? xp : 0); }
    __device__ __forceinline__ int oy(int d) const { return d < 0 ? ym : (d > 0 ? yp : 0); }
    __device__ __forceinline__ int oz(int d) const { return d < 0 ? zm : (d > 0 ? zp : 0); }
    __device__ __forceinline__ int off(int dx, int dy, int dz) const { return ox(dx) + oy(dy) + oz(dz); }
};

__device__ __forceinline__ Nbr make_nbr(const StepParams& P, int x, int y, int z) {
    const int yz = P.ny * P.nz;
    Nbr n;
    n.xm = x > 0 ? -yz : (P.hydro_wrap_lo[0] ? (P.nx - 1) * yz : 0);
    n.xp = x < P.nx - 1 ? yz : (P.hydro_wrap_hi[0] ? -(P.nx - 1) * yz : 0);
    n.ym = y > 0 ? -P.nz : (P.hydro_wrap_lo[1] ? (P.ny - 1) * P.nz : 0);
    n.yp = y < P.ny - 1 ? P.nz : (P.hydro_wrap_hi[1] ? -(P.ny - 1) * P.nz : 0);
    n.zm = z > 0 ? -1 : (P.hydro_wrap_lo[2] ? P.nz - 1 : 0);
    n.zp = z < P.nz - 1 ? 1 : (P.hydro_wrap_hi[2] ? -(P.nz - 1) : 0);
    return n;
}

// ---------------------------------------------------------------- face rules
__device__ __forceinline__ int wrapc(int c, int n, int lo, int hi) {
    if (c < 0 && lo) return c + n;
    if (c >= n && hi) return c - n;
    return c;
}
__device__ __forceinline__ int icl(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// first out-of-domain axis in the reference's order x-, x+, y-, y+, z-, z+ (jni.cpp:743-751)
__device__ __forceinline__ int oob_face(const StepParams& P, int sx, int sy, int sz) {
    if (sx < 0) return 0;
    if (sx >= P.nx) return 1;
    if (sy < 0) return 2;
    if (sy >= P.ny) return 3;
    if (sz < 0) return 4;
    return 5;
}

// A face's rule kind and face-store offset for a per-lane face index, picked from
// the parameter block's own copy of the six faces (wave-uniform values) instead of
// an indexed load from the device face table: a rare cell's face rule then waits
// on no memory load before it can load its inputs.
__device__ __forceinline__ int face_kind(const StepParams& P, int face) {
    int k = P.face[5].hydro_kind;
    k = face == 4 ? P.face[4].hydro_kind : k;
    k = face == 3 ? P.face[3].hydro_kind : k;
    k = face == 2 ? P.face[2].hydro_kind : k;
    k = face == 1 ? P.face[1].hydro_kind : k;
    return face == 0 ? P.face[0].hydro_kind : k;
}
__device__ __forceinline__ long long face_store_off(const StepParams& P, int face) {
    long long o = P.face[5].store_off;
    o = face == 4 ? P.face[4].store_off : o;
    o = face == 3 ? P.face[3].store_off : o;
    o = face == 2 ? P.face[2].store_off : o;
    o = face == 1 ? P.face[1].store_off : o;
    return face == 0 ? P.face[0].store_off : o;
}

// Post-collision value (stored form) of direction q at in-domain cell (sx,sy,sz)
// as the previous step left it: direct A-B read, or the A-A face store of `face`.
template <class St, bool AA>
struct PrevPost {
    const typename St::T* f;  // A-B: previous post-collision buffer
    const float* store;       // A-A: face store of the previous step
    size_t N;
    __device__ __forceinline__ float get(const StepParams& P, int face, int q, int sx, int sy, int sz) const {
        if (!AA) return St::ld(f + (size_t)q * N + (size_t)((sx * P.ny + sy) * P.nz + sz), q);
        const int axis = face >> 1;
        const int layer = (face & 1) ? ((axis == 0 ? P.nx : axis == 1 ? P.ny : P.nz) - 1 -
                                        (axis == 0 ? sx : axis == 1 ? sy : sz))
                                     : (axis == 0 ? sx : axis == 1 ? sy : sz);
        const int ix = q / 9, iy = (q / 3) % 3, iz = q % 3;
        const int j = axis == 0 ? iy * 3 + iz : axis == 1 ? ix * 3 + iz : ix * 3 + iy;
        const int A = axis == 0 ? P.ny : P.nx;
        const int B = axis == 2 ? P.ny : P.nz;
        const int a = axis == 0 ? sy : sx;
        const int b = axis == 2 ? sy : sz;
        // faces without a store (and layers beyond the third) load nothing: the caller
        // computes every rule's candidates and selects afterwards
        const bool ok = ((P.face_store_faces >> face) & 1) && layer >= 0 && layer <= 2;
        return ok ? store[face_store_off(P, face) + ((size_t)(layer * 9 + j) * A + a) * B + b] : 0.0f;
    }
};

// Incoming value for an out-of-domain link q at fluid cell (x,y,z); `bounce` is the
// cell's own post-collision value of opp(q) (the reference's halfway bounce-back).
// The face's rule is picked first and only its inputs are loaded: the mirrored
// source for symmetry, the two layers along the normal for convective outflow,
// nothing for bounce-back and the equilibrium faces.
template <class St, bool AA>
__device__ __forceinline__ float face_link(const StepParams& P, const uint8_t* flags, float ref_p_cell,
                                           const PrevPost<St, AA>& prev, int cell, int x, int y, int z, int q,
                                           float bounce) {
    const int cx = lat_cx(q), cy = lat_cy(q), cz = lat_cz(q);
    const int sx = wrapc(x - cx, P.nx, P.hydro_wrap_lo[0], P.hydro_wrap_hi[0]);
    const int sy = wrapc(y - cy, P.ny, P.hydro_wrap_lo[1], P.hydro_wrap_hi[1]);
    const int sz = wrapc(z - cz, P.nz, P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]);
    const int face = oob_face(P, sx, sy, sz);
    const FaceParams& fc = P.faces_dev[face];  // velocity and pressure of the equilibrium faces
    int kind = face_kind(P, face);
    if (kind == 0 || kind == 1) kind = P.convective_disabled && kind == 0 ? 2 : 6;
    if (kind == 2) return bounce;
    if (kind == 7) {  // symmetry: the mirrored in-domain source (ref: jni.cpp:1582-1587)
        const int ix = icl(sx, 0, P.nx - 1), iy = icl(sy, 0, P.ny - 1), iz = icl(sz, 0, P.nz - 1);
        const bool s0 = flags[(ix * P.ny + iy) * P.nz + iz] & kCellSolid;
        const float p0 = prev.get(P, face, q, ix, iy, iz);
        return s0 ? bounce : p0;
    }
    // convective layers along the face normal (jni.cpp:1494-1516)
    const int dx = (sx < 0 || sx >= P.nx) ? cx : 0;
    const int dy = (sy < 0 || sy >= P.ny) ? cy : 0;
    const int dz = (sz < 0 || sz >= P.nz) ? cz : 0;
    const int x1 = icl(x + dx, 0, P.nx - 1), y1 = icl(y + dy, 0, P.ny - 1), z1 = icl(z + dz, 0, P.nz - 1);
    const int x2 = icl(x + 2 * dx, 0, P.nx - 1), y2 = icl(y + 2 * dy, 0, P.ny - 1), z2 = icl(z + 2 * dz, 0, P.nz - 1);
    const bool conv_rule = !(kind >= 3 && kind <= 5);  // equilibrium faces load nothing
    const bool s1 = conv_rule && (flags[(x1 * P.ny + y1) * P.nz + z1] & kCellSolid);
    const bool s2 = conv_rule && (flags[(x2 * P.ny + y2) * P.nz + z2] & kCellSolid);
    const float p1 = conv_rule ? prev.get(P, face, q, x1, y1, z1) : 0.0f;
    const float p2 = conv_rule ? prev.get(P, face, q, x2, y2, z2) : 0.0f;
    // velocity / pressure / inlet equilibrium (ref: jni.cpp:1518-1561)
    float rho = 1.0f + clampf(ref_p_cell, -0.03f, 0.03f);
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (kind != 5) {
        ux = fc.vel[0];
        uy = fc.vel[1];
        uz = fc.vel[2];
    } else {
        rho = 1.0f + clampf(fc.pressure, -0.03f, 0.03f);
    }
    rho = clampf(rho, 0.97f, 1.03f);
    const float s2v = ux * ux + uy * uy + uz * uz;
    if (s2v > P.kmax_speed * P.kmax_speed && s2v > 0.0f) {
        const float sc = P.kmax_speed / sqrtf(s2v);
        ux *= sc;
        uy *= sc;
        uz *= sc;
    }
    const float eq = feq_s<St::kShifted>(q, rho, ux, uy, uz);
    const float conv = s1 ? bounce : (s2 ? p1 : p1 + 0.15f * (p1 - p2));  // w shift cancels: weights sum to one
    return conv_rule ? conv : eq;
}

// The 9 links of an x-face cell pulled from outside the domain (Side 0: x_min,
// links with c_x = +1; Side 1: x_max, c_x = -1), for the bulk kernel. The same
// rules and arithmetic as face_link, with the face's rule (wave-uniform) chosen
// first so only that rule's values are loaded: equilibrium faces load nothing,
// convective and symmetry faces load their two (one) previous-step values per
// link. g[q] holds the bounce-back candidate on entry.
template <int Side, class St, bool AA>
__device__ __forceinline__ void xface_links(const StepParams& P, const uint8_t* __restrict__ flags, float ref_p_cell,
                                            const PrevPost<St, AA>& prev, int x, int y, int z, float g[27]) {
    constexpr int q0 = Side == 0 ? 18 : 0;
    constexpr int cx = Side == 0 ? 1 : -1;
    const FaceParams& fc = P.faces_dev[Side];
    int kind = fc.hydro_kind;
    if (kind == 0 || kind == 1) kind = P.convective_disabled && kind == 0 ? 2 : 6;
    if (kind == 2) return;
    if (kind >= 3 && kind <= 5) {
        float rho = 1.0f + clampf(ref_p_cell, -0.03f, 0.03f);
        float ux = 0.0f, uy = 0.0f, uz = 0.0f;
        if (kind != 5) {
            ux = fc.vel[0];
            uy = fc.vel[1];
            uz = fc.vel[2];
        } else {
            rho = 1.0f + clampf(fc.pressure, -0.03f, 0.03f);
        }
        rho = clampf(rho, 0.97f, 1.03f);
        const float s2v = ux * ux + uy * uy + uz * uz;
        if (s2v > P.kmax_speed * P.kmax_speed && s2v > 0.0f) {
            const float sc = P.kmax_speed / sqrtf(s2v);
            ux *= sc;
            uy *= sc;
            uz *= sc;
        }
#pragma unroll
        for (int j = 0; j < 9; ++j) g[q0 + j] = feq_s<St::kShifted>(q0 + j, rho, ux, uy, uz);
        return;
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int q = q0 + j;
        const int cy = lat_cy(q), cz = lat_cz(q);
        const int sy = wrapc(y - cy, P.ny, P.hydro_wrap_lo[1], P.hydro_wrap_hi[1]);
        const int sz = wrapc(z - cz, P.nz, P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]);
        if (kind == 7) {
            const int iy = icl(sy, 0, P.ny - 1), iz = icl(sz, 0, P.nz - 1);
            const bool s0 = flags[(x * P.ny + iy) * P.nz + iz] & kCellSolid;
            const float p0 = prev.get(P, Side, q, x, iy, iz);
            g[q] = s0 ? g[q] : p0;
        } else {
            const int dy = (sy < 0 || sy >= P.ny) ? cy : 0;
            const int dz = (sz < 0 || sz >= P.nz) ? cz : 0;
            const int x1 = icl(x + cx, 0, P.nx - 1), y1 = icl(y + dy, 0, P.ny - 1), z1 = icl(z + dz, 0, P.nz - 1);
            const int x2 = icl(x + 2 * cx, 0, P.nx - 1), y2 = icl(y + 2 * dy, 0, P.ny - 1),
                      z2 = icl(z + 2 * dz, 0, P.nz - 1);
            const bool s1 = flags[(x1 * P.ny + y1) * P.nz + z1] & kCellSolid;
            const bool s2 = flags[(x2 * P.ny + y2) * P.nz + z2] & kCellSolid;
            const float p1 = prev.get(P, Side, q, x1, y1, z1);
            const float p2 = prev.get(P, Side, q, x2, y2, z2);
            g[q] = s1 ? g[q] : (s2 ? p1 : p1 + 0.15f * (p1 - p2));
        }
    }
}

// The 9 links of a cell in the y_min (Side 0, c_y = +1) or y_max (Side 1, c_y = -1)
// row of a symmetry face flagged yface_bulk, for the bulk kernel: the rule of
// face_link for kind 7 (the mirrored in-domain source of the previous step, or
// bounce-back from a solid one). Links whose source lies outside in x take the x
// face's rule (the reference's face order x, y, z) and are left alone; x is
// wave-uniform, so that test is a wave-uniform branch. g[q] holds the bounce-back
// candidate on entry.
template <int Side, class St, bool AA>
__device__ __forceinline__ void yface_links(const StepParams& P, const uint8_t* __restrict__ flags,
                                            const PrevPost<St, AA>& prev, int x, int y, int z, float g[27]) {
    constexpr int iy = Side == 0 ? 2 : 0;
    const int ys = Side == 0 ? 0 : P.ny - 1;  // the mirrored source row
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int q = (j / 3) * 9 + iy * 3 + j % 3;
        const int sx = wrapc(x - lat_cx(q), P.nx, P.hydro_wrap_lo[0], P.hydro_wrap_hi[0]);
        if (sx < 0 || sx >= P.nx) continue;
        const int iz = icl(wrapc(z - lat_cz(q), P.nz, P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]), 0, P.nz - 1);
        const bool s0 = flags[(sx * P.ny + ys) * P.nz + iz] & kCellSolid;
        const float p0 = prev.get(P, 2 + Side, q, sx, ys, iz);
        g[q] = s0 ? g[q] : p0;
    }
}

__device__ __forceinline__ bool link_oob(const StepParams& P, int x, int y, int z, int q) {
    const int sx = wrapc(x - lat_cx(q), P.nx, P.hydro_wrap_lo[0], P.hydro_wrap_hi[0]);
    const int sy = wrapc(y - lat_cy(q), P.ny, P.hydro_wrap_lo[1], P.hydro_wrap_hi[1]);
    const int sz = wrapc(z - lat_cz(q), P.nz, P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]);
    return sx < 0 || sy < 0 || sz < 0 || sx >= P.nx || sy >= P.ny || sz >= P.nz;
}

// ---------------------------------------------------------------- temperature
// temperature at (x, y, z) for the stencil of a cell whose own value is `self`:
// periodic wrap, Dirichlet / Neumann thermal faces, solid or adiabatic -> self.
// Every load is issued unconditionally at a clamped in-domain index, so the 14
// stencil points of a cell do not form a chain of dependent loads. Velocities are
// clamped below one lattice unit, so a point lies at most one cell outside the
// domain and one period brings a wrapped point back in.
__device__ __forceinline__ float t_at(const StepParams& P, const float* T, const uint8_t* flags, int x, int y, int z,
                                      float self) {
    if ((x < 0 && P.thermal_wrap_lo[0]) || (x >= P.nx && P.thermal_wrap_hi[0])) x = x < 0 ? x + P.nx : x - P.nx;
    if ((y < 0 && P.thermal_wrap_lo[1]) || (y >= P.ny && P.thermal_wrap_hi[1])) y = y < 0 ? y + P.ny : y - P.ny;
    if ((z < 0 && P.thermal_wrap_lo[2]) || (z >= P.nz && P.thermal_wrap_hi[2])) z = z < 0 ? z + P.nz : z - P.nz;
    const bool inside = x >= 0 && y >= 0 && z >= 0 && x < P.nx && y < P.ny && z < P.nz;
    const int c = (icl(x, 0, P.nx - 1) * P.ny + icl(y, 0, P.ny - 1)) * P.nz + icl(z, 0, P.nz - 1);
    const bool solid = flags[c] & kCellSolid;
    const float tv = T[c];
    const FaceParams& fc = P.faces_dev[oob_face(P, x, y, z)];
    const int kind = fc.thermal_kind;
    const float face_t = kind == 4 ? clampf(fc.temperature, -1.0f, 1.0f)
                                   : (kind == 5 ? clampf(self + fc.heat_flux, -1.0f, 1.0f) : self);
    return inside ? (solid ? self : tv) : face_t;
}

// first-order semi-Lagrangian advection + explicit diffusion + source
// (ref: native/src/aero_lbm_jni.cpp:2145-2188)
__device__ __forceinline__ float temperature_step(const StepParams& P, const float* __restrict__ T,
                                                  const uint8_t* __restrict__ flags, const float* __restrict__ src,
                                                  int cell, int x, int y, int z, const Nbr& nb, bool interior,
                                                  float ux, float uy, float uz) {
    const float tc = T[cell];
    const float px = (float)x - ux, py = (float)y - uy, pz = (float)z - uz;
    const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
    const float fx = px - fx0, fy = py - fy0, fz = pz - fz0;
    float c000, c100, c010, c110, c001, c101, c011, c111, tsum;
    if (interior) {
        // all 26 neighbours are fluid and in the domain: plain indexed reads
        const int ax = (int)fx0 < x ? nb.xm : 0, bx = (int)fx0 < x ? 0 : nb.xp;
        const int ay = (int)fy0 < y ? nb.ym : 0, by = (int)fy0 < y ? 0 : nb.yp;
        const int az = (int)fz0 < z ? nb.zm : 0, bz = (int)fz0 < z ? 0 : nb.zp;
        c000 = T[cell + ax + ay + az];
        c100 = T[cell + bx + ay + az];
        c010 = T[cell + ax + by + az];
        c110 = T[cell + bx + by + az];
        c001 = T[cell + ax + ay + bz];
        c101 = T[cell + bx + ay + bz];
        c011 = T[cell + ax + by + bz];
        c111 = T[cell + bx + by + bz];
        tsum = T[cell + nb.xp] + T[cell + nb.xm] + T[cell + nb.yp] + T[cell + nb.ym] + T[cell + nb.zp] +
               T[cell + nb.zm];
    } else {
        const int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
        c000 = t_at(P, T, flags, x0, y0, z0, tc);
        c100 = t_at(P, T, flags, x0 + 1, y0, z0, tc);
        c010 = t_at(P, T, flags, x0, y0 + 1, z0, tc);
        c110 = t_at(P, T, flags, x0 + 1, y0 + 1, z0, tc);
        c001 = t_at(P, T, flags, x0, y0, z0 + 1, tc);
        c101 = t_at(P, T, flags, x0 + 1, y0, z0 + 1, tc);
        c011 = t_at(P, T, flags, x0, y0 + 1, z0 + 1, tc);
        c111 = t_at(P, T, flags, x0 + 1, y0 + 1, z0 + 1, tc);
        tsum = t_at(P, T, flags, x + 1, y, z, tc) + t_at(P, T, flags, x - 1, y, z, tc) +
               t_at(P, T, flags, x, y + 1, z, tc) + t_at(P, T, flags, x, y - 1, z, tc) +
               t_at(P, T, flags, x, y, z + 1, tc) + t_at(P, T, flags, x, y, z - 1, tc);
    }
    const float c00 = c000 + (c100 - c000) * fx;
    const float c10 = c010 + (c110 - c010) * fx;
    const float c01 = c001 + (c101 - c001) * fx;
    const float c11 = c011 + (c111 - c011) * fx;
    const float c0 = c00 + (c10 - c00) * fy;
    const float c1 = c01 + (c11 - c01) * fy;
    const float adv = c0 + (c1 - c0) * fz;
    const float lap = tsum - 6.0f * tc;
    const float source = (P.source_on && P.in_ch > 9) ? clampf(src[cell], -0.006f, 0.006f) : 0.0f;
    return clampf(adv + (P.alpha * lap + source), -1.0f, 1.0f);
}

// ---------------------------------------------------------------- collision
#define IDX(a, b, c) (((a) * 3 + (b)) * 3 + (c))

// 1-D lattice moments of the weights: sum_c w(c) c^n = 1, 0, 1/3, 0, 1/3 for n = 0..4
__device__ constexpr float mu1(int n) { return n == 0 ? 1.0f : ((n & 1) ? 0.0f : 1.0f / 3.0f); }
__device__ constexpr float emom(int a, int b, int c) { return mu1(a) * mu1(b) * mu1(c); }

// 2^-(number of non-zero indices): the scale of moment (a, b, c) in the halved form
__device__ constexpr float half_scale(int a, int b, int c) {
    return (a ? 0.5f : 1.0f) * (b ? 0.5f : 1.0f) * (c ? 0.5f : 1.0f);
}

// raw moment (a,b,c) of the Guo source S_q = w_q [3 (c-u).F + 9 (c.u)(c.F)] / P for a
// force fa along axis Ax: uFa = u_Ax fa, ua_fa[al] = u_al fa (terms of the other
// force components vanish and are left out)
template <int A, int B, int C, int Ax>
__device__ __forceinline__ float guo_moment(float fa, float uFa, const float ua_fa[3]) {
    float v = 0.0f;
    {
        const float k = 3.0f * emom(A + (Ax == 0), B + (Ax == 1), C + (Ax == 2));
        if (k != 0.0f) v = fmaf(k, fa, v);
    }
    {
        const float k = -3.0f * emom(A, B, C);
        if (k != 0.0f) v = fmaf(k, uFa, v);
    }
#pragma unroll
    for (int al = 0; al < 3; ++al) {
        const float k = 9.0f * emom(A + (al == 0) + (Ax == 0), B + (al == 1) + (Ax == 1), C + (al == 2) + (Ax == 2));
        if (k != 0.0f) v = fmaf(k, ua_fa[al], v);
    }
    return v;
}

// adds the Guo source moments (in the halved form) for a force fa along axis Ax
template <int Ax>
__device__ __forceinline__ void add_guo(float g[27], float fa, float ux, float uy, float uz) {
    const float u3[3] = {ux, uy, uz};
    const float uFa = u3[Ax] * fa;
    const float ua_fa[3] = {ux * fa, uy * fa, uz * fa};
#define GM(a, b, c) g[IDX(a, b, c)] += half_scale(a, b, c) * guo_moment<a, b, c, Ax>(fa, uFa, ua_fa);
    GM(0, 0, 0) GM(0, 0, 1) GM(0, 0, 2) GM(0, 1, 0) GM(0, 1, 1) GM(0, 1, 2) GM(0, 2, 0) GM(0, 2, 1) GM(0, 2, 2)
    GM(1, 0, 0) GM(1, 0, 1) GM(1, 0, 2) GM(1, 1, 0) GM(1, 1, 1) GM(1, 1, 2) GM(1, 2, 0) GM(1, 2, 1) GM(1, 2, 2)
    GM(2, 0, 0) GM(2, 0, 1) GM(2, 0, 2) GM(2, 1, 0) GM(2, 1, 1) GM(2, 1, 2) GM(2, 2, 0) GM(2, 2, 1) GM(2, 2, 2)
#undef GM
}

// The Guo source S_q = w_q [3 (c_q - u).F + 9 (c_q.u)(c_q.F)] for a force (0, 0, fa)
// has 12 non-zero raw moments (weight moments E(0) = 1, E(1) = E(3) = 0,
// E(2) = E(4) = 1/3): (a,b,1) = e_a e_b fa, (1,b,1) = e_b u_x fa, (a,1,1) = e_a u_y fa,
// (a,b,2) = 2 e_a e_b u_z fa for a, b in {0, 2} (e_0 = 1, e_2 = 1/3). Added in the
// halved form (scale 2^-(non-zero indices)).
__device__ __forceinline__ void add_guo_z(float g[27], float fa, float ux, float uy, float uz) {
    const float xf = ux * fa, yf = uy * fa, zf = uz * fa;
    g[IDX(0, 0, 1)] = fmaf(0.5f, fa, g[IDX(0, 0, 1)]);
    g[IDX(2, 0, 1)] = fmaf(0.25f / 3.0f, fa, g[IDX(2, 0, 1)]);
    g[IDX(0, 2, 1)] = fmaf(0.25f / 3.0f, fa, g[IDX(0, 2, 1)]);
    g[IDX(2, 2, 1)] = fmaf(0.125f / 9.0f, fa, g[IDX(2, 2, 1)]);
    g[IDX(1, 0, 1)] = fmaf(0.25f, xf, g[IDX(1, 0, 1)]);
    g[IDX(1, 2, 1)] = fmaf(0.125f / 3.0f, xf, g[IDX(1, 2, 1)]);
    g[IDX(0, 1, 1)] = fmaf(0.25f, yf, g[IDX(0, 1, 1)]);
    g[IDX(2, 1, 1)] = fmaf(0.125f / 3.0f, yf, g[IDX(2, 1, 1)]);
    g[IDX(0, 0, 2)] = fmaf(1.0f, zf, g[IDX(0, 0, 2)]);
    g[IDX(2, 0, 2)] = fmaf(0.5f / 3.0f, zf, g[IDX(2, 0, 2)]);
    g[IDX(0, 2, 2)] = fmaf(0.5f / 3.0f, zf, g[IDX(0, 2, 2)]);
    g[IDX(2, 2, 2)] = fmaf(0.25f / 9.0f, zf, g[IDX(2, 2, 2)]);
}

struct CollideResult {
    float ux, uy, uz;
    bool ok;
};

// Boundary-reference shell (ref: native/src/aero_lbm_jni.cpp:4535-4552 for the
// blend, 6449-6522 for the layer and target rules): a fluid cell `layer` cells deep
// (1 = outermost) relaxes its density and velocity toward the stored reference with
// weight clamp(relax * (1 - (layer - 1) / layers)^2, 0, 0.5). The populations move by
// the difference of the two equilibria, so the blended density and velocity are the
// cell's new conserved moments and the other moments keep their post-collision values.
template <bool Shifted>
__device__ __forceinline__ void shell_blend(const StepParams& P, const uint16_t* __restrict__ state, int cell,
                                            int layer, float g[27], CollideResult& r) {
    const uint2 raw = *reinterpret_cast<const uint2*>(state + (size_t)cell * 4);
    const float tux = (float)__builtin_bit_cast(_Float16, (uint16_t)(raw.x & 0xFFFFu));
    const float tuy = (float)__builtin_bit_cast(_Float16, (uint16_t)(raw.x >> 16));
    const float tuz = (float)__builtin_bit_cast(_Float16, (uint16_t)(raw.y & 0xFFFFu));
    const float tp = (float)__builtin_bit_cast(_Float16, (uint16_t)(raw.y >> 16));
    const float shell_t = 1.0f - (float)(layer - 1) / fmaxf(1.0f, (float)P.shell_layers);
    const float alpha = clampf(P.shell_relax * shell_t * shell_t, 0.0f, 0.5f);
    float rho = Shifted ? 1.0f : 0.0f;
#pragma unroll
    for (int q = 0; q < 27; ++q) rho += g[q];
    const float trho = clampf(1.0f + clampf(tp, -0.03f, 0.03f), 0.97f, 1.03f);
    const float nrho = rho + alpha * (trho - rho);
    const float nux = r.ux + alpha * (tux - r.ux);
    const float nuy = r.uy + alpha * (tuy - r.uy);
    const float nuz = r.uz + alpha * (tuz - r.uz);
    // feq(nrho, nu) - feq(rho, u) = w_q (a + c.b + c.D.c): the equilibrium's moments
    // of order 0, 1 and 2 change by a, b and D
    const float a = (nrho - rho) - 1.5f * (nrho * (nux * nux + nuy * nuy + nuz * nuz) -
                                          rho * (r.ux * r.ux + r.uy * r.uy + r.uz * r.uz));
    const float bx = 3.0f * (nrho * nux - rho * r.ux), by = 3.0f * (nrho * nuy - rho * r.uy),
                bz = 3.0f * (nrho * nuz - rho * r.uz);
    const float dxx = 4.5f * (nrho * nux * nux - rho * r.ux * r.ux), dyy = 4.5f * (nrho * nuy * nuy - rho * r.uy * r.uy),
                dzz = 4.5f * (nrho * nuz * nuz - rho * r.uz * r.uz);
    const float dxy = 9.0f * (nrho * nux * nuy - rho * r.ux * r.uy), dxz = 9.0f * (nrho * nux * nuz - rho * r.ux * r.uz),
                dyz = 9.0f * (nrho * nuy * nuz - rho * r.uy * r.uz);
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const float cx = (float)lat_cx(q), cy = (float)lat_cy(q), cz = (float)lat_cz(q);
        const float quad = cx * cx * dxx + cy * cy * dyy + cz * cz * dzz + cx * cy * dxy + cx * cz * dxz + cy * cz * dyz;
        g[q] += wq(q) * (a + cx * bx + cy * by + cz * bz + quad);
    }
    r.ux = nux;
    r.uy = nuy;
    r.uz = nuz;
}

// g: populations in stored form (f - w when Shifted), overwritten with the post-collision values
// Raw moments of order <= 2 from the 27 populations (separable, only the needed
// lines): m000, the three first moments and the six second moments.
struct LowMoments {
    float m000, m100, m010, m001, m200, m020, m002, m110, m101, m011;
};
__device__ __forceinline__ LowMoments low_moments(const float f[27]) {
    // z lines of the 9 (x, y) columns: Z0 = sum, Z1 = c_z-weighted, Z2 = c_z^2-weighted
    float z0[9], z1[9], z2[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float s = f[i * 3 + 0] + f[i * 3 + 2];
        z1[i] = f[i * 3 + 2] - f[i * 3 + 0];
        z0[i] = s + f[i * 3 + 1];
        z2[i] = s;
    }
    // y lines per x: (ky, kz) in {(0,0), (1,0), (2,0), (0,1), (1,1), (0,2)}
    float y00[3], y10[3], y20[3], y01[3], y11[3], y02[3];
#pragma unroll
    for (int ix = 0; ix < 3; ++ix) {
        const int a = ix * 3, b = ix * 3 + 1, c = ix * 3 + 2;
        const float s0 = z0[a] + z0[c];
        y00[ix] = s0 + z0[b];
        y10[ix] = z0[c] - z0[a];
        y20[ix] = s0;
        y01[ix] = (z1[a] + z1[c]) + z1[b];
        y11[ix] = z1[c] - z1[a];
        y02[ix] = (z2[a] + z2[c]) + z2[b];
    }
    LowMoments m;
    const float sx = y00[0] + y00[2];
    m.m000 = sx + y00[1];
    m.m100 = y00[2] - y00[0];
    m.m200 = sx;
    m.m010 = (y10[0] + y10[2]) + y10[1];
    m.m110 = y10[2] - y10[0];
    m.m001 = (y01[0] + y01[2]) + y01[1];
    m.m101 = y01[2] - y01[0];
    m.m020 = (y20[0] + y20[2]) + y20[1];
    m.m011 = (y11[0] + y11[2]) + y11[1];
    m.m002 = (y02[0] + y02[2]) + y02[1];
    return m;
}

// Central moments about u -> raw moments along one axis, with the outputs of
// order 1 and 2 halved (the inputs of order 1 and 2 arrive halved already): the
// factor the next inverse line needs. hu = u / 2, tu = 2 u, hu2 = u^2 / 2.
__device__ __forceinline__ void unshift_half(float& k0, float& k1, float& k2, float hu, float tu, float hu2) {
    k2 = fmaf(tu, k1, fmaf(hu2, k0, k2));
    k1 = fmaf(hu, k0, k1);
}
// raw moments (order 1 and 2 halved) -> populations (c = -1, 0, +1) of one line
__device__ __forceinline__ void inv_half(float& m0, float& m1, float& m2) {
    const float fm = m2 - m1, fp = m2 + m1;
    m0 = fmaf(-2.0f, m2, m0);
    m1 = fm;  // c = -1 value, moved to slot 0 by the caller
    m2 = fp;
}

__device__ __forceinline__ void inverse_half(float m[27]) {
#pragma unroll
    for (int ky = 0; ky < 3; ++ky)
#pragma unroll
        for (int kz = 0; kz < 3; ++kz) {
            inv_half(m[IDX(0, ky, kz)], m[IDX(1, ky, kz)], m[IDX(2, ky, kz)]);
            const float t = m[IDX(0, ky, kz)];
            m[IDX(0, ky, kz)] = m[IDX(1, ky, kz)];
            m[IDX(1, ky, kz)] = t;
        }
#pragma unroll
    for (int ix = 0; ix < 3; ++ix)
#pragma unroll
        for (int kz = 0; kz < 3; ++kz) {
            inv_half(m[IDX(ix, 0, kz)], m[IDX(ix, 1, kz)], m[IDX(ix, 2, kz)]);
            const float t = m[IDX(ix, 0, kz)];
            m[IDX(ix, 0, kz)] = m[IDX(ix, 1, kz)];
            m[IDX(ix, 1, kz)] = t;
        }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        inv_half(m[i * 3 + 0], m[i * 3 + 1], m[i * 3 + 2]);
        const float t = m[i * 3 + 0];
        m[i * 3 + 0] = m[i * 3 + 1];
        m[i * 3 + 1] = t;
    }
}

// the same for a force (fx, fy, 0) (gravity along x or y: one of the two is zero)
__device__ __forceinline__ void add_guo_xy(float g[27], float fx, float fy, float ux, float uy, float uz) {
    const float ux_fx[3] = {ux * fx, uy * fx, uz * fx};
    const float uy_fy[3] = {ux * fy, uy * fy, uz * fy};
    const float uFx = ux * fx, uFy = uy * fy;
#define GM(a, b, c)                                                                   \
    g[IDX(a, b, c)] += half_scale(a, b, c) * (guo_moment<a, b, c, 0>(fx, uFx, ux_fx) + \
                                              guo_moment<a, b, c, 1>(fy, uFy, uy_fy));
    GM(0, 0, 0) GM(0, 0, 1) GM(0, 0, 2) GM(0, 1, 0) GM(0, 1, 1) GM(0, 1, 2) GM(0, 2, 0) GM(0, 2, 1) GM(0, 2, 2)
    GM(1, 0, 0) GM(1, 0, 1) GM(1, 0, 2) GM(1, 1, 0) GM(1, 1, 1) GM(1, 1, 2) GM(1, 2, 0) GM(1, 2, 1) GM(1, 2, 2)
    GM(2, 0, 0) GM(2, 0, 1) GM(2, 0, 2) GM(2, 1, 0) GM(2, 1, 1) GM(2, 1, 2) GM(2, 2, 0) GM(2, 2, 1) GM(2, 2, 2)
#undef GM
}

// hardware reciprocal, square root and reciprocal square root (about 1 ulp), for
// operands the collision keeps in the normal range (density >= 1e-6, relaxation
// times in [0.5, 0.95], speeds and stress norms away from the denormals)
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }

// Central moments -> raw moments along one axis in the halved form (see unshift_half),
// for the three patterns of known zeros of the post-collision central moments:
// (k0, 0, k2), (0, k1, 0) and (k0, k1, k2).
__device__ __forceinline__ void unshift_k0k2(float k0, float& k1, float& k2, float hu, float hu2) {
    k1 = hu * k0;
    k2 = fmaf(hu2, k0, k2);
}
__device__ __forceinline__ void unshift_k1(float k1, float& k2, float tu) { k2 = tu * k1; }

// g: populations in stored form (f - w when Shifted), overwritten with the post-collision values.
// The collision needs only the raw moments of order <= 2 (density, momentum and the
// second moments, from which the second central moments follow); the higher central
// moments are set by the closure. The post-collision central moments (14 non-zero of
// 27) go back to populations through raw moments kept in the halved form of
// unshift_half / inv_half; the back-transform skips the known zeros, and the Guo
// source enters as its 12 non-zero raw moments.
// GravZ: the caller guarantees the gravity axis is z whenever buoyancy is on (a
// separate instantiation: the x/y branch of the force costs registers)
template <bool Shifted, bool GravZ = false>
__device__ __forceinline__ CollideResult collide(const StepParams& P, float g[27], float T) {
    CollideResult r;
    LowMoments lm = low_moments(g);
    if (Shifted) {
        // the weights' own raw moments: 1 and 1/3 on the even ones of order <= 2
        lm.m000 += 1.0f;
        lm.m200 += 1.0f / 3.0f;
        lm.m020 += 1.0f / 3.0f;
        lm.m002 += 1.0f / 3.0f;
    }
    const float rho = lm.m000;
    const float jx = lm.m100, jy = lm.m010, jz = lm.m001;
    r.ok = finite(rho) && finite(jx) && finite(jy) && finite(jz);
    if (!r.ok) return r;
    const float rho_s = fmaxf(1e-6f, rho);
    const float ir = rcp_hw(rho_s);
    float ux = jx * ir, uy = jy * ir, uz = jz * ir;
    // Boussinesq force along the gravity axis; F/rho = clamp(beta T)
    float b = 0.0f;
    if (P.buoyancy_on) {
        b = clampf(P.beta * T, -0.02f, 0.02f);
        if (P.g_sign > 0) b = -b;
        if (!GravZ && P.g_axis == 0) ux = fmaf(0.5f, b, ux);
        else if (!GravZ && P.g_axis == 1) uy = fmaf(0.5f, b, uy);
        else uz = fmaf(0.5f, b, uz);
    }
    {
        const float s2 = ux * ux + uy * uy + uz * uz;
        if (s2 > P.kmax_speed * P.kmax_speed) {
            const float sc = P.kmax_speed * rsq_hw(s2);
            ux *= sc;
            uy *= sc;
            uz *= sc;
        }
    }
    r.ux = ux;
    r.uy = uy;
    r.uz = uz;
    // second central moments about u: sum f (c_a - u_a)(c_b - u_b)
    const float c200 = fmaf(ux * ux, rho, fmaf(-2.0f * ux, lm.m100, lm.m200));
    const float c020 = fmaf(uy * uy, rho, fmaf(-2.0f * uy, lm.m010, lm.m020));
    const float c002 = fmaf(uz * uz, rho, fmaf(-2.0f * uz, lm.m001, lm.m002));
    const float c110 = fmaf(ux * uy, rho, fmaf(-ux, lm.m010, fmaf(-uy, lm.m100, lm.m110)));
    const float c101 = fmaf(ux * uz, rho, fmaf(-ux, lm.m001, fmaf(-uz, lm.m100, lm.m101)));
    const float c011 = fmaf(uy * uz, rho, fmaf(-uy, lm.m001, fmaf(-uz, lm.m010, lm.m011)));

    const float cs2 = 1.0f / 3.0f;
    float ts = P.tau_shear_base, tn = P.tau_normal_base;
    if (P.sgs_on) {
        // ref: jni.cpp:2346-2366
        const float nu0 = fmaxf(1e-6f, P.nu_base);
        const float e = rho_s * cs2;
        const float nxx = c200 - e, nyy = c020 - e, nzz = c002 - e;
        const float qn2 = nxx * nxx + nyy * nyy + nzz * nzz + 2.0f * (c110 * c110 + c101 * c101 + c011 * c011);
        const float smag = sqrt_hw(fmaxf(0.0f, qn2)) * rcp_hw(fmaxf(1e-6f, 2.0f * rho_s * nu0));
        const float nut = fminf(P.sgs_c2 * smag, P.nut_cap_ratio * nu0);
        const float tau_min = 0.5f + 3.0f * 1.0e-8f;
        ts = clampf(0.5f + 3.0f * clampf(nu0 + nut, 1.0e-8f, (0.95f - 0.5f) / 3.0f), tau_min, 0.95f);
        tn = clampf(P.tau_normal_base + (ts - P.tau_shear_base) * 0.30f, tau_min, 0.95f);
    }
    const float rts = rcp_hw(ts);
    const float sd = clampf(rcp_hw(tn), 0.0f, 1.95f);
    const float so = clampf(rts, 0.0f, 1.95f);
    // relaxation + Gaussian closure (ref: jni.cpp:1384-1430)
    const float eq2 = rho_s * cs2;
    const float d200 = fmaf(sd, eq2 - c200, c200);
    const float d020 = fmaf(sd, eq2 - c020, c020);
    const float d002 = fmaf(sd, eq2 - c002, c002);
    const float om = 1.0f - so;
    const float d110 = om * c110, d101 = om * c101, d011 = om * c011;
    const float t110 = 2.0f * d110, t101 = 2.0f * d101, t011 = 2.0f * d011;
    // post-collision central moments in the halved form (scale 2^-(non-zero indices));
    // the 1/rho of the closure carries the scale of the fourth-order ones
    const float irq = 0.25f * ir, ire = 0.125f * ir;
    const float r0 = rho_s;
    const float A = 0.5f * d200, B = 0.5f * d020, C = 0.5f * d002;
    const float Dxy = 0.25f * d110, Dxz = 0.25f * d101, Dyz = 0.25f * d011;
    const float Exy = (d200 * d020 + t110 * d110) * irq;
    const float Exz = (d200 * d002 + t101 * d101) * irq;
    const float Eyz = (d020 * d002 + t011 * d011) * irq;
    const float Fx = (d200 * d011 + t110 * d101) * ire;
    const float Fy = (d020 * d101 + t110 * d011) * ire;
    const float Fz = (d002 * d110 + t101 * d011) * ire;
    const float G = (d200 * d020 * d002 + t110 * d110 * d002 + t101 * d101 * d020 + t011 * d011 * d200 +
                     8.0f * d110 * d101 * d011) * ire * ir;
    // central -> raw, z lines (a, b): (a,b,0), (a,b,1), (a,b,2); zero where marked
    {
        const float hz = 0.5f * uz, tz = 2.0f * uz, hz2 = 0.5f * uz * uz;
        g[IDX(0, 0, 0)] = r0;   g[IDX(0, 0, 2)] = C;   unshift_k0k2(r0, g[IDX(0, 0, 1)], g[IDX(0, 0, 2)], hz, hz2);
        g[IDX(1, 0, 0)] = 0.0f; g[IDX(1, 0, 1)] = Dxz; unshift_k1(Dxz, g[IDX(1, 0, 2)], tz);
        g[IDX(2, 0, 0)] = A;    g[IDX(2, 0, 2)] = Exz; unshift_k0k2(A, g[IDX(2, 0, 1)], g[IDX(2, 0, 2)], hz, hz2);
        g[IDX(0, 1, 0)] = 0.0f; g[IDX(0, 1, 1)] = Dyz; unshift_k1(Dyz, g[IDX(0, 1, 2)], tz);
        g[IDX(1, 1, 0)] = Dxy;  g[IDX(1, 1, 2)] = Fz;  unshift_k0k2(Dxy, g[IDX(1, 1, 1)], g[IDX(1, 1, 2)], hz, hz2);
        g[IDX(2, 1, 0)] = 0.0f; g[IDX(2, 1, 1)] = Fx;  unshift_k1(Fx, g[IDX(2, 1, 2)], tz);
        g[IDX(0, 2, 0)] = B;    g[IDX(0, 2, 2)] = Eyz; unshift_k0k2(B, g[IDX(0, 2, 1)], g[IDX(0, 2, 2)], hz, hz2);
        g[IDX(1, 2, 0)] = 0.0f; g[IDX(1, 2, 1)] = Fy;  unshift_k1(Fy, g[IDX(1, 2, 2)], tz);
        g[IDX(2, 2, 0)] = Exy;  g[IDX(2, 2, 2)] = G;   unshift_k0k2(Exy, g[IDX(2, 2, 1)], g[IDX(2, 2, 2)], hz, hz2);
    }
    // y lines (a, c); (1,0,0), (0,1,0), (2,1,0), (1,2,0) are still zero
    {
        const float hy = 0.5f * uy, ty = 2.0f * uy, hy2 = 0.5f * uy * uy;
        unshift_k0k2(r0, g[IDX(0, 1, 0)], g[IDX(0, 2, 0)], hy, hy2);
        unshift_k1(Dxy, g[IDX(1, 2, 0)], ty);
        unshift_k0k2(A, g[IDX(2, 1, 0)], g[IDX(2, 2, 0)], hy, hy2);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 1; c < 3; ++c)
                unshift_half(g[IDX(a, 0, c)], g[IDX(a, 1, c)], g[IDX(a, 2, c)], hy, ty, hy2);
    }
    // x lines (b, c); (1,0,0) is still zero
    {
        const float hx = 0.5f * ux, tx = 2.0f * ux, hx2 = 0.5f * ux * ux;
        unshift_k0k2(r0, g[IDX(1, 0, 0)], g[IDX(2, 0, 0)], hx, hx2);
#pragma unroll
        for (int k = 1; k < 9; ++k) unshift_half(g[IDX(0, k / 3, k % 3)], g[IDX(1, k / 3, k % 3)], g[IDX(2, k / 3, k % 3)], hx, tx, hx2);
    }
    // Guo source (ref: jni.cpp:1270-1281, added after the collision) as raw moments
    if (P.buoyancy_on && b != 0.0f) {
        const float pref = fmaxf(0.0f, 1.0f - 0.5f * clampf(rts, 0.0f, 1.999f));
        const float fa = pref * rho_s * b;  // P * F along the axis
        // gravity along z (the common case) with the x and y force terms left out
        if (GravZ || P.g_axis == 2) add_guo_z(g, fa, ux, uy, uz);
        else add_guo_xy(g, P.g_axis == 0 ? fa : 0.0f, P.g_axis == 1 ? fa : 0.0f, ux, uy, uz);
    }
    if (Shifted) {
        // remove the weights' raw moments (1, 1/3, 1/9, 1/27), in the halved form
        g[IDX(0, 0, 0)] -= 1.0f;
        g[IDX(2, 0, 0)] -= 1.0f / 6.0f;
        g[IDX(0, 2, 0)] -= 1.0f / 6.0f;
        g[IDX(0, 0, 2)] -= 1.0f / 6.0f;
        g[IDX(2, 2, 0)] -= 1.0f / 36.0f;
        g[IDX(2, 0, 2)] -= 1.0f / 36.0f;
        g[IDX(0, 2, 2)] -= 1.0f / 36.0f;
        g[IDX(2, 2, 2)] -= 1.0f / 216.0f;
    }
    inverse_half(g);
    return r;
}

// reference reseed for a non-finite cell (jni.cpp:1724-1750)
template <bool Shifted>
__device__ __forceinline__ void reseed(const StepParams& P, const CellFields& C, int cell, float g[27], CollideResult& r) {
    const size_t N = (size_t)P.cells;
    const float p = clampf(finite(C.ref_p[cell]) ? C.ref_p[cell] : 0.0f, -0.03f, 0.03f);
    const float rho = clampf(1.0f + p, 0.97f, 1.03f);
    float ux = C.ref_u[cell], uy = C.ref_u[N + cell], uz = C.ref_u[2 * N + cell];
    ux = finite(ux) ? ux : 0.0f;
    uy = finite(uy) ? uy : 0.0f;
    uz = finite(uz) ? uz : 0.0f;
    const float s2 = ux * ux + uy * uy + uz * uz;
    if (!finite(s2) || s2 > P.kmax_speed * P.kmax_speed) {
        if (!finite(s2) || s2 <= 0.0f) ux = uy = uz = 0.0f;
        else {
            const float sc = P.kmax_speed / sqrtf(s2);
            ux *= sc;
            uy *= sc;
            uz *= sc;
        }
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) g[q] = feq_s<Shifted>(q, rho, ux, uy, uz);
    r.ux = ux;
    r.uy = uy;
    r.uz = uz;
    r.ok = true;
}

template <bool Shifted>
__device__ __forceinline__ void sponge(const StepParams& P, int x, int y, int z, float g[27]) {
    int d = min(min(x, y), z);
    d = min(d, min(min(P.nx - 1 - x, P.ny - 1 - y), P.nz - 1 - z));
    if (d >= 4) return;
    const float eta = (float)(4 - d) * 0.25f;
    const float al = clampf(0.03f * eta * eta, 0.0f, 0.95f);
    const float keep = 1.0f - al;
#pragma unroll
    for (int q = 0; q < 27; ++q) g[q] = keep * g[q] + al * (Shifted ? 0.0f : wq(q));
}

}  // namespace fast
}  // namespace aero
