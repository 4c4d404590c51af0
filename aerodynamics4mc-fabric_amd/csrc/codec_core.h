// Block coder of the analysis codec: one 4x4x4 block of a 3D float grid to the bits
// zfp 1.0.1 writes for it in fixed-accuracy mode (zfp/... citations are to the zfp
// tree the reference vendors, native/third_party/zfp). Host and device code: the
// encoder kernel runs it one block per lane.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <cstdint>

#define AERO_HD __host__ __device__

namespace aero {
namespace codec {

constexpr int kSlotWords = 36;    // scratch words per block: a block codes to at most 2120 bits
constexpr int kTile = 64;         // blocks per encoder workgroup (one wavefront)
constexpr int kMaxBlockBits = 2120;  // 1 + 8 + 63 + 64 * 32 (zfp/src/zfp.c:744-784 for floats)

struct GridDesc {
    const float* values;
    unsigned long long out_word;   // first output word of this grid's stream
    unsigned long long cap_words;  // output words the caller provides for it
    unsigned long long header[3];  // header bits, least significant first
    int nx, ny, nz;
    int bx, by;                    // blocks along x and y
    int nblocks;
    int tile_begin, ntiles;
    int minexp;
    int header_bits;
};

// zfp/src/template/codec3.c:4-87: coefficient order, by total sequency, as
// (i, j, k) -> i + 4 j + 16 k
AERO_HD constexpr int perm3(int n) {
    constexpr unsigned char p[64] = {
        0, 1, 4, 16, 20, 17, 5, 2, 8, 32, 21, 6, 18, 24, 9, 33, 36, 3, 12, 48, 22, 25,
        37, 40, 34, 10, 7, 19, 28, 13, 49, 52, 41, 38, 26, 23, 29, 53, 11, 35, 44, 14, 50, 56,
        42, 27, 39, 45, 30, 54, 57, 60, 51, 15, 43, 46, 58, 61, 55, 31, 62, 59, 47, 63};
    return p[n];
}

// ---- device pieces

AERO_HD inline int wrap_add(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
AERO_HD inline int wrap_sub(int a, int b) { return (int)((unsigned)a - (unsigned)b); }

// zfp/src/template/encode.c:31-56
AERO_HD inline void fwd_lift(int& x, int& y, int& z, int& w) {
    x = wrap_add(x, w); x >>= 1; w = wrap_sub(w, x);
    z = wrap_add(z, y); z >>= 1; y = wrap_sub(y, z);
    x = wrap_add(x, z); x >>= 1; z = wrap_sub(z, x);
    w = wrap_add(w, y); w >>= 1; y = wrap_sub(y, w);
    w = wrap_add(w, y >> 1); y = wrap_sub(y, w >> 1);
}

// in-register transpose of a 32x32 bit matrix: bit c of row r moves to bit r of row c
AERO_HD inline void transpose32(uint32_t* a) {
#pragma unroll
    for (int j = 16, s = 0; j; j >>= 1, s++) {
        const uint32_t m = s == 0 ? 0x0000ffffu : s == 1 ? 0x00ff00ffu : s == 2 ? 0x0f0f0f0fu : s == 3 ? 0x33333333u : 0x55555555u;
#pragma unroll
        for (int r = 0; r < 32; r++) {
            if (r & j) continue;
            const uint32_t t = ((a[r] >> j) ^ a[r + j]) & m;
            a[r + j] ^= t;
            a[r] ^= t << j;
        }
    }
}

// rows 16..31 of the same transpose only (the 16 most significant bit planes), in
// about half the operations: the first stage moves every row's high half into rows
// 16..31, the remaining stages stay inside them
AERO_HD inline void transpose32_high(uint32_t* a) {
#pragma unroll
    for (int r = 0; r < 16; r++) a[r + 16] ^= ((a[r] >> 16) ^ a[r + 16]) & 0x0000ffffu;
#pragma unroll
    for (int j = 8, s = 1; j; j >>= 1, s++) {
        const uint32_t m = s == 1 ? 0x00ff00ffu : s == 2 ? 0x0f0f0f0fu : s == 3 ? 0x33333333u : 0x55555555u;
#pragma unroll
        for (int r = 16; r < 32; r++) {
            if (r & j) continue;
            const uint32_t t = ((a[r] >> j) ^ a[r + j]) & m;
            a[r + j] ^= t;
            a[r] ^= t << j;
        }
    }
}

// true if the condition holds for any lane of the wavefront (the lane itself on the host)
AERO_HD inline bool any_lane(bool v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __any(v);
#else
    return v;
#endif
}

// appends bit fields, least significant first, to a block's scratch slot, whose
// words lie `stride` words apart
struct BitSink {
    unsigned long long* dst;
    unsigned stride;
    unsigned long long acc;
    uint32_t fill;   // bits held in acc, < 64
    uint32_t words;  // words stored
    // 1 <= n <= 64; v holds no bits at or above n
    AERO_HD inline void put(unsigned long long v, uint32_t n) {
        acc |= v << fill;
        const uint32_t f = fill + n;
        if (f >= 64) {
            dst[(size_t)words * stride] = acc;
            words++;
            acc = fill ? v >> (64 - fill) : 0ull;
        }
        fill = f & 63u;
    }
};

// x86-64's truncating float -> int32 conversion: a NaN or a value outside the int32
// range converts to INT32_MIN (the reference build's cast of s * x,
// zfp/src/template/codecf.c:13-20)
AERO_HD inline int cvt_trunc(float f) {
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
}

// codes block b of grid G into the slot at dst (word i at dst[i * stride]); returns
// its length in bits
AERO_HD inline uint32_t encode_block(const GridDesc& G, int b, unsigned long long* dst, unsigned stride) {
    const int bxi = b % G.bx;
    const int rest = b / G.bx;
    const int byi = rest % G.by;
    const int bzi = rest / G.by;
    const int x0 = 4 * bxi, y0 = 4 * byi, z0 = 4 * bzi;
    const size_t sy = (size_t)G.nx, sz = (size_t)G.nx * (size_t)G.ny;
    float v[64];
    if (x0 + 4 <= G.nx && y0 + 4 <= G.ny && z0 + 4 <= G.nz) {
        const float* base = G.values + (size_t)x0 + sy * y0 + sz * z0;
        if (((G.nx & 3) == 0) && ((((uintptr_t)G.values) & 15) == 0)) {
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float4 r = *reinterpret_cast<const float4*>(base + sy * j + sz * k);
                    v[16 * k + 4 * j + 0] = r.x;
                    v[16 * k + 4 * j + 1] = r.y;
                    v[16 * k + 4 * j + 2] = r.z;
                    v[16 * k + 4 * j + 3] = r.w;
                }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int i = 0; i < 4; i++) v[16 * k + 4 * j + i] = base[i + sy * j + sz * k];
        }
    } else {
        // a partial block: zfp pads a row of n < 4 values as (0,0,0,0) for n = 1,
        // (0,1,1,0) for n = 2 and (0,1,2,0) for n = 3, along x, then y, then z, which
        // composes into one index map per axis
        const int nbx = G.nx - x0 < 4 ? G.nx - x0 : 4, nby = G.ny - y0 < 4 ? G.ny - y0 : 4, nbz = G.nz - z0 < 4 ? G.nz - z0 : 4;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int kk = k < nbz ? k : (nbz == 2 && k == 2 ? 1 : 0);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int jj = j < nby ? j : (nby == 2 && j == 2 ? 1 : 0);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int ii = i < nbx ? i : (nbx == 2 && i == 2 ? 1 : 0);
                    v[16 * k + 4 * j + i] = G.values[(size_t)(x0 + ii) + sy * (y0 + jj) + sz * (z0 + kk)];
                }
            }
        }
    }
    // largest magnitude as float bits (a NaN never compares greater,
    // zfp/src/template/encodef.c:31-40); blocks without NaN or infinity skip the
    // NaN filter here and the range checks of the conversion below
    uint32_t mx = 0;
#pragma unroll
    for (int i = 0; i < 64; i++) {
        const uint32_t a = __builtin_bit_cast(uint32_t, v[i]) & 0x7fffffffu;
        mx = mx > a ? mx : a;
    }
    const bool special = any_lane(mx >= 0x7f800000u);
    if (special) {
        mx = 0;
#pragma unroll
        for (int i = 0; i < 64; i++) {
            uint32_t a = __builtin_bit_cast(uint32_t, v[i]) & 0x7fffffffu;
            a = a > 0x7f800000u ? 0u : a;
            mx = mx > a ? mx : a;
        }
    }
    // its frexp exponent, clamped for subnormals; glibc leaves 0 for an infinity
    // (zfp/src/template/encodef.c:11-28)
    const int emax = mx == 0u ? -127 : (mx == 0x7f800000u ? 0 : ((mx >> 23) == 0u ? -126 : (int)(mx >> 23) - 126));
    int prec = emax - G.minexp + 8;  // zfp/src/template/codecf.c:6-13, dims = 3
    prec = prec < 0 ? 0 : (prec > 64 ? 64 : prec);
    if (emax == -127 || prec == 0) {
        dst[0] = 0ull;  // one zero bit: every value is zero at this tolerance
        return 1u;
    }
    BitSink sink{dst, stride, 0ull, 0u, 0u};
    sink.put(2ull * (unsigned long long)(emax + 127) + 1ull, 9);
    const int se = 30 - emax;  // zfp/src/template/encodef.c:43-58: s = 2^(30 - emax) in float
    const float s = se <= 127 ? __builtin_bit_cast(float, (uint32_t)(se + 127) << 23) : __builtin_bit_cast(float, 0x7f800000u);
    int q[64];
    if (special || any_lane(se > 127)) {
#pragma unroll
        for (int i = 0; i < 64; i++) q[i] = cvt_trunc(s * v[i]);
    } else {  // finite values and a finite scale: |s * v| < 2^30
#pragma unroll
        for (int i = 0; i < 64; i++) q[i] = (int)(s * v[i]);
    }
    // zfp/src/template/encode3.c:35-51: along x, then y, then z
#pragma unroll
    for (int z = 0; z < 4; z++)
#pragma unroll
        for (int y = 0; y < 4; y++) fwd_lift(q[16 * z + 4 * y], q[16 * z + 4 * y + 1], q[16 * z + 4 * y + 2], q[16 * z + 4 * y + 3]);
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int z = 0; z < 4; z++) fwd_lift(q[16 * z + x], q[16 * z + x + 4], q[16 * z + x + 8], q[16 * z + x + 12]);
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
        for (int x = 0; x < 4; x++) fwd_lift(q[x + 4 * y], q[x + 4 * y + 16], q[x + 4 * y + 32], q[x + 4 * y + 48]);
    // sequency order and negabinary (zfp/src/template/encode.c:74-89), split into the
    // two 32-coefficient halves whose transposes give the planes' low and high words
    uint32_t lo[32], hi[32];
#pragma unroll
    for (int i = 0; i < 32; i++) {
        lo[i] = ((uint32_t)q[perm3(i)] + 0xaaaaaaaau) ^ 0xaaaaaaaau;
        hi[i] = ((uint32_t)q[perm3(i + 32)] + 0xaaaaaaaau) ^ 0xaaaaaaaau;
    }
    // zfp/src/template/encode.c:178-207, planes 31 down to 32 - prec
    const int kmin = prec < 32 ? 32 - prec : 0;
    if (any_lane(kmin < 16)) {
        transpose32(lo);
        transpose32(hi);
    } else {
        transpose32_high(lo);
        transpose32_high(hi);
    }
    uint32_t n = 0;  // coefficients already significant
#pragma unroll
    for (int k = 31; k >= 0; k--) {
        if (k >= kmin) {
            const unsigned long long x = (unsigned long long)lo[k] | ((unsigned long long)hi[k] << 32);
            if (n) sink.put(n == 64 ? x : (x & ((1ull << n) - 1ull)), n);
            if (n < 64) {
                // group tests: "0" ends the plane; otherwise "1", the zeros before the
                // next one bit and the one itself, whose bit is implied when it is the
                // last coefficient
                unsigned long long r = x >> n;
                bool more = true;
                while (more) {
                    const uint32_t t = r ? (uint32_t)__builtin_ctzll(r) : 0u;
                    const bool last = r && n + t == 63u;
                    more = r && !last;
                    const unsigned long long field = more ? (1ull | (2ull << t)) : (r ? 1ull : 0ull);
                    sink.put(field, r ? (last ? t + 1 : t + 2) : 1u);
                    n = r ? (last ? 64u : n + t + 1) : n;
                    r = more ? r >> (t + 1) : 0ull;
                }
            }
        }
    }
    if (sink.fill) dst[(size_t)sink.words * stride] = sink.acc;
    return sink.words * 64u + sink.fill;
}

// n <= 64 bits of a bit array starting at bit p
AERO_HD inline unsigned long long read_bits(const unsigned long long* w, unsigned long long p, uint32_t n) {
    const unsigned long long i = p >> 6;
    const uint32_t sh = (uint32_t)(p & 63ull);
    unsigned long long v = w[i] >> sh;
    if (sh && sh + n > 64) v |= w[i + 1] << (64 - sh);
    return n >= 64 ? v : (v & ((1ull << n) - 1ull));
}

// n <= 64 bits of slot (tile, lane) from bit p
AERO_HD inline unsigned long long slot_bits(const unsigned long long* scratch, size_t tile, int lane,
                                                        uint32_t p, uint32_t n) {
    const unsigned long long* w = scratch + tile * kSlotWords * kTile + lane;
    const uint32_t i = p >> 6, sh = p & 63u;
    unsigned long long v = w[(size_t)i * kTile] >> sh;
    if (sh && sh + n > 64) v |= w[(size_t)(i + 1) * kTile] << (64 - sh);
    return n >= 64 ? v : (v & ((1ull << n) - 1ull));
}

// the 64 stream bits from bit p, which lies in the header (p < H) or in block (t, j)
// of the grid that starts at bit bstart; zero past the last block
AERO_HD inline unsigned long long gather_word(const GridDesc& G, const unsigned long long* __restrict__ scratch,
                                          const uint32_t* __restrict__ len, unsigned long long p, int t, int j,
                                          unsigned long long bstart) {
    const unsigned long long H = (unsigned long long)G.header_bits;
    unsigned long long word = 0;
    uint32_t filled = 0;
    if (p < H) {
        const uint32_t take = (uint32_t)(H - p < 64ull ? H - p : 64ull);
        word = read_bits(G.header, p, take);
        filled = take;
        p += take;
    }
    while (filled < 64 && t * kTile + j < G.nblocks) {
        const size_t tile = (size_t)(G.tile_begin + t);
        const uint32_t L = len[tile * kTile + j];
        const unsigned long long within = p - bstart;
        if (within < L) {
            const uint32_t take = 64u - filled < L - (uint32_t)within ? 64u - filled : L - (uint32_t)within;
            const unsigned long long bits = slot_bits(scratch, tile, j, (uint32_t)within, take);
            word |= filled ? bits << filled : bits;
            filled += take;
            p += take;
        } else {  // on to the next block
            bstart += L;
            if (++j == kTile) {
                j = 0;
                t++;
            }
        }
    }
    return word;
}

}  // namespace codec
}  // namespace aero
