/* Test infrastructure (the checker, not the product): a plain-C restatement of the
 * zfp 1.0.1 coder for 3D float fields, as the reference's analysis codec drives it
 * (native/src/aero_lbm_analysis_codec.cpp:19-99 compress with
 * zfp_stream_set_accuracy and a full header, :101-172 decompress).  zfp is not a
 * reference source file but a library the reference vendors
 * (native/third_party/zfp, version 1.0.1, native/third_party/zfp/include/zfp/version.h:5-14);
 * the citations below are to those vendored files, abbreviated "zfp/".
 *
 * Written for clarity, one bit at a time, the way the specification reads; the
 * product's encoder builds each bit plane's code from bit counts instead, so the two
 * cross-check each other.  Parity of this file is pinned bit for bit against the
 * reference codec compiled from its own sources (oracle/_ref/libcodec_ref.so) by
 * the CPU tests.
 *
 * Float-to-int conversion: zfp casts s*x to int (zfp/src/template/codecf.c:13-20)
 * and the reference is built for x86-64, where the truncating conversion of a NaN or
 * of a value outside the int32 range gives INT32_MIN; cvt_x86() states that.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- bit stream: bits are stored least significant first in 64-bit words
 * (zfp/include/zfp/bitstream.inl:67-92), i.e. bit i of the stream is bit (i % 8) of
 * byte i / 8 on a little-endian host. */
typedef struct {
    uint8_t* buf;
    uint64_t cap_bits;
    uint64_t pos;
    int overflow;
} bitw;

static void put_bit(bitw* w, unsigned b) {
    if (w->pos >= w->cap_bits) {
        w->overflow = 1;
        return;
    }
    if (b) w->buf[w->pos >> 3] |= (uint8_t)(1u << (w->pos & 7));
    w->pos++;
}

static void put_bits(bitw* w, uint64_t v, unsigned n) {
    for (unsigned i = 0; i < n; i++) put_bit(w, (unsigned)((v >> i) & 1u));
}

typedef struct {
    const uint8_t* buf;
    uint64_t size_bits;
    uint64_t pos;
    int overrun;
} bitr;

static unsigned get_bit(bitr* r) {
    if (r->pos >= r->size_bits) {
        r->overrun = 1;
        r->pos++;
        return 0;
    }
    unsigned b = (r->buf[r->pos >> 3] >> (r->pos & 7)) & 1u;
    r->pos++;
    return b;
}

static uint64_t get_bits(bitr* r, unsigned n) {
    uint64_t v = 0;
    for (unsigned i = 0; i < n; i++) v |= (uint64_t)get_bit(r) << i;
    return v;
}

/* ---- block transform pieces */

/* zfp/src/template/codec3.c:4-87: coefficient order, by total sequency i+j+k, as
 * (i, j, k) triples; the coefficient of (i, j, k) sits at i + 4 j + 16 k */
static const uint8_t kPermIJK[64][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {2, 0, 0},
    {0, 2, 0}, {0, 0, 2}, {1, 1, 1}, {2, 1, 0}, {2, 0, 1}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2},
    {0, 1, 2}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}, {0, 2, 2},
    {2, 0, 2}, {2, 2, 0}, {3, 1, 0}, {3, 0, 1}, {0, 3, 1}, {1, 3, 0}, {1, 0, 3}, {0, 1, 3},
    {1, 2, 2}, {2, 1, 2}, {2, 2, 1}, {3, 1, 1}, {1, 3, 1}, {1, 1, 3}, {3, 2, 0}, {3, 0, 2},
    {0, 3, 2}, {2, 3, 0}, {2, 0, 3}, {0, 2, 3}, {2, 2, 2}, {3, 2, 1}, {3, 1, 2}, {1, 3, 2},
    {2, 3, 1}, {2, 1, 3}, {1, 2, 3}, {0, 3, 3}, {3, 0, 3}, {3, 3, 0}, {3, 2, 2}, {2, 3, 2},
    {2, 2, 3}, {1, 3, 3}, {3, 1, 3}, {3, 3, 1}, {2, 3, 3}, {3, 2, 3}, {3, 3, 2}, {3, 3, 3},
};
static uint8_t kPerm3[64];

static void init_perm(void) {
    for (int n = 0; n < 64; n++) kPerm3[n] = (uint8_t)(kPermIJK[n][0] + 4 * kPermIJK[n][1] + 16 * kPermIJK[n][2]);
}

static int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static int32_t wrap_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

/* zfp/src/template/encode.c:31-56: forward lifting step on 4 values at stride s */
static void fwd_lift(int32_t* p, int s) {
    int32_t x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
    x = wrap_add(x, w); x >>= 1; w = wrap_sub(w, x);
    z = wrap_add(z, y); z >>= 1; y = wrap_sub(y, z);
    x = wrap_add(x, z); x >>= 1; z = wrap_sub(z, x);
    w = wrap_add(w, y); w >>= 1; y = wrap_sub(y, w);
    w = wrap_add(w, y >> 1); y = wrap_sub(y, w >> 1);
    p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

/* zfp/src/template/decode.c:9-45: inverse lifting step */
static void inv_lift(int32_t* p, int s) {
    int32_t x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
    y = wrap_add(y, w >> 1); w = wrap_sub(w, y >> 1);
    y = wrap_add(y, w); w = wrap_sub(w, wrap_sub(y, w));
    z = wrap_add(z, x); x = wrap_sub(x, wrap_sub(z, x));
    y = wrap_add(y, z); z = wrap_sub(z, wrap_sub(y, z));
    w = wrap_add(w, x); x = wrap_sub(x, wrap_sub(w, x));
    p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

/* zfp/src/template/encode3.c:35-51: lifts along x, then y, then z */
static void fwd_xform3(int32_t* p) {
    for (int z = 0; z < 4; z++)
        for (int y = 0; y < 4; y++) fwd_lift(p + 4 * y + 16 * z, 1);
    for (int x = 0; x < 4; x++)
        for (int z = 0; z < 4; z++) fwd_lift(p + 16 * z + x, 4);
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) fwd_lift(p + x + 4 * y, 16);
}

/* zfp/src/template/decode3.c:27-43: lifts along z, then y, then x */
static void inv_xform3(int32_t* p) {
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) inv_lift(p + x + 4 * y, 16);
    for (int x = 0; x < 4; x++)
        for (int z = 0; z < 4; z++) inv_lift(p + 16 * z + x, 4);
    for (int z = 0; z < 4; z++)
        for (int y = 0; y < 4; y++) inv_lift(p + 4 * y + 16 * z, 1);
}

/* zfp/src/template/encode.c:9-29: fill a partial row of n <= 4 values (stride s) */
static void pad_row(float* p, int n, int s) {
    switch (n) {
        case 0: p[0] = 0.0f; /* fall through */
        case 1: p[1 * s] = p[0]; /* fall through */
        case 2: p[2 * s] = p[1 * s]; /* fall through */
        case 3: p[3 * s] = p[0]; /* fall through */
        default: break;
    }
}

static int32_t cvt_x86(float f) {
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int32_t)f;
    return INT32_MIN; /* NaN and out of range */
}

/* zfp/src/template/codecf.c:6-13 with dims = 3 and no rounding mode */
static unsigned block_precision(int emax, unsigned maxprec, int minexp) {
    int p = emax - minexp + 8;
    if (p < 0) p = 0;
    return (unsigned)p < maxprec ? (unsigned)p : maxprec;
}

/* zfp/src/template/encodef.c:11-40: exponent of the block's largest magnitude */
static int block_exponent(const float* v) {
    float mx = 0.0f;
    for (int i = 0; i < 64; i++) {
        float f = fabsf(v[i]);
        if (mx < f) mx = f; /* a NaN never compares greater */
    }
    int e = -127;
    if (mx > 0) {
        if (isinf(mx)) {
            e = 0; /* glibc's frexpf leaves 0 for an infinity */
        } else {
            frexpf(mx, &e);
            if (e < -126) e = -126;
        }
    }
    return e;
}

/* zfp/src/template/encode.c:178-207: embedded coding of 64 negabinary coefficients,
 * bit plane by bit plane from the most significant, without a bit budget */
static void encode_planes(bitw* w, const uint32_t* u, unsigned maxprec) {
    unsigned kmin = 32 > maxprec ? 32 - maxprec : 0;
    unsigned n = 0;
    for (unsigned k = 32; k-- > kmin;) {
        uint64_t x = 0;
        for (unsigned i = 0; i < 64; i++) x |= (uint64_t)((u[i] >> k) & 1u) << i;
        for (unsigned i = 0; i < n; i++) put_bit(w, (unsigned)((x >> i) & 1u));
        x = n < 64 ? x >> n : 0;
        while (n < 64) {
            unsigned any = x != 0;
            put_bit(w, any); /* group test */
            if (!any) break;
            /* scan for the next one bit; the last position needs no bit of its own */
            while (n < 63) {
                unsigned b = (unsigned)(x & 1u);
                put_bit(w, b);
                if (b) break;
                x >>= 1;
                n++;
            }
            x >>= 1;
            n++;
        }
    }
}

/* zfp/src/template/encodef.c:61-88 with zfp/src/template/encode.c:258-280 */
static void encode_block(bitw* w, const float* v, int minexp) {
    int emax = block_exponent(v);
    unsigned maxprec = block_precision(emax, 64, minexp);
    unsigned e = maxprec ? (unsigned)(emax + 127) : 0u;
    if (!e) {
        put_bit(w, 0);
        return;
    }
    put_bits(w, 2u * e + 1u, 9);
    float s = ldexpf(1.0f, 30 - emax);
    int32_t q[64];
    for (int i = 0; i < 64; i++) q[i] = cvt_x86(s * v[i]);
    fwd_xform3(q);
    uint32_t u[64];
    for (int i = 0; i < 64; i++) u[i] = ((uint32_t)q[kPerm3[i]] + 0xaaaaaaaau) ^ 0xaaaaaaaau;
    encode_planes(w, u, maxprec);
}

/* zfp/src/template/encode3.c:14-33 gather of a block with at most 4 values per
 * axis, then padding along x, y and z */
static void gather_block(float* q, const float* f, int nx, int ny, int x0, int y0, int z0, int bx, int by, int bz) {
    for (int z = 0; z < bz; z++) {
        for (int y = 0; y < by; y++) {
            for (int x = 0; x < bx; x++)
                q[16 * z + 4 * y + x] = f[(size_t)(x0 + x) + (size_t)nx * ((size_t)(y0 + y) + (size_t)ny * (size_t)(z0 + z))];
            pad_row(q + 16 * z + 4 * y, bx, 1);
        }
        for (int x = 0; x < 4; x++) pad_row(q + 16 * z + x, by, 4);
    }
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) pad_row(q + 4 * y + x, bz, 16);
}

/* zfp/src/zfp.c:842-857: the tolerance's exponent, 2^minexp <= tolerance */
static int accuracy_minexp(double tolerance) {
    int e;
    frexp(tolerance, &e);
    return e - 1;
}

/* zfp/src/zfp.c:667-734: parameters of fixed-accuracy mode; the short 12-bit form
 * unless minexp is out of its range or the parameters equal zfp's defaults */
static void mode_bits(int minexp, uint64_t* mode, unsigned* nbits) {
    if (minexp != -1074 && minexp <= 843) {
        *mode = (uint64_t)(minexp + 1074) + 2177u;
        *nbits = 12;
        return;
    }
    int me = minexp + 16495;
    if (me < 0) me = 0;
    if (me > 0x7fff) me = 0x7fff;
    uint64_t m = (uint64_t)me;
    m = (m << 7) + 63u;     /* maxprec 64 - 1 */
    m = (m << 15) + 16657u; /* maxbits 16658 - 1 */
    m = (m << 15) + 0u;     /* minbits 1 - 1 */
    m = (m << 12) + 0xfffu;
    *mode = m;
    *nbits = 64;
}

/* Compressed size in bytes (a multiple of 8), or 0 on invalid input or too small a
 * buffer; zeroes and fills out[0 .. capacity). */
long long oracle_codec_compress_3f(const float* values, int nx, int ny, int nz, double tolerance, uint8_t* out,
                                  long long capacity) {
    if (!values || nx <= 0 || ny <= 0 || nz <= 0 || !out || capacity <= 0) return 0;
    if (!(tolerance > 0.0) || !isfinite(tolerance)) return 0;
    if (nx > 65536 || ny > 65536 || nz > 65536) return 0; /* zfp/src/zfp.c:331-337 */
    init_perm();
    memset(out, 0, (size_t)capacity);
    bitw w = {out, (uint64_t)capacity * 8u, 0, 0};
    /* zfp/src/zfp.c:1228-1262 */
    put_bits(&w, 'z', 8);
    put_bits(&w, 'f', 8);
    put_bits(&w, 'p', 8);
    put_bits(&w, 5, 8);
    uint64_t meta = (uint64_t)(nz - 1);
    meta = (meta << 16) + (uint64_t)(ny - 1);
    meta = (meta << 16) + (uint64_t)(nx - 1);
    meta = (meta << 2) + 2u; /* three dimensions */
    meta = (meta << 2) + 2u; /* zfp_type_float - 1 */
    put_bits(&w, meta, 52);
    int minexp = accuracy_minexp(tolerance);
    uint64_t mode;
    unsigned mbits;
    mode_bits(minexp, &mode, &mbits);
    put_bits(&w, mode, mbits);
    /* zfp/src/template/compress.c:60-81: blocks with x fastest, then y, then z */
    float q[64];
    for (int z = 0; z < nz; z += 4)
        for (int y = 0; y < ny; y += 4)
            for (int x = 0; x < nx; x += 4) {
                int bx = nx - x < 4 ? nx - x : 4, by = ny - y < 4 ? ny - y : 4, bz = nz - z < 4 ? nz - z : 4;
                gather_block(q, values, nx, ny, x, y, z, bx, by, bz);
                encode_block(&w, q, minexp);
            }
    if (w.overflow) return 0;
    return (long long)(((w.pos + 63u) / 64u) * 8u); /* zfp/src/zfp.c: stream_flush */
}

/* ---- decoder */

typedef struct {
    unsigned minbits, maxbits, maxprec;
    int minexp;
} params;

/* zfp/src/zfp.c:859-905 */
static int params_from_mode(uint64_t mode, params* p) {
    if (mode <= 4094) {
        if (mode < 2048) {
            p->minbits = p->maxbits = (unsigned)mode + 1;
            p->maxprec = 64;
            p->minexp = -1074;
        } else if (mode < 2048 + 128) {
            p->minbits = 1;
            p->maxbits = 16658;
            p->maxprec = (unsigned)mode + 1 - 2048;
            p->minexp = -1074;
        } else if (mode == 2048 + 128) {
            p->minbits = 1;
            p->maxbits = 16658;
            p->maxprec = 64;
            p->minexp = -1075;
        } else {
            p->minbits = 1;
            p->maxbits = 16658;
            p->maxprec = 64;
            p->minexp = (int)mode - 1074 - 2177;
        }
    } else {
        mode >>= 12; p->minbits = (unsigned)(mode & 0x7fffu) + 1;
        mode >>= 15; p->maxbits = (unsigned)(mode & 0x7fffu) + 1;
        mode >>= 15; p->maxprec = (unsigned)(mode & 0x7fu) + 1;
        mode >>= 7;  p->minexp = (int)(mode & 0x7fffu) - 16495;
    }
    return p->minbits <= p->maxbits && p->maxprec > 0 && p->maxprec <= 64;
}

/* zfp/src/template/decode.c:80-185 (bit budget) and :188-220 (no budget) in one
 * loop: with a budget at least (maxprec + 1) * 64 - 1 bits they read the same bits */
static unsigned decode_planes(bitr* r, unsigned maxbits, unsigned maxprec, uint32_t* u) {
    unsigned kmin = 32 > maxprec ? 32 - maxprec : 0;
    unsigned bits = maxbits;
    unsigned n = 0;
    memset(u, 0, 64 * sizeof(uint32_t));
    for (unsigned k = 32; bits && k-- > kmin;) {
        unsigned m = n < bits ? n : bits;
        bits -= m;
        uint64_t x = get_bits(r, m);
        for (; bits && n < 64; n++) {
            bits--;
            if (get_bit(r)) {
                for (; bits && n < 63; n++) {
                    bits--;
                    if (get_bit(r)) break;
                }
                x += (uint64_t)1 << n;
            } else {
                break;
            }
        }
        for (unsigned i = 0; x; i++, x >>= 1) u[i] += (uint32_t)(x & 1u) << k;
    }
    return maxbits - bits;
}

/* zfp/src/template/decodef.c:7-36 and zfp/src/template/decode.c:283-300 */
static void decode_block(bitr* r, const params* p, float* v) {
    unsigned bits = 1;
    if (!get_bit(r)) {
        for (int i = 0; i < 64; i++) v[i] = 0.0f;
        if (p->minbits > bits) r->pos += p->minbits - bits;
        return;
    }
    bits += 8;
    int emax = (int)get_bits(r, 8) - 127;
    unsigned maxprec = block_precision(emax, p->maxprec, p->minexp);
    unsigned minbits = p->minbits - (bits < p->minbits ? bits : p->minbits);
    uint32_t u[64];
    unsigned used = decode_planes(r, p->maxbits - bits, maxprec, u);
    if (used < minbits) r->pos += minbits - used;
    int32_t q[64];
    for (int i = 0; i < 64; i++) q[kPerm3[i]] = (int32_t)((u[i] ^ 0xaaaaaaaau) - 0xaaaaaaaau);
    inv_xform3(q);
    float s = ldexpf(1.0f, emax - 30); /* zfp/src/template/codecf.c:22-32 */
    for (int i = 0; i < 64; i++) v[i] = s * (float)q[i];
}

/* 1 on success; 0 on a bad header, a reversible-mode stream, a shape mismatch or
 * a stream shorter than its blocks need */
int oracle_codec_decompress_3f(const uint8_t* data, long long size, int nx, int ny, int nz, float* out) {
    if (!data || size <= 0 || !out || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    bitr r = {data, (uint64_t)size * 8u, 0, 0};
    /* zfp/src/zfp.c:1265-1297 */
    if (get_bits(&r, 8) != 'z' || get_bits(&r, 8) != 'f' || get_bits(&r, 8) != 'p' || get_bits(&r, 8) != 5) return 0;
    uint64_t meta = get_bits(&r, 52);
    if ((meta & 3u) != 2u) return 0; /* float */
    meta >>= 2;
    if ((meta & 3u) != 2u) return 0; /* three dimensions */
    meta >>= 2;
    if ((int)(meta & 0xffffu) + 1 != nx || (int)((meta >> 16) & 0xffffu) + 1 != ny ||
        (int)((meta >> 32) & 0xffffu) + 1 != nz)
        return 0;
    uint64_t mode = get_bits(&r, 12);
    if (mode > 4094) mode += get_bits(&r, 52) << 12;
    init_perm();
    params p;
    if (!params_from_mode(mode, &p)) return 0;
    if (p.minexp < -1074) return 0; /* reversible mode is not restated here */
    float q[64];
    for (int z = 0; z < nz; z += 4)
        for (int y = 0; y < ny; y += 4)
            for (int x = 0; x < nx; x += 4) {
                decode_block(&r, &p, q);
                int bx = nx - x < 4 ? nx - x : 4, by = ny - y < 4 ? ny - y : 4, bz = nz - z < 4 ? nz - z : 4;
                for (int k = 0; k < bz; k++)
                    for (int j = 0; j < by; j++)
                        for (int i = 0; i < bx; i++)
                            out[(size_t)(x + i) + (size_t)nx * ((size_t)(y + j) + (size_t)ny * (size_t)(z + k))] =
                                q[16 * k + 4 * j + i];
            }
    return r.overrun ? 0 : 1;
}
