// Analysis codec: zfp streams of 3D float grids (zfp 1.0.1, fixed-accuracy mode,
// full header), byte for byte those of the reference's analysis codec
// (native/src/aero_lbm_analysis_codec.cpp:19-99 compress, :101-172 decompress),
// which drives the zfp library the reference vendors (native/third_party/zfp).
// Citations "zfp/..." are to that vendored tree.
//
// Encoder (device). One lane codes one 4x4x4 block, so a wavefront codes 64
// blocks with the same instruction stream:
//   1. gather the block (partial blocks repeat edge values as zfp pads them,
//      zfp/src/template/encode3.c:14-33 and encode.c:9-29),
//   2. common exponent, quantisation to 30-bit integers relative to it,
//      decorrelating lifting transform along x, y, z, sequency order and
//      negabinary (zfp/src/template/encodef.c:61-88, encode.c:31-80, encode3.c:35-51),
//   3. a 32x32 bit-matrix transpose of each half block turns 64 coefficients into
//      32 bit planes of 64 bits in registers,
//   4. embedded coding of each plane (zfp/src/template/encode.c:178-207): the first n
//      bits verbatim, then group tests, each run "1, zeros, 1" written as one field
//      whose length comes from a count of trailing zeros, into the block's slot of a
//      scratch buffer; the wavefront scans its 64 block lengths.
// A second kernel sums each workgroup's earlier tiles for its bit offset, and every
// block then writes the output words whose first bit it holds, taking the rest of a
// word from the blocks that follow (no atomics and no zero fill: each word is
// written once).
//
// Decoder (host): zfp's stream carries no block index, so block boundaries are only
// known by decoding every block before them; the decoder is a serial C++ pass
// (64-bit reads, trailing-zero counts for the group tests) like the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/aero_lbm_capi.h"
#include "codec_core.h"
#include "runtime_internal.h"

namespace aero {
namespace codec {

constexpr int kMaxGrids = 64;     // grids per batched call
constexpr int kMinExp = -1074;
constexpr int kGroup = 4;         // tiles per assembly workgroup

__device__ __forceinline__ int find_grid(const GridDesc* grids, int ngrids, int tile) {
    int g = 0;
    while (g + 1 < ngrids && tile >= grids[g + 1].tile_begin) g++;
    return g;
}

// one wavefront codes the 64 blocks of a tile into their scratch slots (word i of
// lane j's slot at scratch[(tile * kSlotWords + i) * 64 + j], so a wavefront's stores
// of the same word index are contiguous), then scans their lengths
__global__ void __launch_bounds__(kTile) encode_kernel(const GridDesc* __restrict__ grids, int ngrids,
                                                       unsigned long long* __restrict__ scratch,
                                                       uint32_t* __restrict__ len, uint32_t* __restrict__ lpre,
                                                       uint32_t* __restrict__ tsum) {
    const int tile = blockIdx.x;
    const int lane = threadIdx.x;
    const GridDesc& G = grids[find_grid(grids, ngrids, tile)];
    const int b = (tile - G.tile_begin) * kTile + lane;
    uint32_t bits = 0;
    if (b < G.nblocks) bits = encode_block(G, b, scratch + (size_t)tile * kSlotWords * kTile + lane, kTile);
    uint32_t inc = bits;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    const size_t slot = (size_t)tile * kTile + lane;
    len[slot] = bits;
    lpre[slot] = inc - bits;
    if (lane == 63) tsum[tile] = inc;
}

// workgroup = kGroup tiles of one grid (blockIdx.y): the bit offset of its first tile
// is the sum of the grid's earlier tiles; then every block writes the output words
// whose first bit it holds (block 0 also the header's), each word once
__global__ void __launch_bounds__(64 * kGroup) assemble_kernel(const GridDesc* __restrict__ grids,
                                                               const unsigned long long* __restrict__ scratch,
                                                               const uint32_t* __restrict__ len,
                                                               const uint32_t* __restrict__ lpre,
                                                               const uint32_t* __restrict__ tsum,
                                                               unsigned long long* __restrict__ gbits,
                                                               unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[kGroup];
    __shared__ unsigned long long wave_base[kGroup];
    const GridDesc& G = grids[blockIdx.y];
    const int t0 = (int)blockIdx.x * kGroup;
    if (t0 >= G.ntiles) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long s = 0;
    for (int i = threadIdx.x; i < t0; i += blockDim.x) s += tsum[G.tile_begin + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) part[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long base = part[0] + part[1] + part[2] + part[3];
        for (int w = 0; w < kGroup; w++) {
            wave_base[w] = base;
            if (t0 + w < G.ntiles) base += tsum[G.tile_begin + t0 + w];
        }
        if (t0 + kGroup >= G.ntiles) gbits[blockIdx.y] = (unsigned long long)G.header_bits + base;
    }
    __syncthreads();
    const int t = t0 + wave;
    if (t >= G.ntiles) return;
    const int b = t * kTile + lane;
    if (b >= G.nblocks) return;
    const size_t slot = (size_t)(G.tile_begin + t) * kTile + lane;
    const unsigned long long H = (unsigned long long)G.header_bits;
    const unsigned long long start = H + wave_base[wave] + lpre[slot];
    const unsigned long long end = start + len[slot];
    if (b == 0) {
        for (unsigned long long w = 0; (w << 6) < H && w < G.cap_words; w++)
            out[G.out_word + w] = gather_word(G, scratch, len, w << 6, 0, 0, H);
    }
    for (unsigned long long w = (start + 63) >> 6; (w << 6) < end && w < G.cap_words; w++)
        out[G.out_word + w] = gather_word(G, scratch, len, w << 6, t, lane, start);
}

// ---- host side

struct Workspace {
    unsigned long long* scratch = nullptr;
    size_t scratch_cap = 0;
    uint32_t* len = nullptr;  // per slot
    size_t len_cap = 0;
    uint32_t* lpre = nullptr;  // per slot
    size_t lpre_cap = 0;
    uint32_t* tsum = nullptr;  // per tile
    size_t tsum_cap = 0;
    unsigned long long* gbits = nullptr;  // per grid
    size_t gbits_cap = 0;
    GridDesc* grids = nullptr;
    size_t grids_cap = 0;
    float* in = nullptr;  // staging for host input
    size_t in_cap = 0;
    unsigned long long* out = nullptr;  // staging for host output
    size_t out_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

Workspace& ws() {
    static Workspace w;
    return w;
}

template <class T>
bool grow(T*& p, size_t& have, size_t need, const char* what) {
    if (need <= have && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    have = 0;
    const size_t n = std::max<size_t>(need, 1);
    if (!hip_ok(hipMalloc((void**)&p, n * sizeof(T)), what)) return false;
    have = n;
    return true;
}

long long blocks_of(int nx, int ny, int nz) {
    return (long long)((nx + 3) / 4) * ((ny + 3) / 4) * ((nz + 3) / 4);
}

// zfp/src/zfp.c:744-784 for a 3D float field in fixed-accuracy mode: header
// (at most 148 bits) plus 2120 bits per block, in whole 64-bit words
long long max_bytes(int nx, int ny, int nz) {
    const unsigned long long bits = 148ull + (unsigned long long)blocks_of(nx, ny, nz) * (unsigned long long)kMaxBlockBits;
    return (long long)(((bits + 63ull) / 64ull) * 8ull);
}

// zfp/src/zfp.c:842-857
int accuracy_minexp(double tolerance) {
    int e = 0;
    std::frexp(tolerance, &e);
    return e - 1;
}

struct BitWriter {
    unsigned long long w[3] = {0, 0, 0};
    int pos = 0;
    void put(unsigned long long v, int n) {
        for (int i = 0; i < n; i++, pos++)
            if ((v >> i) & 1ull) w[pos >> 6] |= 1ull << (pos & 63);
    }
};

// zfp/src/zfp.c:1228-1262 (magic, 52-bit metadata, then the mode), with the mode of
// zfp/src/zfp.c:667-734: the 12-bit form unless minexp is outside it or all
// parameters equal zfp's defaults (tolerance below 2^-1073), then 64 bits
void header_of(GridDesc& g, int minexp) {
    BitWriter b;
    b.put('z', 8);
    b.put('f', 8);
    b.put('p', 8);
    b.put(5, 8);  // zfp codec version (zfp/include/zfp/version.h:14)
    unsigned long long meta = (unsigned long long)(g.nz - 1);
    meta = (meta << 16) + (unsigned long long)(g.ny - 1);
    meta = (meta << 16) + (unsigned long long)(g.nx - 1);
    meta = (meta << 2) + 2ull;  // three dimensions
    meta = (meta << 2) + 2ull;  // float
    b.put(meta, 52);
    if (minexp != kMinExp && minexp <= 843) {
        b.put((unsigned long long)(minexp - kMinExp) + 2177ull, 12);
    } else {
        const long long me = std::min(std::max((long long)minexp + 16495, 0ll), 0x7fffll);
        unsigned long long m = (unsigned long long)me;
        m = (m << 7) + 63ull;      // maxprec - 1
        m = (m << 15) + 16657ull;  // maxbits - 1
        m = (m << 15) + 0ull;      // minbits - 1
        m = (m << 12) + 0xfffull;
        b.put(m, 64);
    }
    g.header[0] = b.w[0];
    g.header[1] = b.w[1];
    g.header[2] = b.w[2];
    g.header_bits = b.pos;
}

struct GridRequest {
    const float* values;  // device
    int nx, ny, nz;
    double tolerance;
    unsigned long long out_word, cap_words;
};

bool check_shape(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0; }

// validates one grid the way the reference codec does (its error strings)
bool validate(const void* values, int nx, int ny, int nz, double tolerance) {
    if (!values || !check_shape(nx, ny, nz)) {
        set_error("invalid_input");
        return false;
    }
    if (!(tolerance > 0.0) || !std::isfinite(tolerance)) {
        set_error("invalid_tolerance");
        return false;
    }
    if (nx > 65536 || ny > 65536 || nz > 65536) {  // the header holds 16 bits per extent
        set_error("zfp_write_header_failed");
        return false;
    }
    return true;
}

// codes the grids into `out` (device words); bits[i] = stream length in bits.
// `timed_repeats` > 0 runs the three launches that many times between events.
bool run(const std::vector<GridRequest>& req, unsigned long long* out, std::vector<unsigned long long>& bits,
         int timed_repeats = 0, float* ms = nullptr) {
    Workspace& W = ws();
    const int n = (int)req.size();
    std::vector<GridDesc> g(n);
    int tiles = 0;
    for (int i = 0; i < n; i++) {
        GridDesc& d = g[i];
        d.values = req[i].values;
        d.nx = req[i].nx;
        d.ny = req[i].ny;
        d.nz = req[i].nz;
        d.bx = (d.nx + 3) / 4;
        d.by = (d.ny + 3) / 4;
        const long long nb = blocks_of(d.nx, d.ny, d.nz);
        if (nb > (1ll << 30)) {
            set_error("analysis codec: grid too large");
            return false;
        }
        d.nblocks = (int)nb;
        d.ntiles = (int)((nb + kTile - 1) / kTile);
        d.tile_begin = tiles;
        tiles += d.ntiles;
        d.minexp = accuracy_minexp(req[i].tolerance);
        d.out_word = req[i].out_word;
        d.cap_words = req[i].cap_words;
        header_of(d, d.minexp);
    }
    const size_t slots = (size_t)tiles * kTile;
    if (!grow(W.scratch, W.scratch_cap, slots * kSlotWords, "hipMalloc(codec scratch)") ||
        !grow(W.len, W.len_cap, slots, "hipMalloc(codec lengths)") ||
        !grow(W.lpre, W.lpre_cap, slots, "hipMalloc(codec prefixes)") ||
        !grow(W.tsum, W.tsum_cap, (size_t)tiles, "hipMalloc(codec tile sums)") ||
        !grow(W.gbits, W.gbits_cap, (size_t)kMaxGrids, "hipMalloc(codec totals)") ||
        !grow(W.grids, W.grids_cap, (size_t)kMaxGrids, "hipMalloc(codec grids)"))
        return false;
    hipStream_t st = rt().stream;
    if (!hip_ok(hipMemcpyAsync(W.grids, g.data(), n * sizeof(GridDesc), hipMemcpyHostToDevice, st), "hipMemcpyAsync(codec grids)"))
        return false;
    int max_groups = 1;
    for (const GridDesc& d : g) max_groups = std::max(max_groups, (d.ntiles + kGroup - 1) / kGroup);
    auto launch = [&]() {
        encode_kernel<<<tiles, kTile, 0, st>>>(W.grids, n, W.scratch, W.len, W.lpre, W.tsum);
        assemble_kernel<<<dim3(max_groups, n), 64 * kGroup, 0, st>>>(W.grids, W.scratch, W.len, W.lpre, W.tsum, W.gbits, out);
    };
    if (timed_repeats > 0) {
        if (!W.ev0 && (!hip_ok(hipEventCreate(&W.ev0), "hipEventCreate") || !hip_ok(hipEventCreate(&W.ev1), "hipEventCreate")))
            return false;
        launch();  // warm
        if (!hip_ok(hipEventRecord(W.ev0, st), "hipEventRecord")) return false;
        for (int r = 0; r < timed_repeats; r++) launch();
        if (!hip_ok(hipEventRecord(W.ev1, st), "hipEventRecord")) return false;
    } else {
        launch();
    }
    if (!hip_ok(hipGetLastError(), "codec launch")) return false;
    bits.assign(n, 0ull);
    if (!hip_ok(hipMemcpyAsync(bits.data(), W.gbits, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
                "hipMemcpyAsync(codec totals)"))
        return false;
    if (!hip_ok(hipStreamSynchronize(st), "codec synchronize")) return false;
    if (timed_repeats > 0 && ms) {
        float t = 0.f;
        if (!hip_ok(hipEventElapsedTime(&t, W.ev0, W.ev1), "hipEventElapsedTime")) return false;
        *ms = t / (float)timed_repeats;
    }
    for (int i = 0; i < n; i++) {
        if ((bits[i] + 63ull) / 64ull > g[i].cap_words) {
            set_error("analysis codec: output capacity too small for grid " + std::to_string(i));
            return false;
        }
    }
    return true;
}

// ---- host decoder

struct BitReader {
    const uint8_t* p;
    unsigned long long nbits, pos;
    bool overrun = false;
    BitReader(const uint8_t* d, size_t size) : p(d), nbits((unsigned long long)size * 8ull), pos(0) {}
    unsigned long long word_at(unsigned long long wi) const {  // 64 bits from byte 8 wi, zero past the end
        const unsigned long long b = wi * 8ull, nbytes = nbits / 8ull;
        unsigned long long v = 0;
        if (b + 8 <= nbytes) {
            std::memcpy(&v, p + b, 8);
        } else {
            for (unsigned long long i = 0; b + i < nbytes && i < 8; i++) v |= (unsigned long long)p[b + i] << (8 * i);
        }
        return v;
    }
    unsigned long long peek64() const {
        const unsigned long long i = pos >> 6;
        const unsigned sh = (unsigned)(pos & 63ull);
        unsigned long long v = word_at(i) >> sh;
        if (sh) v |= word_at(i + 1) << (64 - sh);
        return v;
    }
    unsigned long long get(unsigned n) {  // n <= 64
        if (!n) return 0;
        unsigned long long v = peek64();
        if (n < 64) v &= (1ull << n) - 1ull;
        skip(n);
        return v;
    }
    void skip(unsigned long long n) {
        pos += n;
        if (pos > nbits) overrun = true;
    }
};

struct Params {
    unsigned minbits, maxbits, maxprec;
    int minexp;
};

// zfp/src/zfp.c:859-905
bool params_of_mode(unsigned long long mode, Params& p) {
    if (mode <= 4094ull) {
        if (mode < 2048ull) {
            p = {(unsigned)mode + 1u, (unsigned)mode + 1u, 64u, kMinExp};
        } else if (mode < 2048ull + 128ull) {
            p = {1u, 16658u, (unsigned)mode + 1u - 2048u, kMinExp};
        } else if (mode == 2048ull + 128ull) {
            p = {1u, 16658u, 64u, kMinExp - 1};
        } else {
            p = {1u, 16658u, 64u, (int)mode + kMinExp - 2177};
        }
    } else {
        mode >>= 12; p.minbits = (unsigned)(mode & 0x7fffull) + 1u;
        mode >>= 15; p.maxbits = (unsigned)(mode & 0x7fffull) + 1u;
        mode >>= 15; p.maxprec = (unsigned)(mode & 0x7full) + 1u;
        mode >>= 7;  p.minexp = (int)(mode & 0x7fffull) - 16495;
    }
    return p.minbits <= p.maxbits && p.maxprec > 0 && p.maxprec <= 64;
}

// zfp/src/template/decode.c:80-220: bit planes from the most significant, with the
// group tests' zero runs counted 64 bits at a time; returns the bits read
unsigned decode_planes(BitReader& r, unsigned maxbits, unsigned maxprec, uint32_t* u) {
    const unsigned kmin = maxprec < 32 ? 32 - maxprec : 0;
    unsigned bits = maxbits, n = 0;
    std::memset(u, 0, 64 * sizeof(uint32_t));
    for (unsigned k = 32; bits && k-- > kmin;) {
        const unsigned m = std::min(n, bits);
        bits -= m;
        unsigned long long x = r.get(m);
        for (; bits && n < 64; n++) {
            bits--;
            if (!r.get(1)) break;  // negative group test
            const unsigned limit = std::min(63u - n, bits);
            const unsigned long long pk = r.peek64();
            const unsigned z = pk ? (unsigned)__builtin_ctzll(pk) : 64u;
            if (z < limit) {  // z zeros, then the one
                r.skip(z + 1);
                bits -= z + 1;
                n += z;
            } else {  // the run reaches the last coefficient or the budget
                r.skip(limit);
                bits -= limit;
                n += limit;
            }
            x += 1ull << n;
        }
        for (unsigned i = 0; x; i++, x >>= 1) u[i] += (uint32_t)(x & 1ull) << k;
    }
    return maxbits - bits;
}

void inv_lift(int32_t* p, int s) {  // zfp/src/template/decode.c:9-45
    auto add = [](int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); };
    auto sub = [](int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); };
    int32_t x = p[0], y = p[s], z = p[2 * s], w = p[3 * s];
    y = add(y, w >> 1); w = sub(w, y >> 1);
    y = add(y, w); w = sub(w, sub(y, w));
    z = add(z, x); x = sub(x, sub(z, x));
    y = add(y, z); z = sub(z, sub(y, z));
    w = add(w, x); x = sub(x, sub(w, x));
    p[0] = x; p[s] = y; p[2 * s] = z; p[3 * s] = w;
}

// zfp/src/template/decodef.c:7-36 with the inverse transform of
// zfp/src/template/decode3.c:27-43 (z, then y, then x) and codecf.c:22-32
void decode_block(BitReader& r, const Params& P, float* v) {
    unsigned bits = 1;
    if (!r.get(1)) {
        for (int i = 0; i < 64; i++) v[i] = 0.0f;
        if (P.minbits > bits) r.skip(P.minbits - bits);
        return;
    }
    bits += 8;
    const int emax = (int)r.get(8) - 127;
    int pr = emax - P.minexp + 8;
    pr = pr < 0 ? 0 : pr;
    const unsigned maxprec = std::min((unsigned)pr, P.maxprec);
    const unsigned minbits = P.minbits - std::min(bits, P.minbits);
    uint32_t u[64];
    const unsigned used = decode_planes(r, P.maxbits - bits, maxprec, u);
    if (used < minbits) r.skip(minbits - used);
    int32_t q[64];
    for (int i = 0; i < 64; i++) q[perm3(i)] = (int32_t)((u[i] ^ 0xaaaaaaaau) - 0xaaaaaaaau);
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) inv_lift(q + x + 4 * y, 16);
    for (int x = 0; x < 4; x++)
        for (int z = 0; z < 4; z++) inv_lift(q + 16 * z + x, 4);
    for (int z = 0; z < 4; z++)
        for (int y = 0; y < 4; y++) inv_lift(q + 4 * y + 16 * z, 1);
    const float s = std::ldexp(1.0f, emax - 30);
    for (int i = 0; i < 64; i++) v[i] = s * (float)q[i];
}

bool decompress(const uint8_t* data, size_t size, int nx, int ny, int nz, float* out) {
    if (!data || size == 0 || !out || !check_shape(nx, ny, nz)) {
        set_error("invalid_input");
        return false;
    }
    BitReader r(data, size);
    // zfp/src/zfp.c:1265-1297
    if (r.get(8) != 'z' || r.get(8) != 'f' || r.get(8) != 'p' || r.get(8) != 5 || r.overrun) {
        set_error("zfp_read_header_failed");
        return false;
    }
    unsigned long long meta = r.get(52);
    const unsigned type = (unsigned)(meta & 3ull) + 1u;
    meta >>= 2;
    const unsigned dims = (unsigned)(meta & 3ull) + 1u;
    meta >>= 2;
    unsigned long long mode = r.get(12);
    if (mode > 4094ull) mode += r.get(52) << 12;
    Params P{};
    if (r.overrun || !params_of_mode(mode, P)) {
        set_error("zfp_read_header_failed");
        return false;
    }
    if (type != 3u || dims != 3u || (int)(meta & 0xffffull) + 1 != nx || (int)((meta >> 16) & 0xffffull) + 1 != ny ||
        (int)((meta >> 32) & 0xffffull) + 1 != nz) {
        set_error("zfp_header_mismatch");
        return false;
    }
    if (P.minexp < kMinExp) {
        set_error("analysis codec: zfp reversible mode is not supported");
        return false;
    }
    float q[64];
    const size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    for (int z = 0; z < nz; z += 4)
        for (int y = 0; y < ny; y += 4)
            for (int x = 0; x < nx; x += 4) {
                decode_block(r, P, q);
                const int bx = std::min(4, nx - x), by = std::min(4, ny - y), bz = std::min(4, nz - z);
                for (int k = 0; k < bz; k++)
                    for (int j = 0; j < by; j++)
                        for (int i = 0; i < bx; i++) out[(size_t)(x + i) + sy * (y + j) + sz * (z + k)] = q[16 * k + 4 * j + i];
            }
    if (r.overrun) {  // the reference would read past the buffer here
        set_error("zfp_decompress_failed");
        return false;
    }
    return true;
}

}  // namespace codec
}  // namespace aero

using namespace aero;

extern "C" {

AERO_LBM_CAPI_EXPORT long long aero_lbm_analysis_max_compressed_size_3d(int nx, int ny, int nz) {
    if (!codec::check_shape(nx, ny, nz)) return 0;
    return codec::max_bytes(nx, ny, nz);
}

AERO_LBM_CAPI_EXPORT long long aero_lbm_analysis_compress_float_grid_3d(const float* values, int nx, int ny, int nz,
                                                                        double tolerance, uint8_t* out,
                                                                        long long capacity) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!codec::validate(values, nx, ny, nz, tolerance)) return 0;
    if (!out || capacity <= 0) {
        set_error("invalid_input");
        return 0;
    }
    if (!ensure_device()) return 0;
    codec::Workspace& W = codec::ws();
    const size_t cells = (size_t)nx * (size_t)ny * (size_t)nz;
    const unsigned long long cap_words = (unsigned long long)codec::max_bytes(nx, ny, nz) / 8ull;
    if (!codec::grow(W.in, W.in_cap, cells, "hipMalloc(codec input)")) return 0;
    if (!codec::grow(W.out, W.out_cap, (size_t)cap_words, "hipMalloc(codec output)")) return 0;
    hipStream_t st = rt().stream;
    if (!hip_ok(hipMemcpyAsync(W.in, values, cells * sizeof(float), hipMemcpyHostToDevice, st), "hipMemcpyAsync(codec input)"))
        return 0;
    std::vector<codec::GridRequest> req{{W.in, nx, ny, nz, tolerance, 0ull, cap_words}};
    std::vector<unsigned long long> bits;
    if (!codec::run(req, W.out, bits)) return 0;
    const long long bytes = (long long)(((bits[0] + 63ull) / 64ull) * 8ull);
    if (bytes > capacity) {
        set_error("analysis codec: output buffer too small (" + std::to_string(bytes) + " bytes needed)");
        return 0;
    }
    if (!hip_ok(hipMemcpyAsync(out, W.out, (size_t)bytes, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(codec output)"))
        return 0;
    if (!hip_ok(hipStreamSynchronize(st), "codec synchronize")) return 0;
    return bytes;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_analysis_compress_float_grids_3d_device(
    int count, const float* const* values, const int* dims, const double* tolerances, uint8_t* out,
    const long long* out_offsets, const long long* out_capacities, long long* out_sizes) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (count <= 0 || count > codec::kMaxGrids || !values || !dims || !tolerances || !out || !out_offsets ||
        !out_capacities || !out_sizes || (((uintptr_t)out) & 7u)) {
        set_error("invalid_input");
        return 0;
    }
    std::vector<codec::GridRequest> req(count);
    for (int i = 0; i < count; i++) {
        if (!codec::validate(values[i], dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], tolerances[i])) return 0;
        if (out_offsets[i] < 0 || (out_offsets[i] & 7) || out_capacities[i] < 8) {
            set_error("invalid_input");
            return 0;
        }
        req[i] = {values[i], dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], tolerances[i],
                  (unsigned long long)out_offsets[i] / 8ull, (unsigned long long)out_capacities[i] / 8ull};
    }
    if (!ensure_device()) return 0;
    std::vector<unsigned long long> bits;
    if (!codec::run(req, reinterpret_cast<unsigned long long*>(out), bits)) return 0;
    for (int i = 0; i < count; i++) out_sizes[i] = (long long)(((bits[i] + 63ull) / 64ull) * 8ull);
    return 1;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_analysis_compress_grids_timed(
    int count, const float* const* values, const int* dims, const double* tolerances, uint8_t* out,
    const long long* out_offsets, const long long* out_capacities, int repeats, float* out_ms) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (count <= 0 || count > codec::kMaxGrids || !values || !dims || !tolerances || !out || !out_offsets ||
        !out_capacities || repeats <= 0 || !out_ms || (((uintptr_t)out) & 7u)) {
        set_error("invalid_input");
        return 0;
    }
    std::vector<codec::GridRequest> req(count);
    for (int i = 0; i < count; i++) {
        if (!codec::validate(values[i], dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], tolerances[i])) return 0;
        if (out_offsets[i] < 0 || (out_offsets[i] & 7) || out_capacities[i] < 8) {
            set_error("invalid_input");
            return 0;
        }
        req[i] = {values[i], dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], tolerances[i],
                  (unsigned long long)out_offsets[i] / 8ull, (unsigned long long)out_capacities[i] / 8ull};
    }
    if (!ensure_device()) return 0;
    std::vector<unsigned long long> bits;
    return codec::run(req, reinterpret_cast<unsigned long long*>(out), bits, repeats, out_ms) ? 1 : 0;
}

AERO_LBM_CAPI_EXPORT int aero_lbm_analysis_decompress_float_grid_3d(const uint8_t* data, long long size, int nx,
                                                                    int ny, int nz, float* out) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (size <= 0) {
        set_error("invalid_input");
        return 0;
    }
    return codec::decompress(data, (size_t)size, nx, ny, nz, out) ? 1 : 0;
}

}  // extern "C"
