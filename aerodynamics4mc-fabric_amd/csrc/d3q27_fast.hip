// Fast D3Q27 kernels: the same cumulant + SGS + Boussinesq step as the exact
// kernel, organised for throughput on CDNA4.
//
//  * FP32 A-B ("plain"): pull from the previous post-collision buffer, write the
//    next one. 27 x 4 B read + 27 x 4 B written per fluid cell.
//  * FP16 / BF16 (and FP32) A-A in place, equilibrium-shifted for 16-bit: the
//    stored value is g_q = f_q - w_q. Even steps read and write a cell's own 27
//    slots (the collided value of direction q goes to slot opp(q)); odd steps
//    pull the 27 values from the neighbours' opposite slots and push the results
//    to the neighbours (ref: native/src/aero_lbm_jni.cpp:4558-4723 for the A-A
//    indexing and parity flip). 27 x 2 B read + 27 x 2 B written per fluid cell.
//
// Launch shape: 64 (z) x 4 (y) threads per workgroup, one thread per cell, so a
// wavefront reads 64 consecutive cells of a plane. Workgroups are numbered so
// that each of the 8 XCDs (round-robin dispatch) owns one contiguous eighth of
// the brick: neighbouring tiles, whose cache lines overlap by the +-1 shifts of
// the streaming, then share an L2.
//
// Per-cell link mask (built once per payload): bit q set when the pull source
// x - c_q is solid or outside the domain. Waves whose 64 cells have no such link
// take a path with no per-link selects; the face-rule path runs only for cells
// at faces whose rule is not plain bounce-back, and the general temperature
// stencil only next to Dirichlet / Neumann thermal faces.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "fast_core.h"

namespace aero {
namespace fast {

constexpr int kTileZ = 64;  // one wavefront = 64 consecutive z cells of one (x, y) row
constexpr int kTileY = 4;
static_assert(kTileZ == 64, "the step kernel treats x and y as wave-uniform");

struct TileCoord {
    int x, y, z;       // x and y are wave-uniform (scalar registers)
    bool row_valid;    // wave-uniform
    bool valid;
};

// Tile layout: the x planes flagged xface_bulk (whose rows run face rules, longer
// than a bulk row) are dealt out first, an equal share to each XCD; the other
// planes follow, each XCD taking one contiguous range of them. Workgroups are
// dispatched round-robin over the 8 XCDs, so workgroup L runs on XCD L & 7 as its
// (L >> 3)-th tile.
struct TileLayout {
    int nbp;             // tiles per x plane
    int faces;           // flagged x planes (0, 1 or 2)
    int fper, bper;      // face / bulk tiles per XCD
};

__host__ __device__ __forceinline__ TileLayout tile_layout(const StepParams& P) {
    TileLayout T;
    T.nbp = ((P.nz + kTileZ - 1) / kTileZ) * ((P.ny + kTileY - 1) / kTileY);
    const int lo = P.xface_bulk[0] ? 1 : 0;
    const int hi = (P.xface_bulk[1] && !(P.nx == 1 && lo)) ? 1 : 0;  // one plane counts once
    T.faces = lo + hi;
    T.fper = (T.faces * T.nbp + 7) >> 3;
    T.bper = ((P.nx - T.faces) * T.nbp + 7) >> 3;
    return T;
}

__device__ __forceinline__ TileCoord tile_cell(const StepParams& P, int L) {
    const int nbz = (P.nz + kTileZ - 1) / kTileZ;
    const TileLayout T = tile_layout(P);
    const int xcd = L & 7, i = L >> 3;
    int plane, tile;
    bool ok;
    if (i < T.fper) {
        const int f = xcd * T.fper + i;
        ok = f < T.faces * T.nbp;
        plane = f / T.nbp;
        tile = f - plane * T.nbp;
        plane = (plane == 0 && P.xface_bulk[0]) ? 0 : P.nx - 1;
    } else {
        const int b = xcd * T.bper + (i - T.fper);
        ok = b < (P.nx - T.faces) * T.nbp;
        plane = b / T.nbp;
        tile = b - plane * T.nbp;
        plane += P.xface_bulk[0] ? 1 : 0;
    }
    const int by = tile / nbz;
    const int bz = tile - by * nbz;
    TileCoord t;
    t.x = plane;
    // a wavefront holds threads 64w..64w+63 of the 64 x 4 workgroup: one threadIdx.y
    t.y = by * kTileY + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    t.z = bz * kTileZ + (int)threadIdx.x;
    t.row_valid = ok && t.y < P.ny;
    t.valid = t.row_valid && t.z < P.nz;
    return t;
}

// The tile L lies in an x face plane flagged xface_bulk (xface_tile), or holds the
// first or last row of y of a y face flagged yface_bulk (yface_tile): its workgroup
// takes the instantiation of the step body with the face rules, every other
// workgroup the one without them. Kernels launched for contexts with such y faces
// are separate instantiations (template YF), so the others keep their code.
__device__ __forceinline__ bool xface_tile(const StepParams& P, int L) {
    return (L >> 3) < tile_layout(P).fper;
}
__device__ __forceinline__ bool yface_tile(const StepParams& P, int L) {
    if (!(P.yface_bulk[0] | P.yface_bulk[1])) return false;
    const TileLayout T = tile_layout(P);
    const int nbz = (P.nz + kTileZ - 1) / kTileZ, nby = (P.ny + kTileY - 1) / kTileY;
    const int b = (L & 7) * T.bper + ((L >> 3) - T.fper);
    const int by = (b % T.nbp) / nbz;
    return (P.yface_bulk[0] && by == 0) || (P.yface_bulk[1] && by == nby - 1);
}

// Wave-uniform neighbour rows of the wave's (x, y) row: element offset of row
// (x + dx, y + dy) relative to row (x, y) after periodic wrap, and whether that
// row is inside the domain.
struct RowNbr {
    int row;  // element index of cell (x, y, 0)
    int xo[3], yo[3];
    int okbits;  // bit (dx + 1) * 3 + (dy + 1): row (x + dx, y + dy) is inside the domain
    __device__ __forceinline__ int off(int dx, int dy) const { return row + xo[dx + 1] + yo[dy + 1]; }
    __device__ __forceinline__ bool ok(int dx, int dy) const { return (okbits >> ((dx + 1) * 3 + dy + 1)) & 1; }
};

__device__ __forceinline__ RowNbr row_nbr(const StepParams& P, int x, int y) {
    RowNbr R;
    const int yz = P.ny * P.nz;
    R.row = (x * P.ny + y) * P.nz;
    R.xo[1] = R.yo[1] = 0;
    R.xo[0] = x > 0 ? -yz : (P.hydro_wrap_lo[0] ? (P.nx - 1) * yz : 0);
    R.xo[2] = x < P.nx - 1 ? yz : (P.hydro_wrap_hi[0] ? -(P.nx - 1) * yz : 0);
    R.yo[0] = y > 0 ? -P.nz : (P.hydro_wrap_lo[1] ? (P.ny - 1) * P.nz : 0);
    R.yo[2] = y < P.ny - 1 ? P.nz : (P.hydro_wrap_hi[1] ? -(P.ny - 1) * P.nz : 0);
    const int xok = 2 | ((x > 0 || P.hydro_wrap_lo[0]) ? 1 : 0) | ((x < P.nx - 1 || P.hydro_wrap_hi[0]) ? 4 : 0);
    const int yok = 2 | ((y > 0 || P.hydro_wrap_lo[1]) ? 1 : 0) | ((y < P.ny - 1 || P.hydro_wrap_hi[1]) ? 4 : 0);
    R.okbits = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if ((xok >> i) & 1) R.okbits |= yok << (3 * i);
    return R;
}

// An opaque copy of a wave-uniform value: derived values are recomputed from it
// instead of being kept live in scalar registers across the collision.
__device__ __forceinline__ int fresh_s(int v) {
    asm volatile("" : "+s"(v));
    return v;
}

// Per-lane byte offsets of z - 1, z, z + 1 inside a row (periodic wrap); a
// neighbour outside the domain gets kDropByte.
template <class T>
__device__ __forceinline__ void z_bytes(const StepParams& P, int z, int zb[3]) {
    constexpr int s = (int)sizeof(T);
    zb[1] = z * s;
    zb[0] = z > 0 ? (z - 1) * s : (P.hydro_wrap_lo[2] ? (P.nz - 1) * s : kDropByte);
    zb[2] = z < P.nz - 1 ? (z + 1) * s : (P.hydro_wrap_hi[2] ? 0 : kDropByte);
}

// wave-uniform byte offset of row `row_off` of population plane q inside the
// population array (below 2 GiB)
template <class T>
__device__ __forceinline__ int plane_row(int N, int q, int row_off) {
    return (q * N + row_off) * (int)sizeof(T);
}

__host__ inline int tile_grid(const StepParams& P) {
    const TileLayout T = tile_layout(P);
    return 8 * (T.fper + T.bper);
}

// link whose pull source is the neighbour at x + d
__device__ constexpr int link_from(int dx, int dy, int dz) { return ((1 - dx) * 3 + (1 - dy)) * 3 + (1 - dz); }

// face temperature of a stencil point outside the domain (ref: jni.cpp:2145-2188
// with the face rules of t_at): Dirichlet value, self + flux, or self
__device__ __forceinline__ float face_temperature(const FaceParams& fc, float self) {
    return fc.thermal_kind == 4 ? clampf(fc.temperature, -1.0f, 1.0f)
                                : (fc.thermal_kind == 5 ? clampf(self + fc.heat_flux, -1.0f, 1.0f) : self);
}

// first-order semi-Lagrangian advection + explicit diffusion + source
// (ref: native/src/aero_lbm_jni.cpp:2145-2188). Solid or out-of-domain stencil
// points take the cell's own value; the link mask says which they are. XFace: the
// cell lies in an x face plane flagged xface_bulk (xlo / xhi, wave-uniform), whose
// points outside the domain take that face's temperature rule.
template <bool XFace>
__device__ __forceinline__ float temperature_fast(const StepParams& P, const float* __restrict__ T,
                                                  const float* __restrict__ src, int cell, int x, int y, int z,
                                                  const Nbr& nb, uint32_t links, float ux, float uy, float uz,
                                                  bool xlo = false, bool xhi = false) {
    // T, the temperature field, is read as a buffer with 32-bit byte offsets
    const Rsrc rT = make_rsrc(T, P.cells * 4);
    const int cb = cell * 4;
    const float tc = bld_f32(rT, cb);
    float tlo = tc, thi = tc;
    if (XFace) {
        const FaceParams* fp = P.self_dev->faces_dev;
        if (xlo) tlo = face_temperature(fp[0], tc);
        if (xhi) thi = face_temperature(fp[1], tc);
    }
    const float px = (float)x - ux, py = (float)y - uy, pz = (float)z - uz;
    const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
    const float fx = px - fx0, fy = py - fy0, fz = pz - fz0;
    const bool bx = fx0 < (float)x, by = fy0 < (float)y, bz = fz0 < (float)z;  // box starts at -1 on that axis
    // The 2x2x2 box of the departure point starts at (ax, ay, az) in {-1, 0}^3 relative
    // to the cell. Byte offsets of its two layers along each axis (periodic wrap from
    // nb; a neighbour outside the domain maps to the cell, so every index stays in bounds).
    const int X0 = (bx ? nb.xm : 0) * 4, X1 = (bx ? 0 : nb.xp) * 4;
    const int Y0 = (by ? nb.ym : 0) * 4, Y1 = (by ? 0 : nb.yp) * 4;
    const int Z0 = (bz ? nb.zm : 0) * 4, Z1 = (bz ? 0 : nb.zp) * 4;
    // corner (i, j, k) is link link_from(ax + i, ay + j, az + k) = base - (9 i + 3 j + k),
    // base in [13, 26]: after the shift its bit sits at 13 - (9 i + 3 j + k)
    const int base = 26 - (bx ? 0 : 9) - (by ? 0 : 3) - (bz ? 0 : 1);
    const uint32_t sh = links >> (base - 13);
    // a solid or outside point reads the cell itself (the offset is masked to zero,
    // not the loaded value selected, so every load is unconditional)
    auto corner = [&](int i, int j, int k, int off) -> float {
        const float v = bld_f32(rT, cb + (off & ~bit_mask(sh, 13 - 9 * i - 3 * j - k)));
        if (XFace) {
            const bool below = bx && i == 0, above = !bx && i == 1;  // x - 1 / x + 1 layer
            return (xlo && below) ? tlo : ((xhi && above) ? thi : v);
        }
        return v;
    };
    const int o00 = X0 + Y0, o10 = X1 + Y0, o01 = X0 + Y1, o11 = X1 + Y1;
    const float c000 = corner(0, 0, 0, o00 + Z0), c100 = corner(1, 0, 0, o10 + Z0);
    const float c010 = corner(0, 1, 0, o01 + Z0), c110 = corner(1, 1, 0, o11 + Z0);
    const float c001 = corner(0, 0, 1, o00 + Z1), c101 = corner(1, 0, 1, o10 + Z1);
    const float c011 = corner(0, 1, 1, o01 + Z1), c111 = corner(1, 1, 1, o11 + Z1);
    const float c00 = fmaf(c100 - c000, fx, c000), c10 = fmaf(c110 - c010, fx, c010);
    const float c01 = fmaf(c101 - c001, fx, c001), c11 = fmaf(c111 - c011, fx, c011);
    const float c0 = fmaf(c10 - c00, fy, c00), c1 = fmaf(c11 - c01, fy, c01);
    const float adv = fmaf(c1 - c0, fz, c0);
    // the six face neighbours (links 4, 22, 10, 16, 12, 14 pull from +x, -x, +y, -y, +z, -z)
    auto face = [&](int bit, int off, int dx) -> float {
        const float v = bld_f32(rT, cb + ((off * 4) & ~bit_mask(links, bit)));
        if (XFace) return (xlo && dx < 0) ? tlo : ((xhi && dx > 0) ? thi : v);
        return v;
    };
    const float tsum = (face(4, nb.xp, 1) + face(22, nb.xm, -1)) + (face(10, nb.yp, 0) + face(16, nb.ym, 0)) +
                       (face(12, nb.zp, 0) + face(14, nb.zm, 0));
    const float lap = fmaf(-6.0f, tc, tsum);
    const float source =
        (P.source_on && P.in_ch > 9) ? clampf(bld_f32(make_rsrc(src, P.cells * 4), cb), -0.006f, 0.006f) : 0.0f;
    return clampf(adv + fmaf(P.alpha, lap, source), -1.0f, 1.0f);
}

// A-A face store: this step's post-collision values of the cells in the first
// three layers of faces whose rule reads the previous step (symmetry, convective)
__device__ __forceinline__ void face_store(const StepParams& P, int x, int y, int z, const float g[27],
                                           float* __restrict__ store_out) {
#pragma unroll
    for (int face = 0; face < 6; ++face) {
        if (!((P.face_store_faces >> face) & 1)) continue;
        const int axis = face >> 1;
        const int coord = axis == 0 ? x : axis == 1 ? y : z;
        const int n = axis == 0 ? P.nx : axis == 1 ? P.ny : P.nz;
        const int layer = (face & 1) ? n - 1 - coord : coord;
        if (layer > 2) continue;
        const int A = axis == 0 ? P.ny : P.nx;
        const int B = axis == 2 ? P.ny : P.nz;
        const int a = axis == 0 ? y : x;
        const int bb = axis == 2 ? y : z;
        const int cdir = (face & 1) ? 0 : 2;  // incoming through the min face: c = +1 (index 2)
        float* dst = store_out + P.face_store_off[face] + (size_t)(layer * 9) * A * B + (size_t)a * B + bb;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int q = axis == 0 ? cdir * 9 + j : axis == 1 ? (j / 3) * 9 + cdir * 3 + (j % 3) : j * 3 + cdir;
            dst[(size_t)j * A * B] = g[q];
        }
    }
}

// ---------------------------------------------------------------- bulk step kernel
// Mode: 0 = A-B pull, 1 = A-A even (local), 2 = A-A odd (pull + push).
// local_first: the state is the initial one (incoming values in own slots, no
// stream pending): collide in place.
//
// The x face planes flagged xface_bulk run their face rules here. Other cells at
// faces whose rule is not plain bounce-back or next to Dirichlet / Neumann thermal
// faces, cells the sponge touches, and all cells when the thermal periodicity
// differs from the hydrodynamic one carry kMaskRare and are left to the rare-cell
// kernels; for the rest every special link is halfway bounce-back.
// L: the workgroup's tile index within the brick. Shell: the context has a
// boundary-reference shell (a separate instantiation, so contexts without one keep
// the shorter kernel and its register budget).
template <class St, class Uv, int Mode, int StPol, bool Shell, bool XF, bool GravZ = false, bool YF = false>
__device__ __forceinline__ void bulk_body(const StepParams& P, const CellFields& C, const uint32_t* __restrict__ mask,
                                          const typename St::T* __restrict__ fin, typename St::T* fout,
                                          const float* __restrict__ t_in, float* __restrict__ t_out,
                                          typename Uv::T* __restrict__ u_state, const float* __restrict__ store_in,
                                          float* __restrict__ store_out, int local_first, int L) {
    const TileCoord tc = tile_cell(P, L);
    if (!tc.valid) return;
    const int x = tc.x, y = tc.y, z = tc.z;
    const RowNbr R = row_nbr(P, x, y);
    const int cell = R.row + z;
    const uint32_t m = mask[cell];
    // Solid cells take part in the population loads and stores of the local (A-A
    // even) and A-B steps without colliding: the even step swaps their slots q and
    // opp(q), the A-B step writes f_q <- f_opp(q), as the reference's solid cells keep
    // f_post = f and stream it back onto themselves. No fluid cell reads a solid
    // cell's slots, and the wave's stores then fill whole 64-byte sectors (a sector
    // with unwritten bytes costs the memory a read-modify-write). The A-A odd step
    // pushes into the neighbours, so solid cells stay out of it.
    const bool solid = (m & kMaskSolid) != 0;
    if ((m & kMaskRare) || (Mode == 2 && solid)) return;
    const size_t N = (size_t)P.cells;
    const uint32_t links = solid ? kMaskLinks : (m & kMaskLinks);
    using T = typename St::T;
    int zb[3];
    z_bytes<T>(P, z, zb);
    const T* const src = Mode == 0 ? fin : fout;
    const int Ni = P.plane;
    const Rsrc rin = make_rsrc(src, 27 * Ni * (int)sizeof(T));
    const Rsrc rout = make_rsrc(fout, 27 * Ni * (int)sizeof(T));
    const bool xlo = XF && x == 0 && P.xface_bulk[0], xhi = XF && x == P.nx - 1 && P.xface_bulk[1];
    const bool ylo = YF && XF && y == 0 && P.yface_bulk[0], yhi = YF && XF && y == P.ny - 1 && P.yface_bulk[1];
    // temperature first: its stencil loads and arithmetic finish before the 27
    // population loads, which keeps the two sets of values from being live together
    float temp = 0.0f;
    float ux0 = 0.0f, uy0 = 0.0f, uz0 = 0.0f;
    // the carried velocity as a buffer: 3 planes of N values
    const int ub = cell * (int)sizeof(typename Uv::T), us = (int)N * (int)sizeof(typename Uv::T);
    const Rsrc ru = make_rsrc(u_state, 3 * us);
    if (P.thermal_on && !solid) {
        ux0 = Uv::bld(ru, ub, 0);
        uy0 = Uv::bld(ru, ub, us);
        uz0 = Uv::bld(ru, ub, 2 * us);
        if (XF && (xlo || xhi))
            temp = temperature_fast<XF>(P, t_in, C.src, cell, x, y, z, make_nbr(P, x, y, z), links, ux0, uy0, uz0,
                                          xlo, xhi);
        else
            temp = temperature_fast<false>(P, t_in, C.src, cell, x, y, z, make_nbr(P, x, y, z), links, ux0, uy0,
                                           uz0);
    }
    float g[27];
    if (Mode == 1 || local_first) {
#pragma unroll
        for (int q = 0; q < 27; ++q) g[q] = St::template bld<kPolDefault>(rin, zb[1], plane_row<T>(Ni, q, R.row), q);
    } else {
        // Each link has two candidate sources: the neighbour's slot (streaming) and
        // the cell's own slot of the opposite direction (halfway bounce-back, when
        // the mask flags the link). Both lie in one buffer; the scalar offset is the
        // lower of the two and each lane picks its vector offset, so a bounced link
        // costs no second load. Which one is lower is known here: the two lie in
        // different planes (the rest link aside), and a row offset is smaller than
        // the plane stride, so the plane index decides. The pick is a bitfield
        // select on the link's bit spread over the word (0 or all ones).
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int dx = -lat_cx(q), dy = -lat_cy(q), dz = -lat_cz(q);
            const int qa = Mode == 0 ? q : lat_opp(q);  // streamed: plane at the neighbour
            const int qb = Mode == 0 ? lat_opp(q) : q;  // bounced: plane at the cell itself
            const int sb = plane_row<T>(Ni, qb, R.row);
            if (qa == qb) {  // rest link: the cell's own slot
                g[q] = St::template bld<kPolDefault>(rin, zb[1], sb, q);
                continue;
            }
            const int sa = plane_row<T>(Ni, qa, R.ok(dx, dy) ? R.off(dx, dy) : R.row);
            int so, va, vb;  // scalar offset; vector offsets of the streamed and the bounced value
            if (qa < qb) {
                so = sa;
                va = zb[dz + 1];
                vb = zb[1] + (sb - sa);
            } else {
                so = sb;
                va = zb[dz + 1] + (sa - sb);
                vb = zb[1];
            }
            g[q] = St::template bld<kPolDefault>(rin, pick_by_bit(links, q, va, vb), so, q);
        }
    }
    // x face planes flagged xface_bulk (inlet, outflow, symmetry, Dirichlet / Neumann
    // temperature): the 9 links pulled from outside the domain take the face rule,
    // with the loaded value (the cell's own opposite slot) as the bounce-back
    // candidate. x is wave-uniform, so only the waves of these planes branch here.
    if (!solid) {
        if (XF && !local_first && (xlo || xhi)) {
            const StepParams& Q = *P.self_dev;
            const PrevPost<St, Mode != 0> prev{fin, store_in, (size_t)Q.plane};
            const float rp = C.ref_p[cell];
            if (xlo) xface_links<0, St, Mode != 0>(Q, C.flags, rp, prev, x, y, z, g);
            if (xhi) xface_links<1, St, Mode != 0>(Q, C.flags, rp, prev, x, y, z, g);
        }
        if (YF && XF && !local_first && (ylo || yhi)) {
            const StepParams& Q = *P.self_dev;
            const PrevPost<St, Mode != 0> prev{fin, store_in, (size_t)Q.plane};
            if (ylo) yface_links<0, St, Mode != 0>(Q, C.flags, prev, x, y, z, g);
            if (yhi) yface_links<1, St, Mode != 0>(Q, C.flags, prev, x, y, z, g);
        }
        CollideResult r = collide<St::kShifted, GravZ>(P, g, temp);
        if (!r.ok) reseed<St::kShifted>(*P.self_dev, C, cell, g, r);
        if (Shell && P.shell_on) {
            const int layer = C.shell[cell];
            if (layer) shell_blend<St::kShifted>(P, C.shell_state, cell, layer, g, r);
        }
        if (P.thermal_on) {
            bst_f32(make_rsrc(t_out, P.cells * 4), cell * 4, temp);
            Uv::bst(ru, ub, 0, r.ux);
            Uv::bst(ru, ub, us, r.uy);
            Uv::bst(ru, ub, 2 * us, r.uz);
        }
        if (Mode != 0 && P.face_store_faces) face_store(*P.self_dev, x, y, z, g, store_out);
    }
    const RowNbr R2 = row_nbr(P, fresh_s(x), fresh_s(y));
    if (Mode != 2) {
#pragma unroll
        for (int q = 0; q < 27; ++q)
            St::template bst<StPol>(rout, zb[1], plane_row<T>(Ni, Mode == 0 ? q : lat_opp(q), R2.row), g[q], q);
    } else {
        // one store per direction: pushed to x + c_q, or, when that target is solid
        // or outside the domain (the mask's bit opp(q)), bounced back into the cell's
        // own slot opp(q). As for the pull, both candidates lie in one buffer: the
        // scalar offset is the lower of the two and each lane picks its vector offset.
        // A neighbour row outside the domain makes every lane of the wave bounce.
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int dx = lat_cx(q), dy = lat_cy(q), dz = lat_cz(q);
            const int sb = plane_row<T>(Ni, lat_opp(q), R2.row);
            if (!R2.ok(dx, dy)) {
                St::template bst<StPol>(rout, zb[1], sb, g[q], q);
                continue;
            }
            const int sn = plane_row<T>(Ni, q, R2.off(dx, dy));
            int so, vn, vb;
            if (q <= 13) {  // plane q lies below plane opp(q); a row offset is below the plane stride
                so = sn;
                vn = zb[dz + 1];
                vb = zb[1] + (sb - sn);
            } else {
                so = sb;
                vn = zb[dz + 1] + (sn - sb);
                vb = zb[1];
            }
            St::template bst<StPol>(rout, pick_by_bit(links, lat_opp(q), vn, vb), so, g[q], q);
        }
    }
}

template <class St, class Uv, int Mode, int StPol, bool Shell, bool GravZ, bool YF>
__global__ void __launch_bounds__(256) step_kernel(StepParams P, CellFields C, const uint32_t* __restrict__ mask,
                                                   const typename St::T* __restrict__ fin, typename St::T* fout,
                                                   const float* __restrict__ t_in, float* __restrict__ t_out,
                                                   typename Uv::T* __restrict__ u_state,
                                                   const float* __restrict__ store_in, float* __restrict__ store_out,
                                                   int local_first) {
    if (xface_tile(P, blockIdx.x) || (YF && yface_tile(P, blockIdx.x)))
        bulk_body<St, Uv, Mode, StPol, Shell, true, GravZ, YF>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in,
                                                               store_out, local_first, blockIdx.x);
    else
        bulk_body<St, Uv, Mode, StPol, Shell, false, GravZ>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in,
                                                            store_out, local_first, blockIdx.x);
}

// ---------------------------------------------------------------- rare-cell kernels
// The listed rare cells take two launches. The gather launch has one thread per
// (cell, link) and one more per cell for the temperature: every face rule and
// thermal stencil runs as a short independent job, so the launch is wide and its
// loads overlap. The update launch has one thread per cell: it reads the 27
// incoming values and the new temperature back and does the rest of the update.
// A listed cell reads and writes only slots that no other cell touches in the
// same step, so both launches are independent of the bulk kernel.
// i: index in the rare list, q: 0..26 incoming population, 27 temperature
template <class St, class Uv, int Mode>
__device__ __forceinline__ void gather_body(const StepParams& P, const CellFields& C, const uint32_t* __restrict__ mask,
                                            const typename St::T* __restrict__ fin,
                                            const typename St::T* __restrict__ fout, const float* __restrict__ t_in,
                                            const typename Uv::T* __restrict__ u_state,
                                            const float* __restrict__ store_in, int local_first,
                                            const int* __restrict__ list, int count, int cap,
                                            float* __restrict__ vals, int i, int q) {
    if (i >= count) return;
    const int cell = list[i];
    const uint32_t m = mask[cell];
    const int yz = P.ny * P.nz;
    const int x = cell / yz, rem = cell - x * yz;
    const int y = rem / P.nz, z = rem - y * P.nz;
    const size_t N = (size_t)P.cells, Np = (size_t)P.plane;
    const uint32_t links = m & kMaskLinks;
    const Nbr nb = make_nbr(P, x, y, z);
    float v = 0.0f;
    if (q == 27) {
        if (!P.thermal_on) return;
        const float ux0 = Uv::ld(u_state + cell), uy0 = Uv::ld(u_state + N + cell), uz0 = Uv::ld(u_state + 2 * N + cell);
        if ((m & kMaskTherm) || !P.thermal_wrap_matches)
            v = temperature_step(P, t_in, C.flags, C.src, cell, x, y, z, nb, false, ux0, uy0, uz0);
        else
            v = temperature_fast<false>(P, t_in, C.src, cell, x, y, z, nb, links, ux0, uy0, uz0);
    } else {
        const typename St::T* const src = Mode == 0 ? fin : fout;
        const bool sp = (links >> q) & 1u;
        if (Mode == 1 || local_first) {
            v = St::ld(src + q * Np + cell, q);
        } else {
            // special link: the cell's own post-collision value of opp(q) (halfway bounce-back)
            const int qs = Mode == 0 ? (sp ? lat_opp(q) : q) : (sp ? q : lat_opp(q));
            const int s = sp ? cell : cell + nb.off(-lat_cx(q), -lat_cy(q), -lat_cz(q));
            v = St::ld(src + qs * Np + s, q);
        }
        if ((m & kMaskFace) && !local_first && sp && link_oob(P, x, y, z, q)) {
            const PrevPost<St, Mode != 0> prev{fin, store_in, Np};
            v = face_link<St, Mode != 0>(P, C.flags, C.ref_p[cell], prev, cell, x, y, z, q, v);
        }
    }
    vals[(size_t)q * cap + i] = v;
}

template <class St, class Uv, int Mode>
__global__ void __launch_bounds__(256) rare_gather_kernel(StepParams P, CellFields C, const uint32_t* __restrict__ mask,
                                                          const typename St::T* __restrict__ fin,
                                                          const typename St::T* __restrict__ fout,
                                                          const float* __restrict__ t_in,
                                                          const typename Uv::T* __restrict__ u_state,
                                                          const float* __restrict__ store_in, int local_first,
                                                          const int* __restrict__ list, int count, int cap,
                                                          float* __restrict__ vals) {
    gather_body<St, Uv, Mode>(P, C, mask, fin, fout, t_in, u_state, store_in, local_first, list, count, cap, vals,
                              blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

template <class St, class Uv, int Mode>
__device__ __forceinline__ void update_body(const StepParams& P, const CellFields& C, const uint32_t* __restrict__ mask,
                                            typename St::T* fout, float* __restrict__ t_out,
                                            typename Uv::T* __restrict__ u_state, float* __restrict__ store_out,
                                            const int* __restrict__ list, int count, int cap,
                                            const float* __restrict__ vals, int i) {
    if (i >= count) return;
    const int cell = list[i];
    const uint32_t m = mask[cell];
    const int yz = P.ny * P.nz;
    const int x = cell / yz, rem = cell - x * yz;
    const int y = rem / P.nz, z = rem - y * P.nz;
    const size_t N = (size_t)P.cells, Np = (size_t)P.plane;
    const uint32_t links = m & kMaskLinks;
    float g[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) g[q] = vals[(size_t)q * cap + i];
    float temp = 0.0f;
    if (P.thermal_on) {
        temp = vals[(size_t)27 * cap + i];
        t_out[cell] = temp;
    }
    CollideResult r = collide<St::kShifted>(P, g, temp);
    if (!r.ok) reseed<St::kShifted>(P, C, cell, g, r);
    if (P.shell_on) {
        const int layer = C.shell[cell];
        if (layer) shell_blend<St::kShifted>(P, C.shell_state, cell, layer, g, r);
    }
    if (P.sponge_on) sponge<St::kShifted>(P, x, y, z, g);
    if (P.thermal_on) {
        Uv::st(u_state + cell, r.ux);
        Uv::st(u_state + N + cell, r.uy);
        Uv::st(u_state + 2 * N + cell, r.uz);
    }
    if (Mode != 0 && P.face_store_faces) face_store(P, x, y, z, g, store_out);
    if (Mode == 0) {
#pragma unroll
        for (int q = 0; q < 27; ++q) St::st(fout + q * Np + cell, g[q], q);
    } else if (Mode == 1) {
#pragma unroll
        for (int q = 0; q < 27; ++q) St::st(fout + lat_opp(q) * Np + cell, g[q], q);
    } else {
        const Nbr nb = make_nbr(P, x, y, z);
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const bool sp = (links >> lat_opp(q)) & 1u;  // target x + c_q solid or out of the domain
            const int d = sp ? cell : cell + nb.off(lat_cx(q), lat_cy(q), lat_cz(q));
            St::st(fout + (sp ? lat_opp(q) : q) * Np + d, g[q], q);
        }
    }
}

template <class St, class Uv, int Mode>
__global__ void __launch_bounds__(128) rare_update_kernel(StepParams P, CellFields C, const uint32_t* __restrict__ mask,
                                                          typename St::T* fout, float* __restrict__ t_out,
                                                          typename Uv::T* __restrict__ u_state,
                                                          float* __restrict__ store_out, const int* __restrict__ list,
                                                          int count, int cap, const float* __restrict__ vals) {
    update_body<St, Uv, Mode>(P, C, mask, fout, t_out, u_state, store_out, list, count, cap, vals,
                              blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- readout
// layout: 0 = A-B post-collision buffer, 1 = A-A after an even step, 2 = A-A after an
// odd step, 3 = initial state (incoming values in own slots, no face rules)
template <class St>
__device__ __forceinline__ void cell_output_fast(const StepParams& P, const CellFields& C, const uint32_t* mask,
                                                 const typename St::T* f, const float* store, int layout, int x, int y,
                                                 int z, float scale, float o[4]) {
    o[0] = o[1] = o[2] = o[3] = 0.0f;
    const int cell = (x * P.ny + y) * P.nz + z;
    const uint32_t m = mask[cell];
    if (m & kMaskSolid) return;
    const size_t Np = (size_t)P.plane;
    const uint32_t links = m & kMaskLinks;
    const Nbr nb = make_nbr(P, x, y, z);
    float g[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const bool sp = (links >> q) & 1u;
        if (layout == 0) {
            const int src = sp ? cell : cell + nb.off(-lat_cx(q), -lat_cy(q), -lat_cz(q));
            g[q] = St::ld(f + (size_t)(sp ? lat_opp(q) : q) * Np + src, q);
        } else if (layout == 1) {
            const int src = sp ? cell : cell + nb.off(-lat_cx(q), -lat_cy(q), -lat_cz(q));
            g[q] = St::ld(f + (size_t)(sp ? q : lat_opp(q)) * Np + src, q);
        } else {
            g[q] = St::ld(f + (size_t)q * Np + cell, q);
        }
    }
    if ((m & kMaskFace) && layout != 3) {
        const PrevPost<St, true> prev_aa{f, store, Np};
        const PrevPost<St, false> prev_ab{f, store, Np};
        for (int q = 0; q < 27; ++q) {
            if (((links >> q) & 1u) && link_oob(P, x, y, z, q)) {
                float bnc = 0.0f;
#pragma unroll
                for (int k = 0; k < 27; ++k)
                    if (k == q) bnc = g[k];
                const float v = layout == 0
                                    ? face_link<St, false>(P, C.flags, C.ref_p[cell], prev_ab, cell, x, y, z, q, bnc)
                                    : face_link<St, true>(P, C.flags, C.ref_p[cell], prev_aa, cell, x, y, z, q, bnc);
#pragma unroll
                for (int k = 0; k < 27; ++k)
                    if (k == q) g[k] = v;
            }
        }
    }
    float rho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const float v = St::kShifted ? g[q] + wq(q) : g[q];
        if (!finite(v)) return;
        rho += v;
        if (lat_cx(q)) jx += (float)lat_cx(q) * v;
        if (lat_cy(q)) jy += (float)lat_cy(q) * v;
        if (lat_cz(q)) jz += (float)lat_cz(q) * v;
    }
    // output rules of the reference (jni.cpp:2535-2581)
    if (!finite(rho) || !finite(jx) || !finite(jy) || !finite(jz)) return;
    rho = clampf(rho, 0.97f, 1.03f);
    float ux = jx / rho, uy = jy / rho, uz = jz / rho;
    if (!finite(ux) || !finite(uy) || !finite(uz)) ux = uy = uz = 0.0f;
    if (fabsf(rho - 1.0f) < 1e-6f) rho = 1.0f;
    if (fabsf(ux) < 1e-7f) ux = 0.0f;
    if (fabsf(uy) < 1e-7f) uy = 0.0f;
    if (fabsf(uz) < 1e-7f) uz = 0.0f;
    o[0] = ux * scale;
    o[1] = uy * scale;
    o[2] = uz * scale;
    o[3] = clampf(rho - 1.0f, -0.03f, 0.03f);
}

template <class St>
__global__ void __launch_bounds__(256) readout_kernel(StepParams P, CellFields C, const uint32_t* __restrict__ mask,
                                                      const typename St::T* __restrict__ f,
                                                      const float* __restrict__ store, int layout, int stride, int sx,
                                                      int sy, int sz, int out_ch, float scale, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sx * sy * sz) return;
    const int ax = i / (sy * sz);
    const int rem = i - ax * sy * sz;
    const int ay = rem / sz;
    const int az = rem - ay * sz;
    float o[4];
    cell_output_fast<St>(P, C, mask, f, store, layout, min(P.nx - 1, ax * stride), min(P.ny - 1, ay * stride),
                         min(P.nz - 1, az * stride), scale, o);
    float* d = out + (size_t)i * out_ch;
    d[0] = o[0];
    d[1] = o[1];
    d[2] = o[2];
    d[3] = o[3];
    for (int k = 4; k < out_ch; ++k) d[k] = 0.0f;
}

// ---------------------------------------------------------------- batched bricks
// Every workgroup reads its brick's entry of the table (a wave-uniform index, so
// the parameters come through scalar loads) and runs the single-brick code on it.
// At least 5 waves per SIMD (at most 96 VGPRs): the odd (pull + push)
// instantiation once needed ~100 VGPRs; with one store per direction it fits in 80
// (6 waves) and the even one in 72 (7 waves), neither spilling.
template <class St, class Uv, int Mode, int StPol, bool Shell, bool YF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) step_kernel_batch(const BatchBrick* __restrict__ table,
                                                         const uint16_t* __restrict__ owner) {
    using T = typename St::T;
    const int b = __builtin_amdgcn_readfirstlane((int)owner[blockIdx.x]);
    const BatchBrick& B = table[b];
    const int L = (int)blockIdx.x - B.bulk_begin;
    if (xface_tile(B.P, L) || (YF && yface_tile(B.P, L)))
        bulk_body<St, Uv, Mode, StPol, Shell, true, false, YF>(B.P, B.C, B.mask, (const T*)B.f, (T*)B.f, B.t_in,
                                                               B.t_out, (typename Uv::T*)B.u_state, B.store_in,
                                                               B.store_out, 0, L);
    else
        bulk_body<St, Uv, Mode, StPol, Shell, false>(B.P, B.C, B.mask, (const T*)B.f, (T*)B.f, B.t_in, B.t_out,
                                                     (typename Uv::T*)B.u_state, B.store_in, B.store_out, 0, L);
}

template <class St, class Uv, int Mode>
__global__ void __launch_bounds__(256) rare_gather_batch(const BatchBrick* __restrict__ table,
                                                         const uint16_t* __restrict__ owner) {
    using T = typename St::T;
    const int b = __builtin_amdgcn_readfirstlane((int)owner[blockIdx.x]);
    const BatchBrick& B = table[b];
    gather_body<St, Uv, Mode>(B.P, B.C, B.mask, (const T*)B.f, (const T*)B.f, B.t_in,
                              (const typename Uv::T*)B.u_state, B.store_in, 0, B.rare_list, B.rare_count,
                              B.rare_cap, B.rare_vals, ((int)blockIdx.x - B.gather_begin) * 256 + (int)threadIdx.x,
                              blockIdx.y);
}

template <class St, class Uv, int Mode>
__global__ void __launch_bounds__(128) rare_update_batch(const BatchBrick* __restrict__ table,
                                                         const uint16_t* __restrict__ owner) {
    using T = typename St::T;
    const int b = __builtin_amdgcn_readfirstlane((int)owner[blockIdx.x]);
    const BatchBrick& B = table[b];
    update_body<St, Uv, Mode>(B.P, B.C, B.mask, (T*)B.f, B.t_out, (typename Uv::T*)B.u_state, B.store_out,
                              B.rare_list, B.rare_count, B.rare_cap, B.rare_vals,
                              ((int)blockIdx.x - B.update_begin) * 128 + (int)threadIdx.x);
}

// lround of a float (halves away from zero) through double, exact for every float.
// The division runs in double and rounds once to float, which equals the correctly
// rounded float quotient (this file's float divisions are approximate).
__device__ __forceinline__ int16_t quantize_i16(float v, float range) {
    if (!isfinite(v)) return 0;
    const float t = fminf(fmaxf((float)((double)v / (double)range), -1.0f), 1.0f) * 32767.0f;
    const double d = (double)t;
    return (int16_t)(d >= 0.0 ? floor(d + 0.5) : ceil(d - 0.5));
}

// one thread per sampled cell (ref: native/src/aero_lbm_simulation_bridge.cpp:5363-5380,
// quantisation 3122-3128)
template <class St>
__global__ void __launch_bounds__(256) publish_batch(const BatchBrick* __restrict__ table,
                                                     const uint16_t* __restrict__ owner) {
    const int b = __builtin_amdgcn_readfirstlane((int)owner[blockIdx.x]);
    const BatchBrick& B = table[b];
    const StepParams& P = B.P;
    const int stride = B.publish_stride;
    const int sx = (P.nx + stride - 1) / stride, sy = (P.ny + stride - 1) / stride, sz = (P.nz + stride - 1) / stride;
    const int i = ((int)blockIdx.x - B.publish_begin) * 256 + (int)threadIdx.x;
    if (i >= sx * sy * sz) return;
    const int ax = i / (sy * sz), rem = i - ax * sy * sz;
    const int ay = rem / sz, az = rem - ay * sz;
    float o[4];
    cell_output_fast<St>(P, B.C, B.mask, (const typename St::T*)B.f, B.store_out, B.publish_layout,
                         min(P.nx - 1, ax * stride), min(P.ny - 1, ay * stride), min(P.nz - 1, az * stride), 1.0f, o);
    short4 q;
    q.x = quantize_i16(o[0] * B.publish_scale, 5.6f);
    q.y = quantize_i16(o[1] * B.publish_scale, 5.6f);
    q.z = quantize_i16(o[2] * B.publish_scale, 5.6f);
    q.w = quantize_i16(o[3], 0.03f);
    reinterpret_cast<short4*>(B.publish_out)[i] = q;
}

// ---------------------------------------------------------------- init and link masks
// Static-cell patch of an A-A context (flip[i]: kPatch* bits of cells[i]; marks holds
// 1 at every restarted cell). The context's state is the reference's pre-collision
// populations f, as the next step reads them: a restarted fluid cell starts from the
// rest equilibrium, and every other fluid cell keeps the value it holds for the link
// toward a patched neighbour. In layout 2 or 3 (after an odd step, or initial) a
// cell's incoming values are its own slots, so only the restarted cell is written.
// In layout 1 (after an even step) the cell's incoming value of link q is slot
// opp(q) of its neighbour x - c_q, or its own slot q when that link bounces (the
// link mask, rebuilt for the new flags): a restarted cell's incoming slots are
// written there, and for a fluid neighbour x = y + c_q, not restarted, of a cell y
// whose state changed, the value of x's link q moves between its bounce slot (own
// slot q) and its streamed slot (slot opp(q) of y). Each slot is the incoming slot
// of exactly one cell, so the threads write disjoint slots.
__device__ __forceinline__ bool patch_neighbour(const StepParams& P, int x, int y, int z, int q, int* out) {
    const int a = wrapc(x + lat_cx(q), P.nx, P.hydro_wrap_lo[0], P.hydro_wrap_hi[0]);
    const int b = wrapc(y + lat_cy(q), P.ny, P.hydro_wrap_lo[1], P.hydro_wrap_hi[1]);
    const int c = wrapc(z + lat_cz(q), P.nz, P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]);
    if (a < 0 || b < 0 || c < 0 || a >= P.nx || b >= P.ny || c >= P.nz) return false;
    *out = (a * P.ny + b) * P.nz + c;
    return true;
}

template <class St>
__global__ void __launch_bounds__(256) patch_reset_kernel(StepParams P, const uint32_t* __restrict__ mask,
                                                          const uint8_t* __restrict__ flags,
                                                          const uint8_t* __restrict__ marks, typename St::T* f,
                                                          int layout, const int* __restrict__ cells,
                                                          const uint8_t* __restrict__ flip, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int cell = cells[i];
    const int kind = flip[i];
    if (cell < 0 || cell >= P.cells) return;
    const bool restart = (kind & kPatchReset) && !(kind & kPatchSolid);
    const size_t Np = (size_t)P.plane;
    const int yz = P.ny * P.nz;
    const int x = cell / yz, y = (cell - x * yz) / P.nz, z = cell - x * yz - y * P.nz;
    auto rest = [](int q) { return St::kShifted ? 0.0f : wq(q); };
    if (layout != 1) {
        if (restart)
            for (int q = 0; q < 27; ++q) St::st(f + q * Np + cell, rest(q), q);
        return;
    }
    if (kind & kPatchLinks) {
        for (int q = 0; q < 27; ++q) {
            if (q == 13) continue;
            int nb;
            if (!patch_neighbour(P, x, y, z, q, &nb)) continue;
            if ((flags[nb] & kCellSolid) || marks[nb]) continue;  // fluid and not restarted
            // nb pulls its link q from this cell: streamed slot opp(q) here, bounce slot q at nb
            typename St::T* streamed = f + lat_opp(q) * Np + cell;
            typename St::T* bounced = f + q * Np + nb;
            if (kind & kPatchSolid) *bounced = *streamed;
            else *streamed = *bounced;
        }
    }
    if (!restart) return;
    const uint32_t links = mask[cell] & kMaskLinks;
    for (int q = 0; q < 27; ++q) {
        int src;
        const bool own = ((links >> q) & 1u) || !patch_neighbour(P, x, y, z, lat_opp(q), &src);
        if (own) St::st(f + q * Np + cell, rest(q), q);
        else St::st(f + lat_opp(q) * Np + src, rest(q), q);
    }
}

template <class St, class Uv>
__global__ void __launch_bounds__(256) init_kernel(StepParams P, CellFields C, typename St::T* __restrict__ f,
                                                   float* __restrict__ T, float* __restrict__ T_other,
                                                   typename Uv::T* __restrict__ u_state, float t_init) {
    // ref: jni.cpp:1752-1781 (benchmark mode)
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    const size_t N = (size_t)P.cells;
    const bool solid = C.flags[cell] & kCellSolid;
    float rho = 1.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (!solid) {
        rho = clampf(1.0f + C.ref_p[cell], 0.97f, 1.03f);
        ux = C.ref_u[cell];
        uy = C.ref_u[N + cell];
        uz = C.ref_u[2 * N + cell];
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) St::st(f + (size_t)q * P.plane + cell, feq_s<St::kShifted>(q, rho, ux, uy, uz), q);
    float t = T[cell];
    if (fabsf(t) < 1e-8f) t = clampf(t_init, -1.0f, 1.0f);
    else if (!solid) t = clampf(C.ref_t[cell], -1.0f, 1.0f);
    // solid cells are never stepped by the fast kernels and the reference zeroes
    // their temperature on the first update; stencils never read them
    T[cell] = solid ? 0.0f : clampf(t, -1.0f, 1.0f);
    T_other[cell] = 0.0f;
    Uv::st(u_state + cell, ux);
    Uv::st(u_state + N + cell, uy);
    Uv::st(u_state + 2 * N + cell, uz);
}

__global__ void __launch_bounds__(256) mask_kernel(StepParams P, const uint8_t* __restrict__ flags,
                                                   uint32_t* __restrict__ mask) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    if (flags[cell] & kCellSolid) {
        mask[cell] = kMaskSolid;
        return;
    }
    const int yz = P.ny * P.nz;
    const int x = cell / yz;
    const int rem = cell - x * yz;
    const int y = rem / P.nz;
    const int z = rem - y * P.nz;
    uint32_t m = 0;
    bool special_elsewhere = false;  // a face rule the bulk kernel does not run
    for (int q = 0; q < 27; ++q) {
        const int sx = wrapc(x - lat_cx(q), P.nx, P.hydro_wrap_lo[0], P.hydro_wrap_hi[0]);
        const int sy = wrapc(y - lat_cy(q), P.ny, P.hydro_wrap_lo[1], P.hydro_wrap_hi[1]);
        const int sz = wrapc(z - lat_cz(q), P.nz, P.hydro_wrap_lo[2], P.hydro_wrap_hi[2]);
        const bool oob = sx < 0 || sy < 0 || sz < 0 || sx >= P.nx || sy >= P.ny || sz >= P.nz;
        if (oob) {
            m |= 1u << q;
            const int face = oob_face(P, sx, sy, sz);
            const FaceParams fc = P.faces_dev[face];
            const int k = fc.hydro_kind;
            const bool plain_bounce = k == 2 || (k == 0 && P.convective_disabled);
            if (!plain_bounce) m |= kMaskFace;
            // temperature stencil point x - c_q out of the domain (thermal wrap equals hydro wrap
            // whenever this mask drives the fast stencil)
            const bool therm = fc.thermal_kind == 4 || fc.thermal_kind == 5;
            if (therm) m |= kMaskTherm;
            // links leaving through an x face flagged xface_bulk are the bulk kernel's
            // and so are the symmetry links leaving through a y face flagged yface_bulk
            const bool bulk_face = (face < 2 && P.xface_bulk[face]) || (face >= 2 && face < 4 && P.yface_bulk[face - 2]);
            if ((!plain_bounce || therm) && !bulk_face) special_elsewhere = true;
        } else if (flags[(sx * P.ny + sy) * P.nz + sz] & kCellSolid) {
            m |= 1u << q;
        }
    }
    bool rare = !P.thermal_wrap_matches || special_elsewhere;
    const int d = min(min(min(x, y), z), min(min(P.nx - 1 - x, P.ny - 1 - y), P.nz - 1 - z));
    if (P.sponge_on && d < 4) rare = true;
    mask[cell] = m | (rare ? kMaskRare : 0u);
}

// Ordered list of the fluid cells marked kMaskRare (ascending cell index, so a
// wavefront of the rare-cell kernel reads consecutive cells): per-workgroup
// counts, one-workgroup exclusive scan, per-workgroup scatter.
__device__ __forceinline__ bool is_rare(const uint32_t* mask, int cells, int cell) {
    return cell < cells && (mask[cell] & kMaskRare) && !(mask[cell] & kMaskSolid);
}

__global__ void __launch_bounds__(256) rare_count_kernel(int cells, const uint32_t* __restrict__ mask,
                                                         int* __restrict__ counts) {
    __shared__ int wave_n[4];
    const int cell = blockIdx.x * 256 + threadIdx.x;
    const uint64_t b = __ballot(is_rare(mask, cells, cell));
    if ((threadIdx.x & 63) == 0) wave_n[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
}

__global__ void __launch_bounds__(1024) rare_scan_kernel(int blocks, int* __restrict__ counts, int* __restrict__ total) {
    __shared__ int part[1024];
    const int per = (blocks + 1023) / 1024;
    const int b0 = threadIdx.x * per, b1 = min(blocks, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; ++b) sum += counts[b];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < 1024; ++t) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int b = b0; b < b1; ++b) {
        const int v = counts[b];
        counts[b] = run;  // exclusive offset
        run += v;
    }
}

__global__ void __launch_bounds__(256) rare_scatter_kernel(int cells, const uint32_t* __restrict__ mask,
                                                           const int* __restrict__ offsets, int* __restrict__ list) {
    __shared__ int wave_n[4];
    const int cell = blockIdx.x * 256 + threadIdx.x;
    const bool r = is_rare(mask, cells, cell);
    const uint64_t b = __ballot(r);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_n[wave] = __popcll(b);
    __syncthreads();
    int base = offsets[blockIdx.x];
    for (int w = 0; w < wave; ++w) base += wave_n[w];
    if (r) list[base + __popcll(b & ((1ull << lane) - 1ull))] = cell;
}

}  // namespace fast

// ---------------------------------------------------------------- launchers
// 64 cells of padding between planes (measured: +4% on the 128^3 FP32 A-B step,
// neutral for FP16); AERO_LBM_PLANE_PAD overrides it
int fast_plane_stride(int cells) {
    static const int pad = [] {
        const char* e = std::getenv("AERO_LBM_PLANE_PAD");
        return e ? std::max(0, std::atoi(e)) : 64;
    }();
    return cells + pad;
}

static inline dim3 g1(int n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_fast_mask(const StepParams& P, const uint8_t* flags, uint32_t* mask, hipStream_t s) {
    hipLaunchKernelGGL(fast::mask_kernel, g1(P.cells), dim3(256), 0, s, P, flags, mask);
}

void launch_fast_init(const StepParams& P, const CellFields& C, FastStorage st, void* f, float* T, float* T_other,
                      void* u_state, float t_init, hipStream_t s) {
    using namespace fast;
    switch (st) {
        case FastStorage::F32:
            hipLaunchKernelGGL((init_kernel<StF32, UF32>), g1(P.cells), dim3(256), 0, s, P, C, (float*)f, T, T_other,
                               (float*)u_state, t_init);
            break;
        case FastStorage::F16:
            hipLaunchKernelGGL((init_kernel<StF16, UF16>), g1(P.cells), dim3(256), 0, s, P, C, (uint16_t*)f, T,
                               T_other, (uint16_t*)u_state, t_init);
            break;
        case FastStorage::BF16:
            hipLaunchKernelGGL((init_kernel<StBF16, UF16>), g1(P.cells), dim3(256), 0, s, P, C, (uint16_t*)f, T,
                               T_other, (uint16_t*)u_state, t_init);
            break;
        case FastStorage::S16:
            hipLaunchKernelGGL((init_kernel<StS16, UF16>), g1(P.cells), dim3(256), 0, s, P, C, (uint16_t*)f, T,
                               T_other, (uint16_t*)u_state, t_init);
            break;
    }
}

// Store policy: a brick whose working set fits the 256 MB last-level cache
// (Infinity Cache) keeps the default policy, so populations stay resident from
// one step to the next; a larger one streams its stores past the caches.
// AERO_LBM_CACHE_POLICY=default|stream overrides the choice.
constexpr size_t kResidentBytes = 192ull << 20;

static int store_policy(size_t working_set) {
    static const int forced = [] {
        const char* e = std::getenv("AERO_LBM_CACHE_POLICY");
        if (e && std::strcmp(e, "stream") == 0) return fast::kPolStream;
        if (e && std::strcmp(e, "default") == 0) return fast::kPolDefault;
        return -1;
    }();
    if (forced >= 0) return forced;
    return working_set > kResidentBytes ? fast::kPolStream : fast::kPolDefault;
}

template <class St, class Uv, int Mode, int StPol>
static void launch_bulk(const StepParams& P, const CellFields& C, const uint32_t* mask, const void* fin, void* fout,
                        const float* t_in, float* t_out, void* u_state, const float* store_in, float* store_out,
                        int lf, hipStream_t s) {
    using T = typename St::T;
    using U = typename Uv::T;
    const dim3 grid((unsigned)fast::tile_grid(P)), block(fast::kTileZ, fast::kTileY);
    // gravity along z (or no buoyancy): the instantiation without the x/y force branch
    const bool grav_z = !P.buoyancy_on || P.g_axis == 2;
    const bool yf = P.yface_bulk[0] || P.yface_bulk[1];
#define AERO_STEP(SH, GZ, YF)                                                                                      \
    hipLaunchKernelGGL((fast::step_kernel<St, Uv, Mode, StPol, SH, GZ, YF>), grid, block, 0, s, P, C, mask,      \
                       (const T*)fin, (T*)fout, t_in, t_out, (U*)u_state, store_in, store_out, lf)
    if (P.shell_on) {
        if (yf) AERO_STEP(true, false, true);
        else AERO_STEP(true, false, false);
    } else if (grav_z) {
        if (yf) AERO_STEP(false, true, true);
        else AERO_STEP(false, true, false);
    } else {
        if (yf) AERO_STEP(false, false, true);
        else AERO_STEP(false, false, false);
    }
#undef AERO_STEP
}

template <class St, class Uv, int Mode, int StPol>
static void launch_step_m(const StepParams& P, const CellFields& C, const uint32_t* mask, const void* fin, void* fout,
                          const float* t_in, float* t_out, void* u_state, const float* store_in, float* store_out,
                          int local_first, const FastRare& rare, hipStream_t s) {
    using T = typename St::T;
    using U = typename Uv::T;
    const int lf = Mode == 2 ? 0 : local_first;
    if (rare.count > 0) {
        // serial on the stream: a second stream beside the bulk kernel costs more in
        // cross-stream synchronisation than the two short launches take
        const dim3 ggrid((unsigned)((rare.count + 255) / 256), 28), ugrid((unsigned)((rare.count + 127) / 128));
        hipLaunchKernelGGL((fast::rare_gather_kernel<St, Uv, Mode>), ggrid, dim3(256), 0, s, P, C, mask,
                           (const T*)fin, (const T*)fout, t_in, (const U*)u_state, store_in, lf, rare.list,
                           rare.count, rare.cap, rare.vals);
        hipLaunchKernelGGL((fast::rare_update_kernel<St, Uv, Mode>), ugrid, dim3(128), 0, s, P, C, mask, (T*)fout,
                           t_out, (U*)u_state, store_out, rare.list, rare.count, rare.cap, (const float*)rare.vals);
    }
    launch_bulk<St, Uv, Mode, StPol>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, lf, s);
}

template <class St, class Uv, int StPol>
static void launch_step_p(const StepParams& P, const CellFields& C, const uint32_t* mask, const void* fin, void* fout,
                          const float* t_in, float* t_out, void* u_state, const float* store_in, float* store_out,
                          int mode, int local_first, const FastRare& rare, hipStream_t s) {
    if (mode == 0)
        launch_step_m<St, Uv, 0, StPol>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, local_first,
                                        rare, s);
    else if (mode == 1)
        launch_step_m<St, Uv, 1, StPol>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, local_first,
                                        rare, s);
    else
        launch_step_m<St, Uv, 2, StPol>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, 0, rare, s);
}

template <class St, class Uv>
static void launch_step_t(const StepParams& P, const CellFields& C, const uint32_t* mask, const void* fin, void* fout,
                          const float* t_in, float* t_out, void* u_state, const float* store_in, float* store_out,
                          int mode, int local_first, const FastRare& rare, hipStream_t s) {
    const size_t N = (size_t)P.cells;
    // populations (two arrays for A-B), temperature pair, carried velocity, link mask, flags
    const size_t ws = (mode == 0 ? 2 : 1) * 27 * sizeof(typename St::T) * (size_t)P.plane +
                      N * (8 + 3 * sizeof(typename Uv::T) + 5);
    if (store_policy(ws) == fast::kPolStream)
        launch_step_p<St, Uv, fast::kPolStream>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out,
                                                mode, local_first, rare, s);
    else
        launch_step_p<St, Uv, fast::kPolDefault>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out,
                                                 mode, local_first, rare, s);
}

void launch_fast_step(const StepParams& P, const CellFields& C, FastStorage st, const uint32_t* mask, const void* fin,
                      void* fout, const float* t_in, float* t_out, void* u_state, const float* store_in,
                      float* store_out, int mode, int local_first, const FastRare& rare, hipStream_t s) {
    using namespace fast;
    switch (st) {
        case FastStorage::F32:
            launch_step_t<StF32, UF32>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, mode,
                                       local_first, rare, s);
            break;
        case FastStorage::F16:
            launch_step_t<StF16, UF16>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, mode,
                                       local_first, rare, s);
            break;
        case FastStorage::BF16:
            launch_step_t<StBF16, UF16>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, mode,
                                        local_first, rare, s);
            break;
        case FastStorage::S16:
            launch_step_t<StS16, UF16>(P, C, mask, fin, fout, t_in, t_out, u_state, store_in, store_out, mode,
                                       local_first, rare, s);
            break;
    }
}

void launch_fast_rare_list(const StepParams& P, const uint32_t* mask, int* list, int* scratch, int* counter,
                           hipStream_t s) {
    const int blocks = (P.cells + 255) / 256;
    hipLaunchKernelGGL(fast::rare_count_kernel, dim3(blocks), dim3(256), 0, s, P.cells, mask, scratch);
    hipLaunchKernelGGL(fast::rare_scan_kernel, dim3(1), dim3(1024), 0, s, blocks, scratch, counter);
    hipLaunchKernelGGL(fast::rare_scatter_kernel, dim3(blocks), dim3(256), 0, s, P.cells, mask, scratch, list);
}

int fast_tile_blocks(const StepParams& P) { return fast::tile_grid(P); }
int fast_store_policy(size_t working_set_bytes) { return store_policy(working_set_bytes); }

template <class St, class Uv, int Mode, int StPol>
static void batch_step_t(bool shell, const BatchBrick* table, const BatchGrids& g, hipStream_t s) {
    if (g.gather_blocks > 0) {
        hipLaunchKernelGGL((fast::rare_gather_batch<St, Uv, Mode>), dim3((unsigned)g.gather_blocks, 28), dim3(256), 0,
                           s, table, g.gather_owner);
        hipLaunchKernelGGL((fast::rare_update_batch<St, Uv, Mode>), dim3((unsigned)g.update_blocks), dim3(128), 0, s,
                           table, g.update_owner);
    }
    const dim3 block(fast::kTileZ, fast::kTileY), grid((unsigned)g.bulk_blocks);
    if (shell && g.yface)
        hipLaunchKernelGGL((fast::step_kernel_batch<St, Uv, Mode, StPol, true, true>), grid, block, 0, s, table, g.bulk_owner);
    else if (shell)
        hipLaunchKernelGGL((fast::step_kernel_batch<St, Uv, Mode, StPol, true, false>), grid, block, 0, s, table, g.bulk_owner);
    else if (g.yface)
        hipLaunchKernelGGL((fast::step_kernel_batch<St, Uv, Mode, StPol, false, true>), grid, block, 0, s, table, g.bulk_owner);
    else
        hipLaunchKernelGGL((fast::step_kernel_batch<St, Uv, Mode, StPol, false, false>), grid, block, 0, s, table, g.bulk_owner);
}

template <class St, class Uv>
static void batch_step_s(int mode, int pol, bool shell, const BatchBrick* table, const BatchGrids& g, hipStream_t s) {
    if (pol == fast::kPolStream) {
        if (mode == 1) batch_step_t<St, Uv, 1, fast::kPolStream>(shell, table, g, s);
        else batch_step_t<St, Uv, 2, fast::kPolStream>(shell, table, g, s);
    } else {
        if (mode == 1) batch_step_t<St, Uv, 1, fast::kPolDefault>(shell, table, g, s);
        else batch_step_t<St, Uv, 2, fast::kPolDefault>(shell, table, g, s);
    }
}

void launch_batch_step(FastStorage st, int mode, int store_policy_, bool shell, const BatchBrick* table,
                       const BatchGrids& g, hipStream_t s) {
    switch (st) {
        case FastStorage::F32: batch_step_s<fast::StF32, fast::UF32>(mode, store_policy_, shell, table, g, s); break;
        case FastStorage::F16: batch_step_s<fast::StF16, fast::UF16>(mode, store_policy_, shell, table, g, s); break;
        case FastStorage::BF16: batch_step_s<fast::StBF16, fast::UF16>(mode, store_policy_, shell, table, g, s); break;
        case FastStorage::S16: batch_step_s<fast::StS16, fast::UF16>(mode, store_policy_, shell, table, g, s); break;
    }
}

void launch_batch_publish(FastStorage st, const BatchBrick* table, const BatchGrids& g, hipStream_t s) {
    if (g.publish_blocks <= 0) return;
    const dim3 grid((unsigned)g.publish_blocks);
    switch (st) {
        case FastStorage::F32: hipLaunchKernelGGL(fast::publish_batch<fast::StF32>, grid, dim3(256), 0, s, table, g.publish_owner); break;
        case FastStorage::F16: hipLaunchKernelGGL(fast::publish_batch<fast::StF16>, grid, dim3(256), 0, s, table, g.publish_owner); break;
        case FastStorage::BF16: hipLaunchKernelGGL(fast::publish_batch<fast::StBF16>, grid, dim3(256), 0, s, table, g.publish_owner); break;
        case FastStorage::S16: hipLaunchKernelGGL(fast::publish_batch<fast::StS16>, grid, dim3(256), 0, s, table, g.publish_owner); break;
    }
}

void launch_fast_patch_reset(const StepParams& P, FastStorage st, const uint32_t* mask, const uint8_t* flags,
                             const uint8_t* marks, void* f, int layout, const int* cells, const uint8_t* flip,
                             int count, hipStream_t s) {
    using namespace fast;
    if (count <= 0) return;
    const dim3 g((unsigned)((count + 255) / 256));
    switch (st) {
        case FastStorage::F32:
            hipLaunchKernelGGL(patch_reset_kernel<StF32>, g, dim3(256), 0, s, P, mask, flags, marks, (float*)f,
                               layout, cells, flip, count);
            break;
        case FastStorage::F16:
            hipLaunchKernelGGL(patch_reset_kernel<StF16>, g, dim3(256), 0, s, P, mask, flags, marks, (uint16_t*)f,
                               layout, cells, flip, count);
            break;
        case FastStorage::BF16:
            hipLaunchKernelGGL(patch_reset_kernel<StBF16>, g, dim3(256), 0, s, P, mask, flags, marks, (uint16_t*)f,
                               layout, cells, flip, count);
            break;
        case FastStorage::S16:
            hipLaunchKernelGGL(patch_reset_kernel<StS16>, g, dim3(256), 0, s, P, mask, flags, marks, (uint16_t*)f,
                               layout, cells, flip, count);
            break;
    }
}

void launch_fast_readout(const StepParams& P, const CellFields& C, FastStorage st, const uint32_t* mask, const void* f,
                         const float* store, int layout, int stride, int out_ch, float scale, float* out,
                         hipStream_t s) {
    using namespace fast;
    const int sx = (P.nx + stride - 1) / stride, sy = (P.ny + stride - 1) / stride, sz = (P.nz + stride - 1) / stride;
    const dim3 g = g1(sx * sy * sz);
    switch (st) {
        case FastStorage::F32:
            hipLaunchKernelGGL(readout_kernel<StF32>, g, dim3(256), 0, s, P, C, mask, (const float*)f, store, layout,
                               stride, sx, sy, sz, out_ch, scale, out);
            break;
        case FastStorage::F16:
            hipLaunchKernelGGL(readout_kernel<StF16>, g, dim3(256), 0, s, P, C, mask, (const uint16_t*)f, store,
                               layout, stride, sx, sy, sz, out_ch, scale, out);
            break;
        case FastStorage::BF16:
            hipLaunchKernelGGL(readout_kernel<StBF16>, g, dim3(256), 0, s, P, C, mask, (const uint16_t*)f, store,
                               layout, stride, sx, sy, sz, out_ch, scale, out);
            break;
        case FastStorage::S16:
            hipLaunchKernelGGL(readout_kernel<StS16>, g, dim3(256), 0, s, P, C, mask, (const uint16_t*)f, store,
                               layout, stride, sx, sy, sz, out_ch, scale, out);
            break;
    }
}

}  // namespace aero
