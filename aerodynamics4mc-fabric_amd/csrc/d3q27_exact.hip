// Exact FP32 A-B path: one thread per cell, fused pull-stream + temperature
// update + cumulant collision, in the operation order of the reference's CPU
// step so FP32 results match it bit for bit.
//
// Compiled with -ffp-contract=off (no FMA contraction) and IEEE division and
// square root, like the reference's g++ x86-64 build. Terms the reference adds
// that are structurally zero (a population times a zero lattice coefficient)
// are left out: the accumulators start at +0.0f and can never become -0.0f, so
// adding a signed zero never changes them.
//
// Step semantics (ref: native/src/aero_lbm_jni.cpp:8480-8507): the reference
// runs temperature update -> collide -> stream per step and keeps post-stream
// populations f. This kernel keeps post-collision populations f* instead: a
// step gathers f_n = stream(f*_n) (pull, with the reference's link rules), then
// updates T and collides to f*_{n+1}. The very first step after initialisation
// collides the initial populations in place (local_mode = 1). Outputs gather
// f_n on the fly (launch_exact_readout), which is how the reference's
// write_output sees post-stream populations.
#include "kernels.h"

namespace aero {
namespace exact {

struct W27 {
    float w[27];
};
__host__ __device__ constexpr W27 make_weights() {
    W27 r{};
    const float w1[3] = {1.0f / 6.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    int q = 0;
    for (int ix = 0; ix < 3; ++ix)
        for (int iy = 0; iy < 3; ++iy)
            for (int iz = 0; iz < 3; ++iz) r.w[q++] = w1[ix] * w1[iy] * w1[iz];
    return r;
}
__device__ constexpr W27 kW = make_weights();

__device__ __forceinline__ float rmax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float rmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float clampr(float v, float lo, float hi) { return rmin(hi, rmax(lo, v)); }
__device__ __forceinline__ bool fin(float v) { return __builtin_isfinite(v); }

// ref: native/src/aero_lbm_hydro_core.cpp:364-368
__device__ __forceinline__ float feq(int q, float rho, float ux, float uy, float uz) {
    const float cu = 3.0f * ((float)lat_cx(q) * ux + (float)lat_cy(q) * uy + (float)lat_cz(q) * uz);
    const float uu = ux * ux + uy * uy + uz * uz;
    return kW.w[q] * rho * (1.0f + cu + 0.5f * cu * cu - 1.5f * uu);
}

__device__ __forceinline__ int cell_of(const StepParams& P, int x, int y, int z) { return (x * P.ny + y) * P.nz + z; }

__device__ __forceinline__ bool wrap_hydro(const StepParams& P, int& sx, int& sy, int& sz) {
    if (sx < 0) { if (P.hydro_wrap_lo[0]) sx += P.nx; }
    else if (sx >= P.nx) { if (P.hydro_wrap_hi[0]) sx -= P.nx; }
    if (sy < 0) { if (P.hydro_wrap_lo[1]) sy += P.ny; }
    else if (sy >= P.ny) { if (P.hydro_wrap_hi[1]) sy -= P.ny; }
    if (sz < 0) { if (P.hydro_wrap_lo[2]) sz += P.nz; }
    else if (sz >= P.nz) { if (P.hydro_wrap_hi[2]) sz -= P.nz; }
    return sx >= 0 && sy >= 0 && sz >= 0 && sx < P.nx && sy < P.ny && sz < P.nz;
}

// index of the face the reference picks for an out-of-domain source (ref: jni.cpp:743-751)
__device__ __forceinline__ int oob_face(const StepParams& P, int x, int y, int z) {
    if (x < 0) return 0;
    if (x >= P.nx) return 1;
    if (y < 0) return 2;
    if (y >= P.ny) return 3;
    if (z < 0) return 4;
    return 5;
}

__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ FaceParams face_at(const StepParams& P, int i) { return P.faces_dev[i]; }

// ---------------------------------------------------------------- link rules
// All take the "previous post-collision" buffer fp (f*), read-only in a step.
struct Src {
    const float* fp;
    const uint8_t* flags;
    const float* ref_p;
    size_t N;
};

__device__ __forceinline__ float bounce(const Src& s, int cell, int q) {
    // kObstacleBounceBlend = 0 (ref: jni.cpp:140, 1475-1492)
    return s.fp[(size_t)lat_opp(q) * s.N + cell];
}

// ref: jni.cpp:1494-1516
__device__ __forceinline__ float convective(const StepParams& P, const Src& s, int cell, int x, int y, int z, int q, int sx, int sy,
                            int sz) {
    const int dx = (sx < 0 || sx >= P.nx) ? lat_cx(q) : 0;
    const int dy = (sy < 0 || sy >= P.ny) ? lat_cy(q) : 0;
    const int dz = (sz < 0 || sz >= P.nz) ? lat_cz(q) : 0;
    const int s1 = cell_of(P, iclamp(x + dx, 0, P.nx - 1), iclamp(y + dy, 0, P.ny - 1), iclamp(z + dz, 0, P.nz - 1));
    if (s.flags[s1] & kCellSolid) return bounce(s, cell, q);
    const float f1 = s.fp[(size_t)q * s.N + s1];
    const int s2 =
        cell_of(P, iclamp(x + 2 * dx, 0, P.nx - 1), iclamp(y + 2 * dy, 0, P.ny - 1), iclamp(z + 2 * dz, 0, P.nz - 1));
    if (s.flags[s2] & kCellSolid) return f1;
    const float f2 = s.fp[(size_t)q * s.N + s2];
    return f1 + 0.15f * (f1 - f2);
}

// cylinder crossflow preset geometry (ref: jni.cpp:190-228), binary32 as the reference
// defines it: channel height ny - 1, diameter 0.1/2.2 of the length nx - 1 (at least
// 2.5 cells, at most the height - 4), centre at (0.2/2.2, 0.2/0.41) of (length, height)
__device__ __forceinline__ float cylinder_inlet_ux(const StepParams& P, int y, float u_max) {
    const float h = (float)(P.ny - 1 > 1 ? P.ny - 1 : 1);
    const float denom = rmax(h - 0.0f, 1.0e-6f);
    const float s = ((float)y - 0.0f) / denom;
    if (s < 0.0f || s > 1.0f) return 0.0f;
    return 4.0f * u_max * s * (1.0f - s);
}

__device__ __forceinline__ bool cylinder_solid(const StepParams& P, int x, int y) {
    const float len = (float)(P.nx - 1 > 1 ? P.nx - 1 : 1);
    const float h = (float)(P.ny - 1 > 1 ? P.ny - 1 : 1);
    float d = rmax(2.5f, 0.1f / 2.2f * len);
    d = rmin(d, rmax(2.5f, h - 4.0f));
    const float r = 0.5f * d;
    const float cx = 0.2f / 2.2f * len;
    const float cy = 0.0f + 0.2f / 0.41f * h;
    const float dx = (float)x - cx, dy = (float)y - cy;
    return dx * dx + dy * dy <= r * r;
}

// ref: jni.cpp:1518-1561
__device__ __forceinline__ float equilibrium_face(const StepParams& P, const Src& s, int cell, int y, int q, int sx,
                                                  const FaceParams& fc) {
    float rho = 1.0f + clampr(s.ref_p[cell], -0.03f, 0.03f);
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (fc.hydro_kind == 3 || fc.hydro_kind == 4) {
        ux = fc.vel[0];
        uy = fc.vel[1];
        uz = fc.vel[2];
        if (P.cylinder && sx < 0) {  // parabolic inlet (ref: jni.cpp:1540-1544)
            ux = cylinder_inlet_ux(P, y, fc.vel[0]);
            uy = 0.0f;
            uz = 0.0f;
        }
    } else if (fc.hydro_kind == 5) {
        rho = 1.0f + clampr(fc.pressure, -0.03f, 0.03f);
    }
    rho = clampr(rho, 0.97f, 1.03f);
    const float s2 = ux * ux + uy * uy + uz * uz;
    if (s2 > P.kmax_speed * P.kmax_speed && s2 > 0.0f) {
        const float sc = P.kmax_speed / __builtin_sqrtf(s2);
        ux *= sc;
        uy *= sc;
        uz *= sc;
    }
    return feq(q, rho, ux, uy, uz);
}

// ref: jni.cpp:1563-1602
__device__ __forceinline__ float face_value(const StepParams& P, const Src& s, int cell, int x, int y, int z, int q, int sx, int sy,
                            int sz) {
    const FaceParams fc = face_at(P, oob_face(P, sx, sy, sz));
    switch (fc.hydro_kind) {
        case 2:
            return bounce(s, cell, q);
        case 3:
        case 4:
        case 5:
            return equilibrium_face(P, s, cell, y, q, sx, fc);
        case 7: {
            const int src = cell_of(P, iclamp(sx, 0, P.nx - 1), iclamp(sy, 0, P.ny - 1), iclamp(sz, 0, P.nz - 1));
            return (s.flags[src] & kCellSolid) ? bounce(s, cell, q) : s.fp[(size_t)q * s.N + src];
        }
        case 6:
        case 1:
            return convective(P, s, cell, x, y, z, q, sx, sy, sz);
        default:
            if (P.convective_disabled) return bounce(s, cell, q);
            return convective(P, s, cell, x, y, z, q, sx, sy, sz);
    }
}

// f_n at one cell: pull of f* with the reference's link rules (jni.cpp:2415-2452)
__device__ __forceinline__ void gather(const StepParams& P, const Src& s, int cell, int x, int y, int z, bool boundary,
                                       float f[27]) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        int sx = x - lat_cx(q), sy = y - lat_cy(q), sz = z - lat_cz(q);
        const bool in = wrap_hydro(P, sx, sy, sz);
        if (!boundary) {
            f[q] = s.fp[(size_t)q * s.N + cell_of(P, sx, sy, sz)];
        } else if (!in) {
            f[q] = face_value(P, s, cell, x, y, z, q, sx, sy, sz);
        } else {
            const int src = cell_of(P, sx, sy, sz);
            f[q] = (s.flags[src] & kCellSolid) ? bounce(s, cell, q) : s.fp[(size_t)q * s.N + src];
        }
    }
}

// ---------------------------------------------------------------- temperature (jni.cpp:1799-1830, 1910-1945, 2145-2188)
__device__ __forceinline__ float t_neighbor(const StepParams& P, const float* T, const uint8_t* flags, int x, int y,
                                            int z, float self) {
    // periodic wrap as in the reference (coord %= dim, then + dim if negative)
    if ((x < 0 && P.thermal_wrap_lo[0]) || (x >= P.nx && P.thermal_wrap_hi[0])) { x %= P.nx; if (x < 0) x += P.nx; }
    if ((y < 0 && P.thermal_wrap_lo[1]) || (y >= P.ny && P.thermal_wrap_hi[1])) { y %= P.ny; if (y < 0) y += P.ny; }
    if ((z < 0 && P.thermal_wrap_lo[2]) || (z >= P.nz && P.thermal_wrap_hi[2])) { z %= P.nz; if (z < 0) z += P.nz; }
    if (x < 0 || y < 0 || z < 0 || x >= P.nx || y >= P.ny || z >= P.nz) {
        const FaceParams fc = face_at(P, oob_face(P, x, y, z));
        if (fc.thermal_kind == 4) return clampr(fc.temperature, -1.0f, 1.0f);
        if (fc.thermal_kind == 5) return clampr(self + fc.heat_flux, -1.0f, 1.0f);
        return self;
    }
    const int c = cell_of(P, x, y, z);
    if (flags[c] & kCellSolid) return self;
    return T[c];
}

// trilinear sample of `field` at (px, py, pz); corners outside the domain or in solid
// cells read as in t_neighbor with fallback fb (ref: jni.cpp:1910-1945)
__device__ __forceinline__ float trilinear(const StepParams& P, const float* field, const uint8_t* flags, float px,
                                           float py, float pz, float fb) {
    const int x0 = (int)floorf(px), y0 = (int)floorf(py), z0 = (int)floorf(pz);
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const float fx = px - (float)x0, fy = py - (float)y0, fz = pz - (float)z0;
    const float c000 = t_neighbor(P, field, flags, x0, y0, z0, fb);
    const float c100 = t_neighbor(P, field, flags, x1, y0, z0, fb);
    const float c010 = t_neighbor(P, field, flags, x0, y1, z0, fb);
    const float c110 = t_neighbor(P, field, flags, x1, y1, z0, fb);
    const float c001 = t_neighbor(P, field, flags, x0, y0, z1, fb);
    const float c101 = t_neighbor(P, field, flags, x1, y0, z1, fb);
    const float c011 = t_neighbor(P, field, flags, x0, y1, z1, fb);
    const float c111 = t_neighbor(P, field, flags, x1, y1, z1, fb);
    const float c00 = c000 + (c100 - c000) * fx;
    const float c10 = c010 + (c110 - c010) * fx;
    const float c01 = c001 + (c101 - c001) * fx;
    const float c11 = c011 + (c111 - c011) * fx;
    const float c0 = c00 + (c10 - c00) * fy;
    const float c1 = c01 + (c11 - c01) * fy;
    return c0 + (c1 - c0) * fz;
}

__device__ __forceinline__ float update_temperature(const StepParams& P, const float* T, const uint8_t* flags, const float* src, int cell,
                                    int x, int y, int z, float ux, float uy, float uz) {
    const float dt = 1.0f;
    const float tc = T[cell];
    const float adv = trilinear(P, T, flags, (float)x - dt * ux, (float)y - dt * uy, (float)z - dt * uz, tc);
    const float tsum = t_neighbor(P, T, flags, x + 1, y, z, tc) + t_neighbor(P, T, flags, x - 1, y, z, tc) +
                       t_neighbor(P, T, flags, x, y + 1, z, tc) + t_neighbor(P, T, flags, x, y - 1, z, tc) +
                       t_neighbor(P, T, flags, x, y, z + 1, tc) + t_neighbor(P, T, flags, x, y, z - 1, tc);
    const float lap = tsum - 6.0f * tc;
    float source = 0.0f;
    if (P.source_on && P.in_ch > 9) source = clampr(src[cell], -0.006f, 0.006f);
    const float tn = adv + dt * (P.alpha * lap + source - 0.0f * adv);
    return clampr(tn, -1.0f, 1.0f);
}

// ---------------------------------------------------------------- BFECC temperature (jni.cpp:2050-2139)
// Three passes over the cells, each reading only fields the previous pass finished:
// forward advection T -> S, backward advection of S corrected into N = T + (T - B) / 2,
// forward advection of N clamped to the bounds of T around the cell plus diffusion
// and source of T. Velocities are those of the previous step (u_state).
__device__ __forceinline__ void decode_cell(const StepParams& P, int cell, int& x, int& y, int& z) {
    const int yz = P.ny * P.nz;
    x = cell / yz;
    const int rem = cell - x * yz;
    y = rem / P.nz;
    z = rem - y * P.nz;
}

// distance of a cell to the nearest face of the brick
__device__ __forceinline__ int face_distance(const StepParams& P, int x, int y, int z) {
    int d = x;
    d = y < d ? y : d;
    d = z < d ? z : d;
    d = (P.nx - 1 - x) < d ? (P.nx - 1 - x) : d;
    d = (P.ny - 1 - y) < d ? (P.ny - 1 - y) : d;
    d = (P.nz - 1 - z) < d ? (P.nz - 1 - z) : d;
    return d;
}

// effective_runtime_state_nudge (ref: jni.cpp:683-693): 0.08 (1 - d/8)^2 within 8 cells
// of a face in runtime mode, 0 in benchmark mode
__device__ __forceinline__ float state_nudge(const StepParams& P, int x, int y, int z) {
    if (!P.runtime_mode) return 0.0f;
    const int d = face_distance(P, x, y, z);
    if (d >= 8) return 0.0f;
    const float eta = (float)(8 - d) / (float)8;
    return 0.08f * eta * eta;
}

// velocity_component_or_self (ref: jni.cpp:1952-1969): a face neighbour's collision
// velocity, or the cell's own outside the domain (no periodic wrap) or in a solid
__device__ __forceinline__ float vel_or_self(const StepParams& P, const uint8_t* flags, const float* u, int x, int y,
                                             int z, float self) {
    if (x < 0 || y < 0 || z < 0 || x >= P.nx || y >= P.ny || z >= P.nz) return self;
    const int n = cell_of(P, x, y, z);
    if (flags[n] & kCellSolid) return self;
    return u[n];
}

// runtime thermal diffusivity (ref: jni.cpp:1971-2048): alpha0 + nu_t / 0.85 with the
// Smagorinsky viscosity C^2 |S| (capped at 5 nu0) of the central-difference strain of
// the previous step's collision velocities; nu_t = 0 next to a solid cell
__device__ __forceinline__ float runtime_alpha(const StepParams& P, const uint8_t* flags, const float* u_state,
                                               int cell, int x, int y, int z) {
    float nu_t = 0.0f;
    bool adjacent = false;
    {
        const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int a = x + off[k][0], b = y + off[k][1], d = z + off[k][2];
            if (a < 0 || b < 0 || d < 0 || a >= P.nx || b >= P.ny || d >= P.nz) continue;
            if (flags[cell_of(P, a, b, d)] & kCellSolid) adjacent = true;
        }
    }
    if (P.sgs_on && !adjacent) {
        const size_t N = (size_t)P.cells;
        const float* ux = u_state;
        const float* uy = u_state + N;
        const float* uz = u_state + 2 * N;
        const float cx = ux[cell], cy = uy[cell], cz = uz[cell];
        const float dux_dx = 0.5f * (vel_or_self(P, flags, ux, x + 1, y, z, cx) - vel_or_self(P, flags, ux, x - 1, y, z, cx));
        const float dux_dy = 0.5f * (vel_or_self(P, flags, ux, x, y + 1, z, cx) - vel_or_self(P, flags, ux, x, y - 1, z, cx));
        const float dux_dz = 0.5f * (vel_or_self(P, flags, ux, x, y, z + 1, cx) - vel_or_self(P, flags, ux, x, y, z - 1, cx));
        const float duy_dx = 0.5f * (vel_or_self(P, flags, uy, x + 1, y, z, cy) - vel_or_self(P, flags, uy, x - 1, y, z, cy));
        const float duy_dy = 0.5f * (vel_or_self(P, flags, uy, x, y + 1, z, cy) - vel_or_self(P, flags, uy, x, y - 1, z, cy));
        const float duy_dz = 0.5f * (vel_or_self(P, flags, uy, x, y, z + 1, cy) - vel_or_self(P, flags, uy, x, y, z - 1, cy));
        const float duz_dx = 0.5f * (vel_or_self(P, flags, uz, x + 1, y, z, cz) - vel_or_self(P, flags, uz, x - 1, y, z, cz));
        const float duz_dy = 0.5f * (vel_or_self(P, flags, uz, x, y + 1, z, cz) - vel_or_self(P, flags, uz, x, y - 1, z, cz));
        const float duz_dz = 0.5f * (vel_or_self(P, flags, uz, x, y, z + 1, cz) - vel_or_self(P, flags, uz, x, y, z - 1, cz));
        const float sxx = dux_dx, syy = duy_dy, szz = duz_dz;
        const float sxy = 0.5f * (dux_dy + duy_dx);
        const float sxz = 0.5f * (dux_dz + duz_dx);
        const float syz = 0.5f * (duy_dz + duz_dy);
        const float smag = __builtin_sqrtf(rmax(0.0f, 2.0f * (sxx * sxx + syy * syy + szz * szz) +
                                                          4.0f * (sxy * sxy + sxz * sxz + syz * syz)));
        nu_t = P.sgs_c2 * smag;
        nu_t = rmin(nu_t, P.nut_cap_ratio * P.nu_rt0);
        nu_t = rmax(0.0f, nu_t);
    }
    return P.alpha + nu_t / rmax(1.0e-6f, 0.85f);
}

__global__ void __launch_bounds__(256) bfecc_forward_kernel(StepParams P, const uint8_t* __restrict__ flags,
                                                            const float* __restrict__ T,
                                                            const float* __restrict__ u_state, float* __restrict__ S) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    if (flags[cell] & kCellSolid) { S[cell] = 0.0f; return; }
    const size_t N = (size_t)P.cells;
    int x, y, z;
    decode_cell(P, cell, x, y, z);
    const float dt = P.thermal_dt;
    S[cell] = trilinear(P, T, flags, (float)x - dt * u_state[cell], (float)y - dt * u_state[N + cell],
                        (float)z - dt * u_state[2 * N + cell], T[cell]);
}

__global__ void __launch_bounds__(256) bfecc_backward_kernel(StepParams P, const uint8_t* __restrict__ flags,
                                                             const float* __restrict__ T, const float* __restrict__ S,
                                                             const float* __restrict__ u_state,
                                                             float* __restrict__ Nf) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    if (flags[cell] & kCellSolid) { Nf[cell] = 0.0f; return; }
    const size_t N = (size_t)P.cells;
    int x, y, z;
    decode_cell(P, cell, x, y, z);
    const float dt = P.thermal_dt;
    const float tc = T[cell];
    const float back = trilinear(P, S, flags, (float)x + dt * u_state[cell], (float)y + dt * u_state[N + cell],
                                 (float)z + dt * u_state[2 * N + cell], S[cell]);
    Nf[cell] = clampr(tc + 0.5f * (tc - back), -1.0f, 1.0f);
}

__global__ void __launch_bounds__(256) bfecc_correct_kernel(StepParams P, const uint8_t* __restrict__ flags,
                                                            const float* __restrict__ src, const float* __restrict__ T,
                                                            const float* __restrict__ Nf,
                                                            const float* __restrict__ u_state,
                                                            float* __restrict__ out) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    if (flags[cell] & kCellSolid) { out[cell] = 0.0f; return; }
    const size_t N = (size_t)P.cells;
    int x, y, z;
    decode_cell(P, cell, x, y, z);
    const float dt = P.thermal_dt;
    const float tc = T[cell];
    float adv = trilinear(P, Nf, flags, (float)x - dt * u_state[cell], (float)y - dt * u_state[N + cell],
                          (float)z - dt * u_state[2 * N + cell], Nf[cell]);
    // local_temperature_bounds (jni.cpp:1883-1907) of T, centred on the cell
    const float nb[6] = {t_neighbor(P, T, flags, x + 1, y, z, tc), t_neighbor(P, T, flags, x - 1, y, z, tc),
                         t_neighbor(P, T, flags, x, y + 1, z, tc), t_neighbor(P, T, flags, x, y - 1, z, tc),
                         t_neighbor(P, T, flags, x, y, z + 1, tc), t_neighbor(P, T, flags, x, y, z - 1, tc)};
    float lo = tc, hi = tc;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        lo = rmin(lo, nb[i]);
        hi = rmax(hi, nb[i]);
    }
    adv = clampr(adv, lo, hi);
    const float tsum = nb[0] + nb[1] + nb[2] + nb[3] + nb[4] + nb[5];
    const float lap = tsum - 6.0f * tc;
    float source = 0.0f;
    if (P.source_on && P.in_ch > 9) source = clampr(src[cell], -0.006f, 0.006f);
    const float alpha = P.runtime_mode ? runtime_alpha(P, flags, u_state, cell, x, y, z) : P.alpha;
    out[cell] = clampr(adv + dt * (alpha * lap + source - 0.0f * adv), -1.0f, 1.0f);
}

// runtime mode, every step after the first (ref: jni.cpp:8037-8064): solid cells to 0,
// cells within 8 of a face relaxed toward the payload temperature by the state nudge
__global__ void __launch_bounds__(256) t_reference_kernel(StepParams P, CellFields C, float* __restrict__ t) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    if (C.flags[cell] & kCellSolid) {
        t[cell] = 0.0f;
        return;
    }
    int x, y, z;
    decode_cell(P, cell, x, y, z);
    const float tn = state_nudge(P, x, y, z);
    if (tn <= 0.0f) return;
    const float rt = clampr(C.ref_t[cell], -1.0f, 1.0f);
    const float keep = 1.0f - tn;
    t[cell] = clampr(t[cell] * keep + rt * tn, -1.0f, 1.0f);
}

// ---------------------------------------------------------------- drag of the cylinder preset (jni.cpp:2452-2492)
// Momentum exchange 2 f*_q c_q over the in-plane links (c_z = 0) from fluid cells into
// solid cells of the preset's cylinder, in FP64: per-cell sums in q order, a fixed
// tree per block, then one block adds the block partials in a fixed order, so the
// result does not depend on scheduling.
__device__ __forceinline__ void block_sum3(double v[3], double* lds) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 3; ++k) lds[k * 256 + t] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) lds[k * 256 + t] += lds[k * 256 + t + w];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = lds[k * 256];
}

__global__ void __launch_bounds__(256) force_partial_kernel(StepParams P, const uint8_t* __restrict__ flags,
                                                            const float* __restrict__ fpost,
                                                            double* __restrict__ part) {
    __shared__ double lds[3 * 256];
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (cell < P.cells && !(flags[cell] & kCellSolid)) {
        const size_t N = (size_t)P.cells;
        int x, y, z;
        decode_cell(P, cell, x, y, z);
#pragma unroll
        for (int q = 1; q < 27; ++q) {
            if (lat_cz(q) != 0) continue;
            const int ax = x + lat_cx(q), ay = y + lat_cy(q);
            if (ax < 0 || ay < 0 || ax >= P.nx || ay >= P.ny) continue;
            if (!(flags[cell_of(P, ax, ay, z)] & kCellSolid)) continue;
            if (!cylinder_solid(P, ax, ay)) continue;
            const double fq = (double)fpost[(size_t)q * N + cell];
            acc[0] += 2.0 * fq * (double)lat_cx(q);
            acc[1] += 2.0 * fq * (double)lat_cy(q);
            acc[2] += 2.0 * fq * (double)lat_cz(q);
        }
    }
    block_sum3(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) part[3 * (size_t)blockIdx.x + k] = acc[k];
    }
}

__global__ void __launch_bounds__(256) force_final_kernel(StepParams P, const double* __restrict__ part, int blocks,
                                                          float* __restrict__ out) {
    __shared__ double lds[3 * 256];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < blocks; b += 256) {
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += part[3 * (size_t)b + k];
    }
    block_sum3(acc, lds);
    if (threadIdx.x < 3) {
        const double inv_depth = P.nz > 0 ? (1.0 / (double)P.nz) : 1.0;
        out[threadIdx.x] = (float)(acc[threadIdx.x] * inv_depth);
    }
}

// ---------------------------------------------------------------- cumulant collision (jni.cpp:1335-1473, 2209-2413)
template <int A, int B, int C>
__device__ __forceinline__ float raw_moment(const float f[27]) {
    // raw[a][b][c] = sum_q f_q * cx^a * cy^b * cz^c in q order; zero coefficients dropped
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int px = A == 0 ? 1 : (A == 1 ? lat_cx(q) : lat_cx(q) * lat_cx(q));
        const int py = B == 0 ? 1 : (B == 1 ? lat_cy(q) : lat_cy(q) * lat_cy(q));
        const int pz = C == 0 ? 1 : (C == 1 ? lat_cz(q) : lat_cz(q) * lat_cz(q));
        if (px * py * pz != 0) s += f[q] * (float)px * (float)py * (float)pz;
    }
    return s;
}

__device__ __forceinline__ int binom2(int n, int k) { return (n <= 1 || k == 0 || k == n) ? 1 : 2; }

__device__ __forceinline__ void raw_moments(const float f[27], float r[27]) {
#define RM(a, b, c) r[(a * 3 + b) * 3 + c] = raw_moment<a, b, c>(f);
    RM(0, 0, 0) RM(0, 0, 1) RM(0, 0, 2) RM(0, 1, 0) RM(0, 1, 1) RM(0, 1, 2) RM(0, 2, 0) RM(0, 2, 1) RM(0, 2, 2)
    RM(1, 0, 0) RM(1, 0, 1) RM(1, 0, 2) RM(1, 1, 0) RM(1, 1, 1) RM(1, 1, 2) RM(1, 2, 0) RM(1, 2, 1) RM(1, 2, 2)
    RM(2, 0, 0) RM(2, 0, 1) RM(2, 0, 2) RM(2, 1, 0) RM(2, 1, 1) RM(2, 1, 2) RM(2, 2, 0) RM(2, 2, 1) RM(2, 2, 2)
#undef RM
}

// central[a][b][c] = sum_{p<=a,q<=b,r<=c} C(a,p)C(b,q)C(c,r) s_x^(a-p) s_y^(b-q) s_z^(c-r) m[p][q][r],
// each term evaluated ((((coeff*sx)*sy)*sz)*m) as the reference does.
__device__ __forceinline__ void shift_moments(const float in[27], const float sx[3], const float sy[3], const float sz[3],
                                              float out[27]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float sum = 0.0f;
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int q = 0; q < 3; ++q)
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            // constant trip counts so every index folds at compile time
                            if (p <= a && q <= b && r <= c) {
                                const float coeff = (float)(binom2(a, p) * binom2(b, q) * binom2(c, r));
                                sum += coeff * sx[a - p] * sy[b - q] * sz[c - r] * in[(p * 3 + q) * 3 + r];
                            }
                        }
                out[(a * 3 + b) * 3 + c] = sum;
            }
}

__device__ __forceinline__ float inv1d(int i, int a) {
    // kMomentInv1D (ref: jni.cpp:46-50): rows c=-1,0,+1; columns moment order 0,1,2
    return i == 0 ? (a == 0 ? 0.0f : (a == 1 ? -0.5f : 0.5f))
                  : (i == 1 ? (a == 0 ? 1.0f : (a == 1 ? 0.0f : -1.0f)) : (a == 0 ? 0.0f : 0.5f));
}

__device__ __forceinline__ void raw_to_pop(const float m[27], float f[27]) {
#pragma unroll
    for (int ix = 0; ix < 3; ++ix)
#pragma unroll
        for (int iy = 0; iy < 3; ++iy)
#pragma unroll
            for (int iz = 0; iz < 3; ++iz) {
                float sum = 0.0f;
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const float k = inv1d(ix, a) * inv1d(iy, b) * inv1d(iz, c);
                            if (inv1d(ix, a) != 0.0f && inv1d(iy, b) != 0.0f && inv1d(iz, c) != 0.0f)
                                sum += k * m[(a * 3 + b) * 3 + c];
                        }
                f[(ix * 3 + iy) * 3 + iz] = sum;
            }
}

__device__ __forceinline__ float tau_from_nu(float nu) {
    const float nu_max = (0.95f - 0.5f) / 3.0f;
    return 0.5f + 3.0f * clampr(nu, 1.0e-8f, nu_max);
}

// hash_signed_noise (ref: jni.cpp:1318-1327): SplitMix64 finalizer of cell and step
// counter, the low 24 bits mapped to [-1, 1]
__device__ __forceinline__ float fan_noise(unsigned long long cell, unsigned long long tick) {
    unsigned long long x = (cell * 0x9E3779B97F4A7C15ull) ^ (tick * 0xD1B54A32D192ED03ull);
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    const float u = (float)(x & 0x00FFFFFFull) * (1.0f / 16777215.0f);
    return 2.0f * u - 1.0f;
}

struct CollideOut {
    float ux, uy, uz;
    bool reseeded;
};

__device__ __forceinline__ void reseed(const StepParams& P, const CellFields& C, int cell, float fo[27], CollideOut& o) {
    // ref: jni.cpp:1724-1750
    const size_t N = (size_t)P.cells;
    const float p = clampr(fin(C.ref_p[cell]) ? C.ref_p[cell] : 0.0f, -0.03f, 0.03f);
    const float rho = clampr(1.0f + p, 0.97f, 1.03f);
    float ux = fin(C.ref_u[cell]) ? C.ref_u[cell] : 0.0f;
    float uy = fin(C.ref_u[N + cell]) ? C.ref_u[N + cell] : 0.0f;
    float uz = fin(C.ref_u[2 * N + cell]) ? C.ref_u[2 * N + cell] : 0.0f;
    const float s2 = ux * ux + uy * uy + uz * uz;
    if (!fin(s2) || s2 > P.kmax_speed * P.kmax_speed) {
        if (!fin(s2) || s2 <= 0.0f) {
            ux = uy = uz = 0.0f;
        } else {
            const float sc = P.kmax_speed / __builtin_sqrtf(s2);
            ux *= sc;
            uy *= sc;
            uz *= sc;
        }
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) fo[q] = feq(q, rho, ux, uy, uz);
    o.ux = ux;
    o.uy = uy;
    o.uz = uz;
    o.reseeded = true;
}

__device__ __forceinline__ void collide(const StepParams& P, const CellFields& C, int cell, int x, int y, int z, const float fi[27],
                        float T, float fo[27], CollideOut& o) {
    o.reseeded = false;
    float rho = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    bool bad = false;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        if (!fin(fi[q])) bad = true;
    }
    if (!bad) {
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            rho += fi[q];
            if (lat_cx(q) != 0) mx += fi[q] * (float)lat_cx(q);
            if (lat_cy(q) != 0) my += fi[q] * (float)lat_cy(q);
            if (lat_cz(q) != 0) mz += fi[q] * (float)lat_cz(q);
        }
    }
    if (bad || !fin(rho) || !fin(mx) || !fin(my) || !fin(mz)) { reseed(P, C, cell, fo, o); return; }
    const float rho_safe = rmax(1e-6f, rho);
    const float inv_rho = 1.0f / rho_safe;
    float ux = mx * inv_rho, uy = my * inv_rho, uz = mz * inv_rho;
    if (!fin(rho_safe) || !fin(inv_rho) || !fin(ux) || !fin(uy) || !fin(uz)) { reseed(P, C, cell, fo, o); return; }
    float F[3] = {0.0f, 0.0f, 0.0f};
    bool fan_force = false;
    if (P.fan_on) {
        // fan forcing (ref: jni.cpp:2258-2286): push toward the fan speed along its axis,
        // damp the perpendicular velocity, soften above a speed cap
        const size_t N = (size_t)P.cells;
        const float mask = C.fan[cell];
        if (mask > 0.0f) {
            const float fu = C.fan[N + cell], fv = C.fan[2 * N + cell], fw = C.fan[3 * N + cell];
            const float speed_pre = __builtin_sqrtf(ux * ux + uy * uy + uz * uz);
            const float fan_norm = __builtin_sqrtf(fu * fu + fv * fv + fw * fw);
            if (fan_norm > 1e-8f) {
                const float inv_norm = 1.0f / fan_norm;
                const float noise = 1.0f + P.fan_noise_amp * fan_noise((unsigned long long)cell, P.tick);
                const float target = clampr(fan_norm * 0.05f * rmax(0.0f, noise), 0.0f, 0.28f);
                const float nx = fu * inv_norm, ny = fv * inv_norm, nz = fw * inv_norm;
                const float u_para = ux * nx + uy * ny + uz * nz;
                const float upx = ux - u_para * nx, upy = uy - u_para * ny, upz = uz - u_para * nz;
                const float push = rmax(0.0f, target - u_para);
                float damp = 1.0f;
                if (speed_pre > 0.24f) {
                    const float r = (speed_pre - 0.24f) / rmax(1e-4f, 0.04f);
                    damp = 1.0f / (1.0f + r * r);
                }
                const float fb = mask * 0.07f * damp;
                F[0] = fb * rho_safe * (push * nx - 1.0f * upx);
                F[1] = fb * rho_safe * (push * ny - 1.0f * upy);
                F[2] = fb * rho_safe * (push * nz - 1.0f * upz);
                fan_force = true;
            }
        }
    }
    // buoyancy along the gravity axis, added to the fan force; components the
    // reference never touches stay +0.0f
    if (P.buoyancy_on) {
        const float b = clampr(P.beta * T, -0.02f, 0.02f);
        const float fa = P.g_sign < 0 ? rho_safe * b : rho_safe * (-b);
        if (fan_force) {
            F[P.g_axis == 0 ? 0 : (P.g_axis == 1 ? 1 : 2)] += fa;
        } else {
            const float fz0 = 0.0f + fa;
            F[0] = P.g_axis == 0 ? fz0 : 0.0f;
            F[1] = P.g_axis == 1 ? fz0 : 0.0f;
            F[2] = P.g_axis == 2 ? fz0 : 0.0f;
        }
    }
    ux += 0.5f * (F[0] * inv_rho);
    uy += 0.5f * (F[1] * inv_rho);
    uz += 0.5f * (F[2] * inv_rho);
    if (P.runtime_mode) {
        // state nudge toward the reference velocity (ref: jni.cpp:2306-2314); in
        // benchmark mode it is zero and adds nothing
        const size_t N = (size_t)P.cells;
        const float sn = state_nudge(P, x, y, z);
        ux = (1.0f - sn) * ux + sn * C.ref_u[cell];
        uy = (1.0f - sn) * uy + sn * C.ref_u[N + cell];
        uz = (1.0f - sn) * uz + sn * C.ref_u[2 * N + cell];
    }
    if (!fin(ux) || !fin(uy) || !fin(uz)) { reseed(P, C, cell, fo, o); return; }
    const float s2 = ux * ux + uy * uy + uz * uz;
    if (!fin(s2)) { reseed(P, C, cell, fo, o); return; }
    if (s2 > P.kmax_speed * P.kmax_speed) {
        const float sc = P.kmax_speed / __builtin_sqrtf(s2);
        ux *= sc;
        uy *= sc;
        uz *= sc;
    }
    o.ux = ux;
    o.uy = uy;
    o.uz = uz;

    float m[27], c[27];
    raw_moments(fi, m);
    {
        const float sx[3] = {1.0f, -ux, ux * ux};
        const float sy[3] = {1.0f, -uy, uy * uy};
        const float sz[3] = {1.0f, -uz, uz * uz};
        shift_moments(m, sx, sy, sz, c);
    }
    const float cs2 = 1.0f / 3.0f;
    float ts = P.tau_shear_base, tn = P.tau_normal_base;
    if (P.sgs_on) {
        const float nu0 = rmax(1e-6f, P.nu_base);
        const float nxx = c[18] - rho_safe * cs2;  // c200
        const float nyy = c[6] - rho_safe * cs2;   // c020
        const float nzz = c[2] - rho_safe * cs2;   // c002
        const float nxy = c[12], nxz = c[10], nyz = c[4];
        const float qn2 = nxx * nxx + nyy * nyy + nzz * nzz + 2.0f * (nxy * nxy + nxz * nxz + nyz * nyz);
        const float qmag = __builtin_sqrtf(rmax(0.0f, qn2));
        const float smag = qmag / rmax(1e-6f, 2.0f * rho_safe * nu0);
        float nut = P.sgs_c2 * smag;
        nut = rmin(nut, P.nut_cap_ratio * nu0);
        const float tau_min = 0.5f + 3.0f * 1.0e-8f;
        ts = clampr(tau_from_nu(nu0 + nut), tau_min, 0.95f);
        tn = clampr(P.tau_normal_base + (ts - P.tau_shear_base) * 0.30f, tau_min, 0.95f);
    }
    const float om_diag = 1.0f / tn;
    const float om_off = 1.0f / ts;
    // relaxation + Gaussian closure (ref: jni.cpp:1384-1430); output reuses m[]
    {
        const float rr = rho_safe;
        const float eq2 = rr * cs2;
        const float sd = clampr(om_diag, 0.0f, 1.95f);
        const float so = clampr(om_off, 0.0f, 1.95f);
        const float c200 = c[18] + sd * (eq2 - c[18]);
        const float c020 = c[6] + sd * (eq2 - c[6]);
        const float c002 = c[2] + sd * (eq2 - c[2]);
        const float c110 = (1.0f - so) * c[12];
        const float c101 = (1.0f - so) * c[10];
        const float c011 = (1.0f - so) * c[4];
        const float ir = 1.0f / rmax(1e-6f, rr);
#pragma unroll
        for (int k = 0; k < 27; ++k) c[k] = 0.0f;
        c[0] = rr;
        c[18] = c200;
        c[6] = c020;
        c[2] = c002;
        c[12] = c110;
        c[10] = c101;
        c[4] = c011;
        c[24] = c200 * c020 * ir + 2.0f * c110 * c110 * ir;                       // c220
        c[20] = c200 * c002 * ir + 2.0f * c101 * c101 * ir;                       // c202
        c[8] = c020 * c002 * ir + 2.0f * c011 * c011 * ir;                        // c022
        c[22] = (c200 * c011 + 2.0f * c110 * c101) * ir;                          // c211
        c[16] = (c020 * c101 + 2.0f * c110 * c011) * ir;                          // c121
        c[14] = (c002 * c110 + 2.0f * c101 * c011) * ir;                          // c112
        c[26] = (c200 * c020 * c002 + 2.0f * c110 * c110 * c002 + 2.0f * c101 * c101 * c020 +
                 2.0f * c011 * c011 * c200 + 8.0f * c110 * c101 * c011) * ir * ir;  // c222
    }
    {
        const float px[3] = {1.0f, ux, ux * ux};
        const float py[3] = {1.0f, uy, uy * uy};
        const float pz[3] = {1.0f, uz, uz * uz};
        shift_moments(c, px, py, pz, m);
    }
    raw_to_pop(m, fo);
    // exact low-order conservation after inversion (ref: jni.cpp:2376-2394)
    float rc = 0.0f, mxc = 0.0f, myc = 0.0f, mzc = 0.0f;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        rc += fo[q];
        if (lat_cx(q) != 0) mxc += fo[q] * (float)lat_cx(q);
        if (lat_cy(q) != 0) myc += fo[q] * (float)lat_cy(q);
        if (lat_cz(q) != 0) mzc += fo[q] * (float)lat_cz(q);
    }
    const float rcs = rmax(1e-6f, rc);
    const float uxc = mxc / rcs, uyc = myc / rcs, uzc = mzc / rcs;
#pragma unroll
    for (int q = 0; q < 27; ++q) fo[q] += feq(q, rho_safe, ux, uy, uz) - feq(q, rcs, uxc, uyc, uzc);
    // Guo source (ref: jni.cpp:1270-1281); identically +-0 when F = 0
    if (F[0] != 0.0f || F[1] != 0.0f || F[2] != 0.0f) {
        const float inv_cs2 = 1.0f / (1.0f / 3.0f);
        const float inv_cs4 = inv_cs2 * inv_cs2;
        const float pref = rmax(0.0f, 1.0f - 0.5f * clampr(om_off, 0.0f, 1.999f));
        const float uf = ux * F[0] + uy * F[1] + uz * F[2];
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const float cx = (float)lat_cx(q), cy = (float)lat_cy(q), cz = (float)lat_cz(q);
            const float cu = cx * ux + cy * uy + cz * uz;
            const float cf = cx * F[0] + cy * F[1] + cz * F[2];
            fo[q] = fo[q] + pref * kW.w[q] * ((cf - uf) * inv_cs2 + cu * cf * inv_cs4);
        }
    }
    if (P.sponge_on) {
        // ref: jni.cpp:767-774, 2401-2411
        int d = x;
        d = y < d ? y : d;
        d = z < d ? z : d;
        d = (P.nx - 1 - x) < d ? (P.nx - 1 - x) : d;
        d = (P.ny - 1 - y) < d ? (P.ny - 1 - y) : d;
        d = (P.nz - 1 - z) < d ? (P.nz - 1 - z) : d;
        if (d < 4) {
            const float eta = (float)(4 - d) / (float)4;
            const float al = clampr(0.03f * eta * eta, 0.0f, 0.95f);
            if (al > 0.0f) {
                const float keep = 1.0f - al;
#pragma unroll
                for (int q = 0; q < 27; ++q) fo[q] = keep * fo[q] + al * feq(q, 1.0f, 0.0f, 0.0f, 0.0f);
            }
        }
    }
}

// ---------------------------------------------------------------- kernels
__global__ void __launch_bounds__(256) init_kernel(StepParams P, CellFields C, float* __restrict__ f,
                                                   float* __restrict__ T, float* __restrict__ u_state, float t_init) {
    // ref: jni.cpp:1752-1781 (benchmark mode)
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    const size_t N = (size_t)P.cells;
    const bool solid = C.flags[cell] & kCellSolid;
    float rho = 1.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (!solid) {
        rho = clampr(1.0f + C.ref_p[cell], 0.97f, 1.03f);
        ux = C.ref_u[cell];
        uy = C.ref_u[N + cell];
        uz = C.ref_u[2 * N + cell];
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) f[(size_t)q * N + cell] = feq(q, rho, ux, uy, uz);
    float t = T[cell];
    // benchmark mode seeds a zero temperature with the configured initial one; otherwise a
    // fluid cell takes the payload's temperature (ref: jni.cpp:1771-1776)
    if (!P.runtime_mode && fabsf(t) < 1e-8f) t = clampr(t_init, -1.0f, 1.0f);
    else if (!solid) t = clampr(C.ref_t[cell], -1.0f, 1.0f);
    T[cell] = clampr(t, -1.0f, 1.0f);
    u_state[cell] = ux;
    u_state[N + cell] = uy;
    u_state[2 * N + cell] = uz;
}

__global__ void __launch_bounds__(256) step_kernel(StepParams P, CellFields C, const float* __restrict__ fin,
                                                   float* __restrict__ fout, const float* __restrict__ t_in,
                                                   float* __restrict__ t_out, float* __restrict__ u_state,
                                                   int local_mode) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    const size_t N = (size_t)P.cells;
    const uint8_t fl = C.flags[cell];
    if (fl & kCellSolid) {
        // collide copies, stream swaps to the opposite slot (ref: jni.cpp:2212-2218, 2421-2424)
#pragma unroll
        for (int q = 0; q < 27; ++q) fout[(size_t)q * N + cell] = fin[(size_t)(local_mode ? q : lat_opp(q)) * N + cell];
        if (P.thermal_on) t_out[cell] = 0.0f;
        u_state[cell] = 0.0f;
        u_state[N + cell] = 0.0f;
        u_state[2 * N + cell] = 0.0f;
        return;
    }
    const int yz = P.ny * P.nz;
    const int x = cell / yz;
    const int rem = cell - x * yz;
    const int y = rem / P.nz;
    const int z = rem - y * P.nz;
    float f[27];
    if (local_mode) {
#pragma unroll
        for (int q = 0; q < 27; ++q) f[q] = fin[(size_t)q * N + cell];
    } else {
        const Src s{fin, C.flags, C.ref_p, N};
        gather(P, s, cell, x, y, z, (fl & kCellBoundary) != 0, f);
    }
    float T = 0.0f;
    if (P.thermal_on) {
        if (!P.thermal_update) {
            T = t_in[cell];  // runtime mode: no temperature update this step (stride 2)
            t_out[cell] = T;
        } else if (P.thermal_bfecc) {
            T = t_out[cell];  // the BFECC passes already wrote this step's temperature
        } else {
            T = update_temperature(P, t_in, C.flags, C.src, cell, x, y, z, u_state[cell], u_state[N + cell],
                                   u_state[2 * N + cell]);
            t_out[cell] = T;
        }
    }
    float fo[27];
    CollideOut o;
    collide(P, C, cell, x, y, z, f, T, fo, o);
#pragma unroll
    for (int q = 0; q < 27; ++q) fout[(size_t)q * N + cell] = fo[q];
    u_state[cell] = o.ux;
    u_state[N + cell] = o.uy;
    u_state[2 * N + cell] = o.uz;
}

// Stream only: f_n = stream(f*_n) into fout. Used when the payload changes between
// steps, so the pending stream runs with the payload the reference streamed with.
__global__ void __launch_bounds__(256) stream_kernel(StepParams P, CellFields C, const float* __restrict__ fin,
                                                     float* __restrict__ fout) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.cells) return;
    const size_t N = (size_t)P.cells;
    const uint8_t fl = C.flags[cell];
    if (fl & kCellSolid) {
#pragma unroll
        for (int q = 0; q < 27; ++q) fout[(size_t)q * N + cell] = fin[(size_t)lat_opp(q) * N + cell];
        return;
    }
    const int yz = P.ny * P.nz;
    const int x = cell / yz;
    const int rem = cell - x * yz;
    const int y = rem / P.nz;
    const int z = rem - y * P.nz;
    float f[27];
    const Src s{fin, C.flags, C.ref_p, N};
    gather(P, s, cell, x, y, z, (fl & kCellBoundary) != 0, f);
#pragma unroll
    for (int q = 0; q < 27; ++q) fout[(size_t)q * N + cell] = f[q];
}

// moments of f_n -> (u*scale, clamp(rho-1)) with the reference's output rules (jni.cpp:2535-2581)
__device__ __forceinline__ void cell_output(const float f[27], float scale, float o[4]) {
    o[0] = o[1] = o[2] = o[3] = 0.0f;
    float rho = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
#pragma unroll
    for (int q = 0; q < 27; ++q)
        if (!fin(f[q])) return;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        rho += f[q];
        if (lat_cx(q) != 0) ux += f[q] * (float)lat_cx(q);
        if (lat_cy(q) != 0) uy += f[q] * (float)lat_cy(q);
        if (lat_cz(q) != 0) uz += f[q] * (float)lat_cz(q);
    }
    if (!fin(rho) || !fin(ux) || !fin(uy) || !fin(uz)) return;
    rho = clampr(rho, 0.97f, 1.03f);
    if (!fin(rho) || rho <= 1e-6f) return;
    ux /= rho;
    uy /= rho;
    uz /= rho;
    if (!fin(ux) || !fin(uy) || !fin(uz)) ux = uy = uz = 0.0f;
    if (fabsf(rho - 1.0f) < 1e-6f) rho = 1.0f;
    if (fabsf(ux) < 1e-7f) ux = 0.0f;
    if (fabsf(uy) < 1e-7f) uy = 0.0f;
    if (fabsf(uz) < 1e-7f) uz = 0.0f;
    o[0] = ux * scale;
    o[1] = uy * scale;
    o[2] = uz * scale;
    o[3] = clampr(rho - 1.0f, -0.03f, 0.03f);
}

__device__ __forceinline__ void output_at(const StepParams& P, const CellFields& C, const float* fp, int post_collision, int x, int y,
                          int z, float scale, float o[4]) {
    const int cell = cell_of(P, x, y, z);
    const uint8_t fl = C.flags[cell];
    if (fl & kCellSolid) {
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        return;
    }
    float f[27];
    if (post_collision) {
        const Src s{fp, C.flags, C.ref_p, (size_t)P.cells};
        gather(P, s, cell, x, y, z, (fl & kCellBoundary) != 0, f);
    } else {
#pragma unroll
        for (int q = 0; q < 27; ++q) f[q] = fp[(size_t)q * P.cells + cell];
    }
    cell_output(f, scale, o);
}

__global__ void __launch_bounds__(256) readout_kernel(StepParams P, CellFields C, const float* __restrict__ fp,
                                                      int post_collision, int stride, int sx, int sy, int sz,
                                                      int out_ch, float scale, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sx * sy * sz) return;
    const int ax = i / (sy * sz);
    const int rem = i - ax * sy * sz;
    const int ay = rem / sz;
    const int az = rem - ay * sz;
    const int gx = min(P.nx - 1, ax * stride), gy = min(P.ny - 1, ay * stride), gz = min(P.nz - 1, az * stride);
    float o[4];
    output_at(P, C, fp, post_collision, gx, gy, gz, scale, o);
    float* dst = out + (size_t)i * out_ch;
    dst[0] = o[0];
    dst[1] = o[1];
    dst[2] = o[2];
    dst[3] = o[3];
    for (int k = 4; k < out_ch; ++k) dst[k] = 0.0f;
}

__global__ void sample_kernel(StepParams P, CellFields C, const float* __restrict__ fp, int post_collision, int x,
                              int y, int z, float* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float o[4];
    output_at(P, C, fp, post_collision, x, y, z, 1.0f, o);
    out[0] = o[0];
    out[1] = o[1];
    out[2] = o[2];
    out[3] = o[3];
}

}  // namespace exact

static inline dim3 grid1(int n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_exact_init(const StepParams& P, const CellFields& C, float* f, float* t, float* u_state,
                       float initial_temperature, hipStream_t s) {
    hipLaunchKernelGGL(exact::init_kernel, grid1(P.cells), dim3(256), 0, s, P, C, f, t, u_state, initial_temperature);
}

void launch_exact_step(const StepParams& P, const CellFields& C, const float* fin, float* fout, const float* t_in,
                       float* t_out, float* u_state, int local_mode, hipStream_t s) {
    hipLaunchKernelGGL(exact::step_kernel, grid1(P.cells), dim3(256), 0, s, P, C, fin, fout, t_in, t_out, u_state,
                       local_mode);
}

void launch_exact_bfecc(const StepParams& P, const CellFields& C, const float* t_in, float* t_out, float* t_mid,
                        const float* u_state, hipStream_t s) {
    // t_out holds S until the last pass overwrites it with the new temperature
    hipLaunchKernelGGL(exact::bfecc_forward_kernel, grid1(P.cells), dim3(256), 0, s, P, C.flags, t_in, u_state, t_out);
    hipLaunchKernelGGL(exact::bfecc_backward_kernel, grid1(P.cells), dim3(256), 0, s, P, C.flags, t_in,
                       (const float*)t_out, u_state, t_mid);
    hipLaunchKernelGGL(exact::bfecc_correct_kernel, grid1(P.cells), dim3(256), 0, s, P, C.flags, C.src, t_in,
                       (const float*)t_mid, u_state, t_out);
}

void launch_exact_t_reference(const StepParams& P, const CellFields& C, float* t, hipStream_t s) {
    hipLaunchKernelGGL(exact::t_reference_kernel, grid1(P.cells), dim3(256), 0, s, P, C, t);
}

void launch_exact_force(const StepParams& P, const CellFields& C, const float* fpost, double* part, hipStream_t s) {
    const int blocks = (P.cells + 255) / 256;
    hipLaunchKernelGGL(exact::force_partial_kernel, dim3(blocks), dim3(256), 0, s, P, C.flags, fpost, part);
    hipLaunchKernelGGL(exact::force_final_kernel, dim3(1), dim3(256), 0, s, P, (const double*)part, blocks,
                       (float*)(part + 3 * (size_t)blocks));
}

void launch_exact_stream(const StepParams& P, const CellFields& C, const float* fin, float* fout, hipStream_t s) {
    hipLaunchKernelGGL(exact::stream_kernel, grid1(P.cells), dim3(256), 0, s, P, C, fin, fout);
}

void launch_exact_readout(const StepParams& P, const CellFields& C, const float* f, int post_collision, int stride,
                          int out_ch, float scale, float* out, hipStream_t s) {
    const int sx = (P.nx + stride - 1) / stride, sy = (P.ny + stride - 1) / stride, sz = (P.nz + stride - 1) / stride;
    hipLaunchKernelGGL(exact::readout_kernel, grid1(sx * sy * sz), dim3(256), 0, s, P, C, f, post_collision, stride, sx,
                       sy, sz, out_ch, scale, out);
}

void launch_exact_sample(const StepParams& P, const CellFields& C, const float* f, int post_collision, int x, int y,
                         int z, float* out4, hipStream_t s) {
    hipLaunchKernelGGL(exact::sample_kernel, dim3(1), dim3(64), 0, s, P, C, f, post_collision, x, y, z, out4);
}

}  // namespace aero
