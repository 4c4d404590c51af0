// FP64 forward-mode (dual-number) directional derivative of the adjoint model's
// loss -- TEST INFRASTRUCTURE ONLY, the checker of the hand-written reverse pass.
// It runs the model of aerodynamics4mc-fabric_amd/csrc/adjoint_core.h on the host
// with every state value a Dual<double>: the tangent seeded on the parameters
// along a direction (d_inflow, d_heat) propagates through every step, so the
// final tangent of the loss is dL/dp . d, with no finite-difference truncation.
// Host loops restate the device launches of adjoint.hip (initial state, source
// upsampling, forward step).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../aerodynamics4mc-fabric_amd/csrc/adjoint_core.h"

using namespace aero::adj;
using D = Dual<double>;

extern "C" int ad_dual_directional(int n, int m, int steps, double nu, float beta, float alpha, const float* u0,
                                   const uint8_t* solid, int nprobe, const int* probes, const double* target,
                                   const double* inflow, const double* heat, const double* d_inflow,
                                   const double* d_heat, double* out_loss, double* out_dloss) {
    Params P{};
    P.nx = P.ny = P.nz = n;
    P.tau_shear = P.tau_normal = 0.5 + 3.0 * nu;
    P.beta = beta;
    P.g_axis = 2;
    P.g_sign = -1;
    P.alpha = alpha;
    P.kmax = 0.57735026919f * 0.60f;
    P.rho_in = 1.0f;
    for (int k = 0; k < 3; ++k) P.u0[k] = u0[k];
    const int N = n * n * n;
    std::vector<D> uin(3 * (size_t)n * n), src(N);
    for (size_t i = 0; i < uin.size(); ++i) uin[i] = D(inflow[i], d_inflow[i]);
    // heat source: trilinear sample of the m^3 field at the cell centre, clamped
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
            for (int z = 0; z < n; ++z) {
                const int xyz[3] = {x, y, z};
                int lo[3];
                float t[3];
                for (int a = 0; a < 3; ++a) {
                    float p;
                    source_weights(n, m, xyz[a], &p);
                    lo[a] = std::min((int)std::floor(p), std::max(0, m - 2));
                    t[a] = p - (float)lo[a];
                }
                D v(0.0);
                for (int k = 0; k < 8; ++k) {
                    const int i = std::min(lo[0] + (k >> 2), m - 1), j = std::min(lo[1] + ((k >> 1) & 1), m - 1),
                              l = std::min(lo[2] + (k & 1), m - 1);
                    const float w = ((k >> 2) ? t[0] : 1.0f - t[0]) * (((k >> 1) & 1) ? t[1] : 1.0f - t[1]) *
                                    ((k & 1) ? t[2] : 1.0f - t[2]);
                    const size_t h = (size_t)(i * m + j) * m + l;
                    v = v + D((double)w) * D(heat[h], d_heat[h]);
                }
                src[cidx(P, x, y, z)] = rclamp(v, -0.006f, 0.006f);
            }
    std::vector<D> s0((size_t)31 * N), s1((size_t)31 * N);
    for (int c = 0; c < N; ++c) {
        const bool sol = solid[c] != 0;
        const D ux(sol ? 0.0 : (double)u0[0]), uy(sol ? 0.0 : (double)u0[1]), uz(sol ? 0.0 : (double)u0[2]);
        for (int q = 0; q < 27; ++q) s0[(size_t)q * N + c] = feq<D>(q, D(1.0), ux, uy, uz);
        s0[(size_t)27 * N + c] = D(0.0);
        s0[(size_t)28 * N + c] = ux;
        s0[(size_t)29 * N + c] = uy;
        s0[(size_t)30 * N + c] = uz;
    }
    for (int st = 0; st < steps; ++st) {
        const D* si = s0.data();
        D* so = s1.data();
#pragma omp parallel for schedule(static)
        for (int c = 0; c < N; ++c) {
            if (solid[c]) {
                for (int q = 0; q < 28; ++q) so[(size_t)q * N + c] = si[(size_t)q * N + c];
                for (int k = 0; k < 3; ++k) so[(size_t)(28 + k) * N + c] = D(0.0);
                continue;
            }
            const int x = c / (n * n), y = (c / n) % n, z = c % n;
            int idx[14];
            D w[8], pre;
            const D Tn = temperature(P, si + (size_t)27 * N, solid, src[c], x, y, z, si[(size_t)28 * N + c],
                                     si[(size_t)29 * N + c], si[(size_t)30 * N + c], idx, w, &pre);
            D g[27], u[3];
            // the initial populations are collided where they are (no stream before the first collision)
            for (int q = 0; q < 27; ++q)
                g[q] = st == 0 ? si[(size_t)q * N + c] : incoming(P, si, solid, uin.data(), x, y, z, q);
            collide(P, g, Tn, u);
            for (int q = 0; q < 27; ++q) so[(size_t)q * N + c] = g[q];
            so[(size_t)27 * N + c] = Tn;
            for (int k = 0; k < 3; ++k) so[(size_t)(28 + k) * N + c] = u[k];
        }
        s0.swap(s1);
    }
    D L(0.0);
    for (int i = 0; i < nprobe; ++i)
        for (int k = 0; k < 3; ++k) {
            const D d = s0[(size_t)(28 + k) * N + probes[i]] - D(target[3 * i + k]);
            L = L + d * d;
        }
    *out_loss = L.v;
    *out_dloss = L.d;
    return 1;
}

// The hand reverse mode of the low-moment collision (collide_low_adjoint) against
// forward-mode dual numbers on the same map, in FP64: for inputs m[10], T and output
// adjoints fbar[27], ubar[3], writes the reverse result (mbar_rev[10], tbar_rev) and
// the one assembled from 11 dual evaluations (mbar_dual[10], tbar_dual).
extern "C" void ad_collide_vjp(double nu, float beta, int g_axis, int g_sign, float kmax, const double* m,
                               double T, const double* fbar, const double* ubar, double* mbar_rev,
                               double* tbar_rev, double* mbar_dual, double* tbar_dual) {
    Params P{};
    P.nx = P.ny = P.nz = 4;
    P.tau_shear = P.tau_normal = 0.5 + 3.0 * nu;
    P.beta = beta;
    P.g_axis = g_axis;
    P.g_sign = g_sign;
    P.kmax = kmax;
    P.rho_in = 1.0f;
    collide_low_adjoint<double>(P, m, T, fbar, ubar, mbar_rev, tbar_rev);
    for (int i = 0; i <= kLow; ++i) {
        D md[kLow], fd[27], ud[3];
        for (int k = 0; k < kLow; ++k) md[k] = D(m[k], k == i ? 1.0 : 0.0);
        collide_low(P, md, D(T, i == kLow ? 1.0 : 0.0), fd, ud);
        double acc = 0.0;
        for (int q = 0; q < 27; ++q) acc += fd[q].d * fbar[q];
        for (int k = 0; k < 3; ++k) acc += ud[k].d * ubar[k];
        if (i < kLow) mbar_dual[i] = acc;
        else *tbar_dual = acc;
    }
}
