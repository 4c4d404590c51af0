/*
 * D3Q19 mesoscale oracle body -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's CPU mesoscale step,
 * native/src/aero_lbm_mesoscale.cpp:110-738 (tables 146-176, seeding 298-333,
 * targets 393-530, step 548-738) and the create/step C ABI (1397-1509).
 * Included twice by meso_oracle.c (R = float in the reference's operation
 * order, R = double for the FP64 oracle). Distributions are AoS
 * (cell*19 + q) as in the reference; y is the vertical axis.
 */

static inline R MS(smax)(R a, R b) { return (a < b) ? b : a; }
static inline R MS(smin)(R a, R b) { return (b < a) ? b : a; }
/* std::clamp(v, lo, hi) = v < lo ? lo : (hi < v ? hi : v) */
static inline R MS(sclamp)(R v, R lo, R hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

typedef struct MS(Ctx) {
    int nx, ny, nz, cells;
    float dx, dt, nu, pr, prt;
    R tau_min;  /* lower clamp of both relaxation times (reference: 0.58) */
    int initialized;
    R *rho, *ambient, *deep, *surface, *wx, *wy, *wz;
    R *f, *f_next, *g, *g_next, *post_f, *post_g, *next_deep, *next_surface;
    R w[19];
} MS(Ctx);

static inline R MS(hfeq)(const MS(Ctx) * c, int q, R rho, R ux, R uy, R uz) {
    const R cu = K(3.0) * ((R)kMCX[q] * ux + (R)kMCY[q] * uy + (R)kMCZ[q] * uz);
    const R uu = ux * ux + uy * uy + uz * uz;
    return c->w[q] * rho * (K(1.0) + cu + K(0.5) * cu * cu - K(1.5) * uu);
}
static inline R MS(tfeq)(const MS(Ctx) * c, int q, R t, R ux, R uy, R uz) {
    const R cu = K(3.0) * ((R)kMCX[q] * ux + (R)kMCY[q] * uy + (R)kMCZ[q] * uz);
    return c->w[q] * t * (K(1.0) + cu);
}
static inline R MS(clamp_speed)(R u) { return MS(sclamp)(u, K(-0.35), K(0.35)); }
static inline R MS(odd_tau)(R even) {
    const R d = MS(smax)(K(0.03), even - K(0.5));
    return MS(sclamp)(K(0.5) + K(0.25) / d, K(0.55), K(2.5));
}
static inline int MS(idx)(const MS(Ctx) * c, int x, int y, int z) { return (x * c->ny + y) * c->nz + z; }
static inline int MS(solid)(const float* fc, int cell) { return fc[cell * 25 + 24] > 0.5f; }
static inline R MS(wall_t)(const float* fc, int cell) { return (R)fc[cell * 25 + 4] + (R)fc[cell * 25 + 23]; }

/* hydro_core.h:122-127 and thermal_core.h:257-262: lattice transport of the domain */
static void MS(transport)(const MS(Ctx) * c, R* vscale, R* tau_h, R* tau_t) {
    const R dx = (R)c->dx, dt = (R)c->dt, nu = (R)c->nu, pr = (R)c->pr;
    *vscale = dx / dt;
    const R nu_l = nu * dt / (dx * dx);
    *tau_h = K(0.5) + K(3.0) * nu_l;
    const R alpha = nu / pr;
    const R alpha_l = alpha * dt / (dx * dx);
    *tau_t = K(0.5) + alpha_l / K(0.25);
}

static void MS(seed)(MS(Ctx) * c, const float* fc, R vscale) {
    for (int i = 0; i < c->cells; ++i) {
        const float* p = fc + (size_t)i * 25;
        const R amb_t = (R)p[2] + (R)p[22];
        const R surf_t = (R)p[4] + (R)p[23];
        const int solid = MS(solid)(fc, i);
        const R bx = (R)p[6] + (R)p[20];
        const R bz = (R)p[7] + (R)p[21];
        const R by = (R)p[19] + (R)p[18];
        const R den = MS(smax)(K(1.0e-6), vscale);
        const R bxl = MS(clamp_speed)((solid ? K(0.0) : bx) / den);
        const R byl = MS(clamp_speed)((solid ? K(0.0) : by) / den);
        const R bzl = MS(clamp_speed)((solid ? K(0.0) : bz) / den);
        c->rho[i] = K(1.0);
        c->ambient[i] = solid ? surf_t : amb_t;
        c->deep[i] = (R)p[3];
        c->surface[i] = surf_t;
        c->wx[i] = solid ? K(0.0) : bx;
        c->wy[i] = solid ? K(0.0) : by;
        c->wz[i] = solid ? K(0.0) : bz;
        for (int q = 0; q < 19; ++q) {
            c->f[i * 19 + q] = MS(hfeq)(c, q, K(1.0), bxl, byl, bzl);
            c->g[i * 19 + q] = MS(tfeq)(c, q, c->ambient[i], bxl, byl, bzl);
        }
    }
}

static R MS(hlap)(const MS(Ctx) * c, const R* fld, int x, int y, int z) {
    const R ce = fld[MS(idx)(c, x, y, z)];
    const R xp = fld[MS(idx)(c, x + 1 < c->nx - 1 ? x + 1 : c->nx - 1, y, z)];
    const R xm = fld[MS(idx)(c, x - 1 > 0 ? x - 1 : 0, y, z)];
    const R zp = fld[MS(idx)(c, x, y, z + 1 < c->nz - 1 ? z + 1 : c->nz - 1)];
    const R zm = fld[MS(idx)(c, x, y, z - 1 > 0 ? z - 1 : 0)];
    return xp + xm + zp + zm - K(4.0) * ce;
}
static R MS(vlap)(const MS(Ctx) * c, const R* fld, int x, int y, int z) {
    if (c->ny <= 1) return K(0.0);
    const R ce = fld[MS(idx)(c, x, y, z)];
    const R yp = fld[MS(idx)(c, x, y + 1 < c->ny - 1 ? y + 1 : c->ny - 1, z)];
    const R ym = fld[MS(idx)(c, x, y - 1 > 0 ? y - 1 : 0, z)];
    return yp + ym - K(2.0) * ce;
}
static R MS(slope)(const MS(Ctx) * c, const float* fc, int x, int y, int z, int axis) {
    const int xp = x + 1 < c->nx - 1 ? x + 1 : c->nx - 1, xm = x - 1 > 0 ? x - 1 : 0;
    const int zp = z + 1 < c->nz - 1 ? z + 1 : c->nz - 1, zm = z - 1 > 0 ? z - 1 : 0;
    if (axis == 0) {
        const R hp = (R)fc[MS(idx)(c, xp, y, z) * 25 + 0], hm = (R)fc[MS(idx)(c, xm, y, z) * 25 + 0];
        return K(0.5) * (hp - hm);
    }
    const R hp = (R)fc[MS(idx)(c, x, y, zp) * 25 + 0], hm = (R)fc[MS(idx)(c, x, y, zm) * 25 + 0];
    return K(0.5) * (hp - hm);
}

typedef struct { R rho, ux, uy, uz, teq, nsurf, ndeep; } MS(Targets);

/* aero_lbm_mesoscale.cpp:393-530 */
static MS(Targets) MS(targets)(const MS(Ctx) * c, const float* fc, R vscale, int x, int y, int z) {
    const int i = MS(idx)(c, x, y, z);
    const float* p = fc + (size_t)i * 25;
    const R dt = MS(smax)(K(1.0e-3), (R)c->dt);
    const R amb_relax = MS(sclamp)(dt / K(600.0), K(0.0), K(1.0));
    const R deep_relax = MS(sclamp)(dt / K(2400.0), K(0.0), K(1.0));
    const R surf_relax = MS(sclamp)(dt / K(300.0), K(0.0), K(1.0));
    const R wind_relax = MS(sclamp)(dt / K(60.0), K(0.0), K(1.0));
    R rho = K(0.0), ux = K(0.0), uy = K(0.0), uz = K(0.0), amb = K(0.0);
    for (int q = 0; q < 19; ++q) {
        const R fq = c->f[i * 19 + q];
        rho += fq;
        ux += fq * (R)kMCX[q];
        uy += fq * (R)kMCY[q];
        uz += fq * (R)kMCZ[q];
        amb += c->g[i * 19 + q];
    }
    rho = MS(smax)(K(1.0e-6), rho);
    ux /= rho;
    uy /= rho;
    uz /= rho;
    const R amb_t = (R)p[2] + (R)p[22];
    const R deep_t = (R)p[3];
    const R surf_t = (R)p[4] + (R)p[23];
    const R rough = MS(smax)(K(0.0), (R)p[5]);
    const R conv_heat = (R)p[10], conv_moist = (R)p[11], conv_in_x = (R)p[12], conv_in_z = (R)p[13];
    const R tor_x = (R)p[14], tor_z = (R)p[15], tor_heat = (R)p[16], tor_moist = (R)p[17], tor_up = (R)p[18];
    const R nest_up = (R)p[19], nest_dx = (R)p[20], nest_dz = (R)p[21];
    const R bx = (R)p[6] + nest_dx + conv_in_x + tor_x;
    const R bz = (R)p[7] + nest_dz + conv_in_z + tor_z;
    const R by = nest_up + tor_up;
    MS(Targets) o;
    if (MS(solid)(fc, i)) {
        o.rho = K(1.0); o.ux = o.uy = o.uz = K(0.0);
        o.teq = surf_t; o.nsurf = surf_t; o.ndeep = deep_t;
        return o;
    }
    const R hum = MS(sclamp)((R)p[9] + conv_moist + tor_moist, K(0.0), K(1.0));
    const R sx = MS(slope)(c, fc, x, y, z, 0);
    const R sz = MS(slope)(c, fc, x, y, z, 1);
    const R s_hl = MS(hlap)(c, c->surface, x, y, z);
    const R s_vl = MS(vlap)(c, c->surface, x, y, z);
    const R d_vl = MS(vlap)(c, c->deep, x, y, z);
    const R a_hl = MS(hlap)(c, c->ambient, x, y, z);
    const R a_vl = MS(vlap)(c, c->ambient, x, y, z);
    const R ndeep = c->deep[i] + deep_relax * (deep_t - c->deep[i]) + K(0.01) * d_vl;
    const R damp = K(1.0) - K(0.12) * hum;
    const R nsurf = c->surface[i] + surf_relax * damp * (surf_t - c->surface[i]) + K(0.10) * (ndeep - c->surface[i]) +
                    K(0.04) * conv_heat + K(0.08) * tor_heat + K(0.05) * s_hl + K(0.025) * s_vl;
    const R den = MS(smax)(K(1.0e-6), vscale);
    const R bxl = MS(clamp_speed)(bx / den);
    const R byl = MS(clamp_speed)(by / den);
    const R bzl = MS(clamp_speed)(bz / den);
    const R push = K(0.0008) * (nsurf - amb);
    const R nsw = K(1.0) - (R)y / MS(smax)(K(1.0), (R)(c->ny - 1));
    const R lift = K(0.0015) * (bxl * sx + bzl * sz) * nsw;
    const R blift = K(0.0009) * (nsurf - amb) * nsw;
    const R steer_x = K(-0.005) * sx * bzl;
    const R steer_z = K(0.005) * sz * bxl;
    const R drag = MS(sclamp)(rough * K(0.02), K(0.0), K(0.25));
    const R tux = MS(clamp_speed)(bxl + steer_x + push * sx - drag * ux);
    const R tuz = MS(clamp_speed)(bzl + steer_z + push * sz - drag * uz);
    const R tuy = MS(clamp_speed)(byl + lift + blift - drag * uy);
    o.ux = MS(clamp_speed)((K(1.0) - wind_relax) * ux + wind_relax * tux);
    o.uy = MS(clamp_speed)((K(1.0) - wind_relax) * uy + wind_relax * tuy);
    o.uz = MS(clamp_speed)((K(1.0) - wind_relax) * uz + wind_relax * tuz);
    const R src = amb_relax * (amb_t - amb) + (K(0.018) + K(0.004) * hum) * (nsurf - amb) + K(0.01) * (ndeep - amb) +
                  K(0.08) * conv_heat * (K(0.5) + K(0.5) * hum) + K(0.10) * tor_heat * (K(0.45) + K(0.55) * hum) +
                  K(0.06) * tor_up + K(0.06) * nest_up + K(0.01) * a_hl + K(0.02) * a_vl;
    o.rho = rho;
    o.teq = amb + src;
    o.nsurf = nsurf;
    o.ndeep = ndeep;
    return o;
}

/* aero_lbm_mesoscale.cpp:548-738 */
static void MS(step)(MS(Ctx) * c, const float* fc) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    R vscale, tau_h0, tau_t0;
    MS(transport)(c, &vscale, &tau_h0, &tau_t0);
    const R tau_h = MS(sclamp)(tau_h0, c->tau_min, K(2.5));
    const R tau_ho = MS(odd_tau)(tau_h);
    const R oh_e = K(1.0) / tau_h, oh_o = K(1.0) / tau_ho;
    const R tau_t = MS(sclamp)(tau_t0, c->tau_min, K(2.5));
    const R tau_to = MS(odd_tau)(tau_t);
    const R ot_e = K(1.0) / tau_t, ot_o = K(1.0) / tau_to;
    const R max_speed = K(0.35) * vscale;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int i = MS(idx)(c, x, y, z);
                const MS(Targets) t = MS(targets)(c, fc, vscale, x, y, z);
                c->next_deep[i] = t.ndeep;
                c->next_surface[i] = t.nsurf;
                for (int q = 0; q < 19; ++q) {
                    const int qo = kMOPP[q];
                    if (q > qo) continue;
                    const int d = i * 19 + q;
                    if (q == qo) {
                        const R fe = MS(hfeq)(c, q, t.rho, t.ux, t.uy, t.uz);
                        const R ge = MS(tfeq)(c, q, t.teq, t.ux, t.uy, t.uz);
                        c->post_f[d] = c->f[d] - oh_e * (c->f[d] - fe);
                        c->post_g[d] = c->g[d] - ot_e * (c->g[d] - ge);
                        continue;
                    }
                    const int od = i * 19 + qo;
                    const R fe = MS(hfeq)(c, q, t.rho, t.ux, t.uy, t.uz);
                    const R feo = MS(hfeq)(c, qo, t.rho, t.ux, t.uy, t.uz);
                    const R ge = MS(tfeq)(c, q, t.teq, t.ux, t.uy, t.uz);
                    const R geo = MS(tfeq)(c, qo, t.teq, t.ux, t.uy, t.uz);
                    const R f_ev = K(0.5) * (c->f[d] + c->f[od]);
                    const R f_od = K(0.5) * (c->f[d] - c->f[od]);
                    const R fe_ev = K(0.5) * (fe + feo);
                    const R fe_od = K(0.5) * (fe - feo);
                    const R f_evp = f_ev - oh_e * (f_ev - fe_ev);
                    const R f_odp = f_od - oh_o * (f_od - fe_od);
                    c->post_f[d] = f_evp + f_odp;
                    c->post_f[od] = f_evp - f_odp;
                    const R g_ev = K(0.5) * (c->g[d] + c->g[od]);
                    const R g_od = K(0.5) * (c->g[d] - c->g[od]);
                    const R ge_ev = K(0.5) * (ge + geo);
                    const R ge_od = K(0.5) * (ge - geo);
                    const R g_evp = g_ev - ot_e * (g_ev - ge_ev);
                    const R g_odp = g_od - ot_o * (g_od - ge_od);
                    c->post_g[d] = g_evp + g_odp;
                    c->post_g[od] = g_evp - g_odp;
                }
            }
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int i = MS(idx)(c, x, y, z);
                const float* p = fc + (size_t)i * 25;
                if (MS(solid)(fc, i)) {
                    const R wt = MS(wall_t)(fc, i);
                    for (int q = 0; q < 19; ++q) {
                        c->f_next[i * 19 + q] = MS(hfeq)(c, q, K(1.0), K(0.0), K(0.0), K(0.0));
                        c->g_next[i * 19 + q] = MS(tfeq)(c, q, wt, K(0.0), K(0.0), K(0.0));
                    }
                    continue;
                }
                const R den = MS(smax)(K(1.0e-6), vscale);
                const R bbx = MS(clamp_speed)(((R)p[6] + (R)p[20] + (R)p[12] + (R)p[14]) / den);
                const R bbz = MS(clamp_speed)(((R)p[7] + (R)p[21] + (R)p[13] + (R)p[15]) / den);
                const R btemp = (R)p[2] + (R)p[22];
                const R bby = MS(clamp_speed)(((R)p[19] + (R)p[18]) / den);
                int ed = x < z ? x : z;
                const int ed2 = (nx - 1 - x) < (nz - 1 - z) ? (nx - 1 - x) : (nz - 1 - z);
                ed = ed < ed2 ? ed : ed2;
                const int top = ny - 1 - y;
                const int sd = ed < top ? ed : top;
                const R ea = sd <= 0 ? K(1.0) : (sd < 2 ? K(0.5) : K(0.0));
                for (int q = 0; q < 19; ++q) {
                    const int sx = x - kMCX[q], sy = y - kMCY[q], sz = z - kMCZ[q];
                    const int d = i * 19 + q;
                    R fv, gv;
                    int bounced = 0;
                    if (sx >= 0 && sx < nx && sy >= 0 && sy < ny && sz >= 0 && sz < nz) {
                        const int src = MS(idx)(c, sx, sy, sz);
                        if (MS(solid)(fc, src)) {
                            fv = c->post_f[i * 19 + kMOPP[q]];
                            gv = MS(tfeq)(c, q, MS(wall_t)(fc, src), K(0.0), K(0.0), K(0.0));
                            bounced = 1;
                        } else {
                            fv = c->post_f[src * 19 + q];
                            gv = c->post_g[src * 19 + q];
                        }
                    } else if (sy < 0) {
                        fv = c->post_f[i * 19 + kMOPP[q]];
                        gv = MS(tfeq)(c, q, MS(wall_t)(fc, i), K(0.0), K(0.0), K(0.0));
                        bounced = 1;
                    } else {
                        fv = MS(hfeq)(c, q, K(1.0), bbx, bby, bbz);
                        gv = MS(tfeq)(c, q, btemp, bbx, bby, bbz);
                    }
                    if (ea > K(0.0) && !bounced) {
                        const R fb = MS(hfeq)(c, q, K(1.0), bbx, bby, bbz);
                        const R gb = MS(tfeq)(c, q, btemp, bbx, bby, bbz);
                        fv = (K(1.0) - ea) * fv + ea * fb;
                        gv = (K(1.0) - ea) * gv + ea * gb;
                    }
                    c->f_next[d] = fv;
                    c->g_next[d] = gv;
                }
            }
    R* s;
    s = c->f; c->f = c->f_next; c->f_next = s;
    s = c->g; c->g = c->g_next; c->g_next = s;
    s = c->deep; c->deep = c->next_deep; c->next_deep = s;
    s = c->surface; c->surface = c->next_surface; c->next_surface = s;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->cells; ++i) {
        if (MS(solid)(fc, i)) {
            c->rho[i] = K(1.0);
            c->ambient[i] = MS(wall_t)(fc, i);
            c->wx[i] = c->wy[i] = c->wz[i] = K(0.0);
            continue;
        }
        R rho = K(0.0), ux = K(0.0), uy = K(0.0), uz = K(0.0), amb = K(0.0);
        for (int q = 0; q < 19; ++q) {
            const R fq = c->f[i * 19 + q];
            rho += fq;
            ux += fq * (R)kMCX[q];
            uy += fq * (R)kMCY[q];
            uz += fq * (R)kMCZ[q];
            amb += c->g[i * 19 + q];
        }
        rho = MS(smax)(K(1.0e-6), rho);
        ux /= rho;
        uy /= rho;
        uz /= rho;
        c->rho[i] = rho;
        c->ambient[i] = amb;
        c->wx[i] = MS(sclamp)(ux * vscale, -max_speed, max_speed);
        c->wy[i] = MS(sclamp)(uy * vscale, -max_speed, max_speed);
        c->wz[i] = MS(sclamp)(uz * vscale, -max_speed, max_speed);
    }
}

ORACLE_API void* MS(meso_create)(int nx, int ny, int nz, float dx, float dt, float nu, float pr, float prt, float tau_min) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return NULL;
    MS(Ctx)* c = (MS(Ctx)*)calloc(1, sizeof(MS(Ctx)));
    c->nx = nx; c->ny = ny; c->nz = nz; c->cells = nx * ny * nz;
    c->dx = dx; c->dt = dt; c->nu = nu; c->pr = pr; c->prt = prt;
    c->tau_min = (R)tau_min;
    const R w0 = K(1.0) / K(3.0), w1 = K(1.0) / K(18.0), w2 = K(1.0) / K(36.0);
    for (int q = 0; q < 19; ++q) c->w[q] = q == 0 ? w0 : (q < 7 ? w1 : w2);
    const size_t n = (size_t)c->cells;
    R** fields[] = {&c->rho, &c->ambient, &c->deep, &c->surface, &c->wx, &c->wy, &c->wz, &c->next_deep, &c->next_surface};
    for (size_t k = 0; k < sizeof(fields) / sizeof(fields[0]); ++k) *fields[k] = (R*)calloc(n, sizeof(R));
    R** dists[] = {&c->f, &c->f_next, &c->g, &c->g_next, &c->post_f, &c->post_g};
    for (size_t k = 0; k < sizeof(dists) / sizeof(dists[0]); ++k) *dists[k] = (R*)calloc(n * 19, sizeof(R));
    return c;
}

ORACLE_API void MS(meso_destroy)(void* h) {
    MS(Ctx)* c = (MS(Ctx)*)h;
    if (!c) return;
    R* all[] = {c->rho, c->ambient, c->deep, c->surface, c->wx, c->wy, c->wz, c->next_deep, c->next_surface,
                c->f, c->f_next, c->g, c->g_next, c->post_f, c->post_g};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) free(all[k]);
    free(c);
}

/* aero_lbm_mesoscale_step_context (1452-1498): seed on first use, step, export */
ORACLE_API int MS(meso_step)(void* h, const float* forcing, float* out_state) {
    MS(Ctx)* c = (MS(Ctx)*)h;
    if (!c || !forcing) return 0;
    if (!c->initialized) {
        R vscale, th, tt;
        MS(transport)(c, &vscale, &th, &tt);
        MS(seed)(c, forcing, vscale);
        c->initialized = 1;
    }
    MS(step)(c, forcing);
    if (out_state) {
        for (int i = 0; i < c->cells; ++i) {
            float* o = out_state + (size_t)i * 6;
            o[0] = (float)c->ambient[i]; o[1] = (float)c->deep[i]; o[2] = (float)c->surface[i];
            o[3] = (float)c->wx[i]; o[4] = (float)c->wy[i]; o[5] = (float)c->wz[i];
        }
    }
    return 1;
}

ORACLE_API void MS(meso_get_f)(void* h, R* f, R* g) {
    MS(Ctx)* c = (MS(Ctx)*)h;
    memcpy(f, c->f, (size_t)c->cells * 19 * sizeof(R));
    memcpy(g, c->g, (size_t)c->cells * 19 * sizeof(R));
}

/* test hooks: overwrite the distributions (marks the context initialised) and
 * read the macroscopic fields and the derived transport */
ORACLE_API void MS(meso_set_f)(void* h, const R* f, const R* g) {
    MS(Ctx)* c = (MS(Ctx)*)h;
    memcpy(c->f, f, (size_t)c->cells * 19 * sizeof(R));
    memcpy(c->g, g, (size_t)c->cells * 19 * sizeof(R));
    c->initialized = 1;
}

ORACLE_API void MS(meso_get_fields)(void* h, R* ambient, R* surface, R* deep) {
    MS(Ctx)* c = (MS(Ctx)*)h;
    memcpy(ambient, c->ambient, (size_t)c->cells * sizeof(R));
    memcpy(surface, c->surface, (size_t)c->cells * sizeof(R));
    memcpy(deep, c->deep, (size_t)c->cells * sizeof(R));
}

ORACLE_API void MS(meso_transport)(void* h, R* out3) {
    MS(Ctx)* c = (MS(Ctx)*)h;
    MS(transport)(c, &out3[0], &out3[1], &out3[2]);
}
