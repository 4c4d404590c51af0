/*
 * D3Q27 oracle body -- TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's
 * cpu_baseline). Never linked into or called by libaero_lbm.
 *
 * Plain-C restatement of the reference's single-threaded CPU step in benchmark
 * mode, native/src/aero_lbm_jni.cpp (run_cpu_step, 8480-8507): ingest ->
 * initialise (first step) -> temperature update -> collide -> stream -> output.
 * Included twice by d3q27_oracle.c: once with R = float (operation order of the
 * reference, so FP32 results are bit-identical to the reference CPU path when
 * built with -ffp-contract=off on x86-64 SSE), once with R = double (the FP64
 * oracle used for accuracy claims).
 *
 * Expected macros: R, SFX(name), K(literal), RSQRT_(x) sqrt, RFLOOR_(x),
 * RFABS_(x), RISFINITE_(x).
 */

/* std::max / std::min / clampf as the reference defines them
 * (jni.cpp:1193-1195): min(a,b) = (b<a)?b:a, max(a,b) = (a<b)?b:a. */
static inline R SFX(rmax)(R a, R b) { return (a < b) ? b : a; }
static inline R SFX(rmin)(R a, R b) { return (b < a) ? b : a; }
static inline R SFX(clampr)(R v, R lo, R hi) { return SFX(rmin)(hi, SFX(rmax)(lo, v)); }
static inline R SFX(finite_or)(R v, R fb) { return RISFINITE_(v) ? v : fb; }

/* jni.cpp:46-50 kMomentInv1D */
static const R SFX(kInv1D)[3][3] = {
    {K(0.0), K(-0.5), K(0.5)},
    {K(1.0), K(0.0), K(-1.0)},
    {K(0.0), K(0.5), K(0.5)},
};

typedef struct SFX(Ctx) {
    int nx, ny, nz;
    size_t cells;
    int in_ch, out_ch;
    AeroLbmBenchmarkConfig cfg;
    /* resolved physics */
    R tau_shear_base, tau_normal_base, nu_base, alpha, beta;
    R sgs_c2, nut_cap_ratio;
    int g_axis, g_sign;
    int initialized;
    uint64_t step_counter;
    R *f, *f_post;
    R *rho, *ux, *uy, *uz;
    R *ref_ux, *ref_uy, *ref_uz, *ref_p, *ref_T, *src;
    unsigned char* obstacle;
    R *T, *T_next, *T_scr;
    int bfecc; /* BFECC thermal advection (heated-cavity presets; runtime mode) */
    /* runtime (non-benchmark) mode, jni.cpp:675-681: state and temperature nudge near
     * the faces, fan forcing, turbulent thermal diffusivity, temperature every
     * second step with dt 2 */
    int runtime;
    int thermal_stride;
    R nu_rt0;   /* runtime molecular viscosity (lattice) */
    R *fan_mask, *fan_ux, *fan_uy, *fan_uz;
    int cylinder; /* cylinder crossflow preset: parabolic inlet, drag integral */
    float force[3]; /* last step's drag integral (ctx.last_force) */
    R w[27];
    R kmax_speed;
    /* boundary-reference shell (jni.cpp:4535-4552, 6449-6522): per-cell layer, FP16-
     * rounded targets (u_x, u_y, u_z, p) decoded to R, depth and outer relaxation */
    unsigned char* shell_layer;
    R* shell_target;
    int shell_layers;
    R shell_relax;
} SFX(Ctx);

/* hydro_core.h:130-205 (tables): q = (ix*3+iy)*3+iz, c = ix-1, ..., w = product of
 * the 1-D weights {1/6, 2/3, 1/6} evaluated left to right in R. */
static void SFX(init_weights)(R w[27]) {
    const R w1[3] = {K(1.0) / K(6.0), K(2.0) / K(3.0), K(1.0) / K(6.0)};
    int q = 0;
    for (int ix = 0; ix < 3; ++ix)
        for (int iy = 0; iy < 3; ++iy)
            for (int iz = 0; iz < 3; ++iz) w[q++] = w1[ix] * w1[iy] * w1[iz];
}

/* hydro_core.cpp:364-368 equilibrium_d3q27 */
static inline R SFX(feq)(const SFX(Ctx) * c, int q, R rho, R ux, R uy, R uz) {
    const R cu = K(3.0) * ((R)kCX[q] * ux + (R)kCY[q] * uy + (R)kCZ[q] * uz);
    const R uu = ux * ux + uy * uy + uz * uz;
    return c->w[q] * rho * (K(1.0) + cu + K(0.5) * cu * cu - K(1.5) * uu);
}

static inline size_t SFX(cell_index)(const SFX(Ctx) * c, int x, int y, int z) {
    return ((size_t)x * (size_t)c->ny + (size_t)y) * (size_t)c->nz + (size_t)z;
}

static inline const AeroLbmBoundaryFaceConfig* SFX(face)(const SFX(Ctx) * c, int idx) {
    switch (idx) {
        case 0: return &c->cfg.x_min;
        case 1: return &c->cfg.x_max;
        case 2: return &c->cfg.y_min;
        case 3: return &c->cfg.y_max;
        case 4: return &c->cfg.z_min;
        default: return &c->cfg.z_max;
    }
}

/* jni.cpp:729-741 wrap_axis_periodic */
static int SFX(wrap_axis)(const SFX(Ctx) * c, int* coord, int dim, int min_face, int max_face, int thermal) {
    if (*coord >= 0 && *coord < dim) return 0;
    const AeroLbmBoundaryFaceConfig* f = SFX(face)(c, *coord < 0 ? min_face : max_face);
    const int kind = thermal ? f->thermal_kind : f->hydrodynamic_kind;
    const int periodic = thermal ? AERO_LBM_THERMAL_BOUNDARY_PERIODIC : AERO_LBM_HYDRO_BOUNDARY_PERIODIC;
    if (kind != periodic) return 0;
    *coord %= dim;
    if (*coord < 0) *coord += dim;
    return 1;
}

static void SFX(remap)(const SFX(Ctx) * c, int* x, int* y, int* z, int thermal) {
    SFX(wrap_axis)(c, x, c->nx, 0, 1, thermal);
    SFX(wrap_axis)(c, y, c->ny, 2, 3, thermal);
    SFX(wrap_axis)(c, z, c->nz, 4, 5, thermal);
}

/* jni.cpp:743-761 face lookup order x, y, z */
static const AeroLbmBoundaryFaceConfig* SFX(face_for_oob)(const SFX(Ctx) * c, int x, int y, int z) {
    if (x < 0) return &c->cfg.x_min;
    if (x >= c->nx) return &c->cfg.x_max;
    if (y < 0) return &c->cfg.y_min;
    if (y >= c->ny) return &c->cfg.y_max;
    if (z < 0) return &c->cfg.z_min;
    if (z >= c->nz) return &c->cfg.z_max;
    return NULL;
}

static inline int SFX(flag)(const SFX(Ctx) * c, uint32_t bit) {
    return c->cfg.enabled != 0 && (c->cfg.flags & bit) != 0;
}

/* ---------------- parameter resolution (jni.cpp:787-993) ---------------- */
static R SFX(clamp_nu)(R nu) {
    const R nu_min = K(1.0e-8);
    const R nu_max = (K(0.95) - K(0.5)) / K(3.0);
    return SFX(clampr)(nu, nu_min, nu_max);
}
static R SFX(tau_from_nu)(R nu) { return K(0.5) + K(3.0) * SFX(clamp_nu)(nu); }

static int SFX(heated_cavity)(const AeroLbmBenchmarkConfig* g) {
    /* jni.cpp:865-869 */
    return g->preset == AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_2D ||
           g->preset == AERO_LBM_BENCHMARK_PRESET_DIFFERENTIALLY_HEATED_CAVITY_3D;
}

static R SFX(finite_or_f)(float v, R fb) { return isfinite(v) ? (R)v : fb; }

static R SFX(reference_speed)(const AeroLbmBenchmarkConfig* g) {
    /* jni.cpp:823-859 */
    if (SFX(heated_cavity)(g)) {
        const R ma = SFX(clampr)(SFX(finite_or_f)(g->mach_number, K(0.05)), K(1.0e-4), K(0.30));
        return ma * RSQRT_(K(1.0) / K(3.0));
    }
    const float* iv = g->initial_velocity;
    R speed = RSQRT_((R)iv[0] * (R)iv[0] + (R)iv[1] * (R)iv[1] + (R)iv[2] * (R)iv[2]);
    if (g->preset == AERO_LBM_BENCHMARK_PRESET_CYLINDER_CROSSFLOW_2D) {
        if (speed > K(1.0e-8)) return speed;
        const float* v = g->x_min.velocity;
        const R fs = RSQRT_((R)v[0] * (R)v[0] + (R)v[1] * (R)v[1] + (R)v[2] * (R)v[2]);
        return (K(2.0) / K(3.0)) * fs;
    }
    const AeroLbmBoundaryFaceConfig* faces[6] = {&g->x_min, &g->x_max, &g->y_min, &g->y_max, &g->z_min, &g->z_max};
    for (int i = 0; i < 6; ++i) {
        const float* v = faces[i]->velocity;
        const R fs = RSQRT_((R)v[0] * (R)v[0] + (R)v[1] * (R)v[1] + (R)v[2] * (R)v[2]);
        speed = SFX(rmax)(speed, fs);
    }
    return speed;
}

/* jni.cpp:881-899: span of the Dirichlet face temperatures (1 without any) */
static R SFX(temperature_span)(const AeroLbmBenchmarkConfig* g) {
    const AeroLbmBoundaryFaceConfig* faces[6] = {&g->x_min, &g->x_max, &g->y_min, &g->y_max, &g->z_min, &g->z_max};
    R lo = K(1.0e9), hi = K(-1.0e9);
    int any = 0;
    for (int i = 0; i < 6; ++i) {
        if (faces[i]->thermal_kind != AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET) continue;
        any = 1;
        lo = SFX(rmin)(lo, (R)faces[i]->temperature);
        hi = SFX(rmax)(hi, (R)faces[i]->temperature);
    }
    if (!any) return K(1.0);
    return SFX(rmax)(K(1.0e-6), hi - lo);
}

static int SFX(resolve)(SFX(Ctx) * c, const AeroLbmSolverOverrides* ov) {
    const AeroLbmBenchmarkConfig* g = &c->cfg;
    /* runtime molecular nu0 = 1.5e-5 m^2/s * 0.05 s / 1 m^2 (jni.cpp:100-104) */
    c->nu_rt0 = SFX(clamp_nu)(K(1.50e-5) * (K(1.0) / K(20.0)) / (K(1.0) * K(1.0)));
    c->thermal_stride = 1;
    if (!g->enabled) {
        /* runtime mode (jni.cpp:675-681, 896-993): molecular viscosity and thermal
         * diffusivity of air at 1 m cells and 0.05 s steps, tau_normal = tau(nu0),
         * BFECC every second step (dt 2), Boussinesq beta of thermal_core.h:268-279 */
        c->runtime = 1;
        c->nu_base = c->nu_rt0;
        const R tau_min = K(0.5) + K(3.0) * K(1.0e-8);
        c->tau_shear_base = SFX(clampr)(SFX(tau_from_nu)(c->nu_base), tau_min, K(0.95));
        c->tau_normal_base = SFX(clampr)(K(0.5) + K(3.0) * (K(1.50e-5) * (K(1.0) / K(20.0)) / (K(1.0) * K(1.0))),
                                         tau_min, K(0.95));
        const R alpha_phys = K(1.50e-5) / K(0.71); /* thermal_core.h:22-24 nu / Pr */
        c->alpha = SFX(rmax)(K(1.0e-8), alpha_phys * (K(1.0) / K(20.0)) / (K(1.0) * K(1.0)));
        c->beta = ((K(9.81) * (K(1.0) / K(300.0))) * K(20.0)) * (((K(1.0) / K(20.0)) * (K(1.0) / K(20.0))) / K(1.0));
        c->bfecc = 1;
        c->thermal_stride = 2;
        c->cylinder = 0;
        R sgs_c = K(0.025);
        c->nut_cap_ratio = K(5.0);
        c->g_axis = 1;
        c->g_sign = -1;
        if (ov) {
            if (!isnan(ov->sgs_constant)) sgs_c = (R)ov->sgs_constant;
            if (!isnan(ov->sgs_nut_cap_ratio)) c->nut_cap_ratio = (R)ov->sgs_nut_cap_ratio;
            if (!isnan(ov->boussinesq_beta)) c->beta = (R)ov->boussinesq_beta;
            if (!isnan(ov->thermal_diffusivity)) c->alpha = (R)ov->thermal_diffusivity;
            if (ov->gravity_axis >= 0 && ov->gravity_axis <= 2) c->g_axis = ov->gravity_axis;
            if (ov->gravity_sign == 1 || ov->gravity_sign == -1) c->g_sign = ov->gravity_sign;
        }
        c->sgs_c2 = sgs_c * sgs_c;
        c->kmax_speed = K(0.57735026919) * K(0.60);
        return 1;
    }
    /* jni.cpp:917-929 effective_base_nu_shear */
    const R re = SFX(finite_or_f)(g->reynolds_number, K(0.0));
    const R ref_speed = SFX(reference_speed)(g);
    const R ref_len = SFX(rmax)(K(1.0e-6), SFX(finite_or_f)(g->reference_length, K(1.0)));
    const R nu_rt = c->nu_rt0;
    const int heated = SFX(heated_cavity)(g);
    const R ra = SFX(rmax)(K(1.0e-6), SFX(finite_or_f)(g->rayleigh_number, K(1.0e5)));
    const R pr = SFX(rmax)(K(1.0e-6), SFX(finite_or_f)(g->prandtl_number, K(0.71)));
    R nu;
    if (heated) nu = SFX(clamp_nu)(ref_speed * ref_len * RSQRT_(pr / ra)); /* jni.cpp:937-943 */
    else if (re <= K(1.0e-6) || ref_speed <= K(1.0e-8)) nu = nu_rt;
    else nu = SFX(clamp_nu)(ref_speed * ref_len / re);
    c->nu_base = nu;
    const R tau_min = K(0.5) + K(3.0) * K(1.0e-8);
    c->tau_shear_base = SFX(clampr)(SFX(tau_from_nu)(nu), tau_min, K(0.95));
    c->tau_normal_base = SFX(clampr)(SFX(tau_from_nu)(SFX(clamp_nu)(nu)), tau_min, K(0.95));
    /* jni.cpp:966-971 benchmark thermal diffusivity (heated cavity: nu / Pr) */
    c->alpha = heated ? nu / pr : K(0.035);
    /* thermal_core.h:268-279 with jni.cpp:161-167 constants; heated cavity:
     * U^2 / (L * dT) (jni.cpp:986-993) */
    if (heated)
        c->beta = (ref_speed * ref_speed) / SFX(rmax)(K(1.0e-6), ref_len * SFX(temperature_span)(g));
    else
        c->beta = ((K(9.81) * (K(1.0) / K(300.0))) * K(20.0)) * (((K(1.0) / K(20.0)) * (K(1.0) / K(20.0))) / K(1.0));
    c->bfecc = heated; /* jni.cpp:871-873 */
    c->cylinder = g->preset == AERO_LBM_BENCHMARK_PRESET_CYLINDER_CROSSFLOW_2D;
    R sgs_c = K(0.025);
    c->nut_cap_ratio = K(5.0);
    c->g_axis = 1;
    c->g_sign = -1;
    if (ov) {
        if (!isnan(ov->sgs_constant)) sgs_c = (R)ov->sgs_constant;
        if (!isnan(ov->sgs_nut_cap_ratio)) c->nut_cap_ratio = (R)ov->sgs_nut_cap_ratio;
        if (!isnan(ov->boussinesq_beta)) c->beta = (R)ov->boussinesq_beta;
        if (!isnan(ov->thermal_diffusivity)) c->alpha = (R)ov->thermal_diffusivity;
        if (ov->gravity_axis >= 0 && ov->gravity_axis <= 2) c->g_axis = ov->gravity_axis;
        if (ov->gravity_sign == 1 || ov->gravity_sign == -1) c->g_sign = ov->gravity_sign;
    }
    c->sgs_c2 = sgs_c * sgs_c;
    c->kmax_speed = K(0.57735026919) * K(0.60);
    return 1;
}

/* ---------------- per-cell helpers ---------------- */
static void SFX(reseed)(SFX(Ctx) * c, size_t cell) {
    /* jni.cpp:1724-1750 */
    const R p = SFX(clampr)(SFX(finite_or)(c->ref_p[cell], K(0.0)), K(-0.03), K(0.03));
    const R rho = SFX(clampr)(K(1.0) + p, K(0.97), K(1.03));
    R ux = SFX(finite_or)(c->ref_ux[cell], K(0.0));
    R uy = SFX(finite_or)(c->ref_uy[cell], K(0.0));
    R uz = SFX(finite_or)(c->ref_uz[cell], K(0.0));
    const R s2 = ux * ux + uy * uy + uz * uz;
    if (!RISFINITE_(s2) || s2 > c->kmax_speed * c->kmax_speed) {
        if (!RISFINITE_(s2) || s2 <= K(0.0)) {
            ux = uy = uz = K(0.0);
        } else {
            const R sc = c->kmax_speed / RSQRT_(s2);
            ux *= sc; uy *= sc; uz *= sc;
        }
    }
    c->rho[cell] = rho;
    c->ux[cell] = ux; c->uy[cell] = uy; c->uz[cell] = uz;
    for (int q = 0; q < 27; ++q) c->f_post[(size_t)q * c->cells + cell] = SFX(feq)(c, q, rho, ux, uy, uz);
}

static int SFX(binom)(int n, int k) {
    if (k < 0 || k > n) return 0;
    if (n <= 1 || k == 0 || k == n) return 1;
    return 2;
}

/* jni.cpp:1335-1357 */
static void SFX(raw_moments)(const R f[27], R raw[3][3][3]) {
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) for (int cc = 0; cc < 3; ++cc) raw[a][b][cc] = K(0.0);
    for (int q = 0; q < 27; ++q) {
        const R fq = f[q];
        const R px[3] = {K(1.0), (R)kCX[q], (R)(kCX[q] * kCX[q])};
        const R py[3] = {K(1.0), (R)kCY[q], (R)(kCY[q] * kCY[q])};
        const R pz[3] = {K(1.0), (R)kCZ[q], (R)(kCZ[q] * kCZ[q])};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int cc = 0; cc < 3; ++cc) raw[a][b][cc] += fq * px[a] * py[b] * pz[cc];
    }
}

/* jni.cpp:1359-1382 */
static void SFX(central_moments)(const R raw[3][3][3], R ux, R uy, R uz, R cen[3][3][3]) {
    const R sx[3] = {K(1.0), -ux, ux * ux};
    const R sy[3] = {K(1.0), -uy, uy * uy};
    const R sz[3] = {K(1.0), -uz, uz * uz};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int cc = 0; cc < 3; ++cc) {
                R sum = K(0.0);
                for (int p = 0; p <= a; ++p)
                    for (int q = 0; q <= b; ++q)
                        for (int r = 0; r <= cc; ++r) {
                            const R coeff = (R)(SFX(binom)(a, p) * SFX(binom)(b, q) * SFX(binom)(cc, r));
                            sum += coeff * sx[a - p] * sy[b - q] * sz[cc - r] * raw[p][q][r];
                        }
                cen[a][b][cc] = sum;
            }
}

/* jni.cpp:1384-1430 */
static void SFX(relax_closure)(R rho, R in[3][3][3], R om_diag, R om_off, R out[3][3][3]) {
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) for (int cc = 0; cc < 3; ++cc) out[a][b][cc] = K(0.0);
    out[0][0][0] = rho;
    const R cs2 = K(1.0) / K(3.0);
    const R eq2 = rho * cs2;
    const R sd = SFX(clampr)(om_diag, K(0.0), K(1.95));
    const R so = SFX(clampr)(om_off, K(0.0), K(1.95));
    out[2][0][0] = in[2][0][0] + sd * (eq2 - in[2][0][0]);
    out[0][2][0] = in[0][2][0] + sd * (eq2 - in[0][2][0]);
    out[0][0][2] = in[0][0][2] + sd * (eq2 - in[0][0][2]);
    out[1][1][0] = (K(1.0) - so) * in[1][1][0];
    out[1][0][1] = (K(1.0) - so) * in[1][0][1];
    out[0][1][1] = (K(1.0) - so) * in[0][1][1];
    const R c200 = out[2][0][0], c020 = out[0][2][0], c002 = out[0][0][2];
    const R c110 = out[1][1][0], c101 = out[1][0][1], c011 = out[0][1][1];
    const R ir = K(1.0) / SFX(rmax)(K(1e-6), rho);
    out[2][2][0] = c200 * c020 * ir + K(2.0) * c110 * c110 * ir;
    out[2][0][2] = c200 * c002 * ir + K(2.0) * c101 * c101 * ir;
    out[0][2][2] = c020 * c002 * ir + K(2.0) * c011 * c011 * ir;
    out[2][1][1] = (c200 * c011 + K(2.0) * c110 * c101) * ir;
    out[1][2][1] = (c020 * c101 + K(2.0) * c110 * c011) * ir;
    out[1][1][2] = (c002 * c110 + K(2.0) * c101 * c011) * ir;
    out[2][2][2] = (c200 * c020 * c002 + K(2.0) * c110 * c110 * c002 + K(2.0) * c101 * c101 * c020 +
                    K(2.0) * c011 * c011 * c200 + K(8.0) * c110 * c101 * c011) * ir * ir;
}

/* jni.cpp:1432-1455 */
static void SFX(central_to_raw)(const R cen[3][3][3], R ux, R uy, R uz, R raw[3][3][3]) {
    const R px[3] = {K(1.0), ux, ux * ux};
    const R py[3] = {K(1.0), uy, uy * uy};
    const R pz[3] = {K(1.0), uz, uz * uz};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int cc = 0; cc < 3; ++cc) {
                R sum = K(0.0);
                for (int p = 0; p <= a; ++p)
                    for (int q = 0; q <= b; ++q)
                        for (int r = 0; r <= cc; ++r) {
                            const R coeff = (R)(SFX(binom)(a, p) * SFX(binom)(b, q) * SFX(binom)(cc, r));
                            sum += coeff * px[a - p] * py[b - q] * pz[cc - r] * cen[p][q][r];
                        }
                raw[a][b][cc] = sum;
            }
}

/* jni.cpp:1457-1473 */
static void SFX(raw_to_pop)(const R raw[3][3][3], R f[27]) {
    for (int ix = 0; ix < 3; ++ix)
        for (int iy = 0; iy < 3; ++iy)
            for (int iz = 0; iz < 3; ++iz) {
                R sum = K(0.0);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int cc = 0; cc < 3; ++cc)
                            sum += SFX(kInv1D)[ix][a] * SFX(kInv1D)[iy][b] * SFX(kInv1D)[iz][cc] * raw[a][b][cc];
                f[(ix * 3 + iy) * 3 + iz] = sum;
            }
}

/* jni.cpp:1270-1281 */
static R SFX(guo)(const SFX(Ctx) * c, int q, R ux, R uy, R uz, R fx, R fy, R fz, R omega) {
    const R cx = (R)kCX[q], cy = (R)kCY[q], cz = (R)kCZ[q];
    const R cu = cx * ux + cy * uy + cz * uz;
    const R cf = cx * fx + cy * fy + cz * fz;
    const R uf = ux * fx + uy * fy + uz * fz;
    const R inv_cs2 = K(1.0) / (K(1.0) / K(3.0));
    const R inv_cs4 = inv_cs2 * inv_cs2;
    const R pref = SFX(rmax)(K(0.0), K(1.0) - K(0.5) * SFX(clampr)(omega, K(0.0), K(1.999)));
    return pref * c->w[q] * ((cf - uf) * inv_cs2 + cu * cf * inv_cs4);
}

/* jni.cpp:767-774 effective_sponge_alpha (benchmark flag gates it) */
static R SFX(sponge_alpha)(const SFX(Ctx) * c, int x, int y, int z) {
    if (SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_SPONGE)) return K(0.0);
    int d = x;
    if (y < d) d = y;
    if (z < d) d = z;
    if (c->nx - 1 - x < d) d = c->nx - 1 - x;
    if (c->ny - 1 - y < d) d = c->ny - 1 - y;
    if (c->nz - 1 - z < d) d = c->nz - 1 - z;
    if (d >= 4) return K(0.0);
    const R eta = (R)(4 - d) / (R)4;
    return SFX(clampr)(K(0.03) * eta * eta, K(0.0), K(0.95));
}

static void SFX(decode)(const SFX(Ctx) * c, size_t cell, int* x, int* y, int* z) {
    const int yz = c->ny * c->nz;
    *x = (int)(cell / (size_t)yz);
    const int rem = (int)(cell - (size_t)*x * (size_t)yz);
    *y = rem / c->nz;
    *z = rem - *y * c->nz;
}

/* hash_signed_noise (jni.cpp:1318-1327): SplitMix64 finalizer of cell and tick,
 * 24 low bits mapped to [-1, 1] */
static R SFX(hash_noise)(uint64_t cell, uint64_t tick) {
    uint64_t x = (cell * 0x9E3779B97F4A7C15ull) ^ (tick * 0xD1B54A32D192ED03ull);
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    const R u = (R)(float)(x & 0x00FFFFFFull) * (K(1.0) / K(16777215.0));
    return K(2.0) * u - K(1.0);
}

/* effective_runtime_state_nudge (jni.cpp:683-693): 0.08 (1 - d/8)^2 within 8 cells
 * of a face in runtime mode, 0 in benchmark mode */
static R SFX(state_nudge)(const SFX(Ctx) * c, int x, int y, int z) {
    if (!c->runtime) return K(0.0);
    int d = x;
    if (y < d) d = y;
    if (z < d) d = z;
    if (c->nx - 1 - x < d) d = c->nx - 1 - x;
    if (c->ny - 1 - y < d) d = c->ny - 1 - y;
    if (c->nz - 1 - z < d) d = c->nz - 1 - z;
    if (d >= 8) return K(0.0);
    const R eta = (R)(8 - d) / (R)8;
    return K(0.08) * eta * eta;
}

/* ---------------- collide (jni.cpp:2209-2413) ---------------- */
static void SFX(collide)(SFX(Ctx) * c) {
    const size_t N = c->cells;
    const int sgs_on = !SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_SGS);
    const int buoy_on = !SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_BUOYANCY);
    /* effective_enable_fan_forcing / effective_fan_noise_amp (jni.cpp:799-809) */
    const int fan_on = !SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING);
    const R noise_amp = SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_NOISE) ? K(0.0) : K(0.02);
    const R cs2 = K(1.0) / K(3.0);
    /* cells are independent: threads change nothing but wall time */
#pragma omp parallel for schedule(static)
    for (size_t cell = 0; cell < N; ++cell) {
        if (c->obstacle[cell]) {
            c->rho[cell] = K(1.0);
            c->ux[cell] = c->uy[cell] = c->uz[cell] = K(0.0);
            for (int q = 0; q < 27; ++q) c->f_post[(size_t)q * N + cell] = c->f[(size_t)q * N + cell];
            continue;
        }
        R fl[27];
        R rho = K(0.0), mx = K(0.0), my = K(0.0), mz = K(0.0);
        int bad = 0;
        for (int q = 0; q < 27; ++q) {
            const R fq = c->f[(size_t)q * N + cell];
            if (!RISFINITE_(fq)) { bad = 1; break; }
            fl[q] = fq;
            rho += fq;
            mx += fq * (R)kCX[q];
            my += fq * (R)kCY[q];
            mz += fq * (R)kCZ[q];
        }
        if (bad || !RISFINITE_(rho) || !RISFINITE_(mx) || !RISFINITE_(my) || !RISFINITE_(mz)) {
            SFX(reseed)(c, cell);
            continue;
        }
        const R rho_safe = SFX(rmax)(K(1e-6), rho);
        const R inv_rho = K(1.0) / rho_safe;
        R ux = mx * inv_rho, uy = my * inv_rho, uz = mz * inv_rho;
        if (!RISFINITE_(rho_safe) || !RISFINITE_(inv_rho) || !RISFINITE_(ux) || !RISFINITE_(uy) || !RISFINITE_(uz)) {
            SFX(reseed)(c, cell);
            continue;
        }
        const R speed_pre = RSQRT_(ux * ux + uy * uy + uz * uz);
        R F[3] = {K(0.0), K(0.0), K(0.0)};
        /* fan forcing (jni.cpp:2258-2286): push toward the fan speed along its axis,
         * damp the perpendicular velocity, soften above a speed cap */
        if (fan_on && c->fan_mask[cell] > K(0.0)) {
            const R fu = c->fan_ux[cell], fv = c->fan_uy[cell], fw = c->fan_uz[cell];
            const R fan_norm = RSQRT_(fu * fu + fv * fv + fw * fw);
            if (fan_norm > K(1e-8)) {
                const R inv_norm = K(1.0) / fan_norm;
                const R noise = K(1.0) + noise_amp * SFX(hash_noise)((uint64_t)cell, c->step_counter);
                const R target = SFX(clampr)(fan_norm * K(0.05) * SFX(rmax)(K(0.0), noise), K(0.0), K(0.28));
                const R nx = fu * inv_norm, ny = fv * inv_norm, nz = fw * inv_norm;
                const R u_para = ux * nx + uy * ny + uz * nz;
                const R upx = ux - u_para * nx, upy = uy - u_para * ny, upz = uz - u_para * nz;
                const R push = SFX(rmax)(K(0.0), target - u_para);
                R damp = K(1.0);
                if (speed_pre > K(0.24)) {
                    const R r = (speed_pre - K(0.24)) / SFX(rmax)(K(1e-4), K(0.04));
                    damp = K(1.0) / (K(1.0) + r * r);
                }
                const R fb = c->fan_mask[cell] * K(0.07) * damp;
                F[0] = fb * rho_safe * (push * nx - K(1.0) * upx);
                F[1] = fb * rho_safe * (push * ny - K(1.0) * upy);
                F[2] = fb * rho_safe * (push * nz - K(1.0) * upz);
            }
        }
        if (buoy_on) {
            const R b = SFX(clampr)(c->beta * c->T[cell], K(-0.02), K(0.02));
            if (c->g_sign < 0) F[c->g_axis] += rho_safe * b;
            else F[c->g_axis] += rho_safe * (-b);
        }
        const R dux = F[0] * inv_rho, duy = F[1] * inv_rho, duz = F[2] * inv_rho;
        ux += K(0.5) * dux;
        uy += K(0.5) * duy;
        uz += K(0.5) * duz;
        {
            /* state nudge toward the reference velocity (jni.cpp:2306-2314); 0 in benchmark mode */
            int x, y, z;
            SFX(decode)(c, cell, &x, &y, &z);
            const R sn = SFX(state_nudge)(c, x, y, z);
            ux = (K(1.0) - sn) * ux + sn * c->ref_ux[cell];
            uy = (K(1.0) - sn) * uy + sn * c->ref_uy[cell];
            uz = (K(1.0) - sn) * uz + sn * c->ref_uz[cell];
        }
        if (!RISFINITE_(ux) || !RISFINITE_(uy) || !RISFINITE_(uz)) { SFX(reseed)(c, cell); continue; }
        const R s2 = ux * ux + uy * uy + uz * uz;
        if (!RISFINITE_(s2)) { SFX(reseed)(c, cell); continue; }
        if (s2 > c->kmax_speed * c->kmax_speed) {
            const R sc = c->kmax_speed / RSQRT_(s2);
            ux *= sc; uy *= sc; uz *= sc;
        }
        c->rho[cell] = rho;
        c->ux[cell] = ux; c->uy[cell] = uy; c->uz[cell] = uz;

        R raw[3][3][3], cen[3][3][3], cpost[3][3][3], rpost[3][3][3];
        SFX(raw_moments)(fl, raw);
        SFX(central_moments)(raw, ux, uy, uz, cen);

        R ts = c->tau_shear_base, tn = c->tau_normal_base;
        if (sgs_on) {
            const R nu0 = SFX(rmax)(K(1e-6), c->nu_base);
            const R nxx = cen[2][0][0] - rho_safe * cs2;
            const R nyy = cen[0][2][0] - rho_safe * cs2;
            const R nzz = cen[0][0][2] - rho_safe * cs2;
            const R nxy = cen[1][1][0], nxz = cen[1][0][1], nyz = cen[0][1][1];
            const R qn2 = nxx * nxx + nyy * nyy + nzz * nzz + K(2.0) * (nxy * nxy + nxz * nxz + nyz * nyz);
            const R qmag = RSQRT_(SFX(rmax)(K(0.0), qn2));
            const R smag = qmag / SFX(rmax)(K(1e-6), K(2.0) * rho_safe * nu0);
            R nut = c->sgs_c2 * smag;
            nut = SFX(rmin)(nut, c->nut_cap_ratio * nu0);
            const R tau_min = K(0.5) + K(3.0) * K(1.0e-8);
            ts = SFX(clampr)(SFX(tau_from_nu)(nu0 + nut), tau_min, K(0.95));
            tn = SFX(clampr)(c->tau_normal_base + (ts - c->tau_shear_base) * K(0.30), tau_min, K(0.95));
        }
        const R om_diag = K(1.0) / tn;
        const R om_off = K(1.0) / ts;
        SFX(relax_closure)(rho_safe, cen, om_diag, om_off, cpost);
        SFX(central_to_raw)(cpost, ux, uy, uz, rpost);
        R fc[27];
        SFX(raw_to_pop)(rpost, fc);
        R rc = K(0.0), mxc = K(0.0), myc = K(0.0), mzc = K(0.0);
        for (int q = 0; q < 27; ++q) {
            const R fq = fc[q];
            rc += fq;
            mxc += fq * (R)kCX[q];
            myc += fq * (R)kCY[q];
            mzc += fq * (R)kCZ[q];
        }
        const R rcs = SFX(rmax)(K(1e-6), rc);
        const R uxc = mxc / rcs, uyc = myc / rcs, uzc = mzc / rcs;
        for (int q = 0; q < 27; ++q)
            fc[q] += SFX(feq)(c, q, rho_safe, ux, uy, uz) - SFX(feq)(c, q, rcs, uxc, uyc, uzc);
        for (int q = 0; q < 27; ++q) {
            const R s = SFX(guo)(c, q, ux, uy, uz, F[0], F[1], F[2], om_off);
            c->f_post[(size_t)q * N + cell] = fc[q] + s;
        }
        if (c->shell_layer && c->shell_layer[cell]) {
            /* the density and velocity of the post-collision state relax toward the
             * reference by alpha = clamp(relax * (1 - (layer-1)/layers)^2, 0, 0.5);
             * the populations move by the difference of the two equilibria */
            const R* tg = c->shell_target + 4 * cell;
            const R st = K(1.0) - (R)(c->shell_layer[cell] - 1) / SFX(rmax)(K(1.0), (R)c->shell_layers);
            const R al = SFX(clampr)(c->shell_relax * st * st, K(0.0), K(0.5));
            R rp = K(0.0);
            for (int q = 0; q < 27; ++q) rp += c->f_post[(size_t)q * N + cell];
            const R trho = SFX(clampr)(K(1.0) + SFX(clampr)(tg[3], K(-0.03), K(0.03)), K(0.97), K(1.03));
            const R nr = rp + al * (trho - rp);
            const R nx = ux + al * (tg[0] - ux), ny = uy + al * (tg[1] - uy), nz = uz + al * (tg[2] - uz);
            for (int q = 0; q < 27; ++q)
                c->f_post[(size_t)q * N + cell] += SFX(feq)(c, q, nr, nx, ny, nz) - SFX(feq)(c, q, rp, ux, uy, uz);
        }
        {
            const int yz = c->ny * c->nz;
            const int x = (int)(cell / (size_t)yz);
            const int rem = (int)(cell - (size_t)x * (size_t)yz);
            const int y = rem / c->nz;
            const int z = rem - y * c->nz;
            const R al = SFX(sponge_alpha)(c, x, y, z);
            if (al > K(0.0)) {
                const R keep = K(1.0) - al;
                for (int q = 0; q < 27; ++q) {
                    const R fe = SFX(feq)(c, q, K(1.0), K(0.0), K(0.0), K(0.0));
                    R* fr = &c->f_post[(size_t)q * N + cell];
                    *fr = keep * *fr + al * fe;
                }
            }
        }
    }
}

/* ---------------- streaming (jni.cpp:1475-1602, 2415-2452) ---------------- */
static R SFX(bounce)(const SFX(Ctx) * c, size_t cell, int q) {
    /* kObstacleBounceBlend = 0 (jni.cpp:140) so the blend never applies */
    return c->f_post[(size_t)kOPP[q] * c->cells + cell];
}

static int SFX(iclamp)(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static R SFX(convective)(const SFX(Ctx) * c, size_t cell, int x, int y, int z, int q, int sx, int sy, int sz) {
    const int dx = (sx < 0 || sx >= c->nx) ? kCX[q] : 0;
    const int dy = (sy < 0 || sy >= c->ny) ? kCY[q] : 0;
    const int dz = (sz < 0 || sz >= c->nz) ? kCZ[q] : 0;
    const size_t s1 = SFX(cell_index)(c, SFX(iclamp)(x + dx, 0, c->nx - 1), SFX(iclamp)(y + dy, 0, c->ny - 1),
                                      SFX(iclamp)(z + dz, 0, c->nz - 1));
    if (c->obstacle[s1]) return SFX(bounce)(c, cell, q);
    const R f1 = c->f_post[(size_t)q * c->cells + s1];
    const size_t s2 = SFX(cell_index)(c, SFX(iclamp)(x + 2 * dx, 0, c->nx - 1), SFX(iclamp)(y + 2 * dy, 0, c->ny - 1),
                                      SFX(iclamp)(z + 2 * dz, 0, c->nz - 1));
    if (c->obstacle[s2]) return f1;
    const R f2 = c->f_post[(size_t)q * c->cells + s2];
    return f1 + K(0.15) * (f1 - f2);
}

/* parabolic inlet of the cylinder preset (jni.cpp:214-220): 4 u_max s (1 - s),
 * s = y / (ny - 1) */
static R SFX(cylinder_inlet_ux)(const SFX(Ctx) * c, int y, R u_max) {
    const R h = (R)(c->ny - 1 > 1 ? c->ny - 1 : 1);
    const R denom = SFX(rmax)(h, K(1.0e-6));
    const R s = ((R)y - K(0.0)) / denom;
    if (s < K(0.0) || s > K(1.0)) return K(0.0);
    return K(4.0) * u_max * s * (K(1.0) - s);
}

static R SFX(equilibrium_face)(const SFX(Ctx) * c, size_t cell, int y, int q, int sx,
                               const AeroLbmBoundaryFaceConfig* fc) {
    R rho = K(1.0) + SFX(clampr)(c->ref_p[cell], K(-0.03), K(0.03));
    R ux = K(0.0), uy = K(0.0), uz = K(0.0);
    switch (fc->hydrodynamic_kind) {
        case AERO_LBM_HYDRO_BOUNDARY_MOVING_WALL:
        case AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET:
            ux = (R)fc->velocity[0]; uy = (R)fc->velocity[1]; uz = (R)fc->velocity[2];
            if (c->cylinder && sx < 0) { /* jni.cpp:1540-1544 */
                ux = SFX(cylinder_inlet_ux)(c, y, (R)fc->velocity[0]);
                uy = K(0.0);
                uz = K(0.0);
            }
            break;
        case AERO_LBM_HYDRO_BOUNDARY_PRESSURE_DIRICHLET:
            rho = K(1.0) + SFX(clampr)((R)fc->pressure, K(-0.03), K(0.03));
            break;
        default: break;
    }
    rho = SFX(clampr)(rho, K(0.97), K(1.03));
    const R s2 = ux * ux + uy * uy + uz * uz;
    if (s2 > c->kmax_speed * c->kmax_speed && s2 > K(0.0)) {
        const R sc = c->kmax_speed / RSQRT_(s2);
        ux *= sc; uy *= sc; uz *= sc;
    }
    return SFX(feq)(c, q, rho, ux, uy, uz);
}

static R SFX(face_value)(const SFX(Ctx) * c, size_t cell, int x, int y, int z, int q, int sx, int sy, int sz) {
    const AeroLbmBoundaryFaceConfig* fc = SFX(face_for_oob)(c, sx, sy, sz);
    const int kind = fc ? fc->hydrodynamic_kind : AERO_LBM_HYDRO_BOUNDARY_INHERIT_GAME;
    switch (kind) {
        case AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK:
            return SFX(bounce)(c, cell, q);
        case AERO_LBM_HYDRO_BOUNDARY_MOVING_WALL:
        case AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET:
        case AERO_LBM_HYDRO_BOUNDARY_PRESSURE_DIRICHLET:
            return SFX(equilibrium_face)(c, cell, y, q, sx, fc);
        case AERO_LBM_HYDRO_BOUNDARY_SYMMETRY: {
            const size_t s = SFX(cell_index)(c, SFX(iclamp)(sx, 0, c->nx - 1), SFX(iclamp)(sy, 0, c->ny - 1),
                                             SFX(iclamp)(sz, 0, c->nz - 1));
            return c->obstacle[s] ? SFX(bounce)(c, cell, q) : c->f_post[(size_t)q * c->cells + s];
        }
        case AERO_LBM_HYDRO_BOUNDARY_CONVECTIVE_OUTFLOW:
            return SFX(convective)(c, cell, x, y, z, q, sx, sy, sz);
        case AERO_LBM_HYDRO_BOUNDARY_PERIODIC:
            break;
        default:
            if (SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_CONVECTIVE_OUTFLOW)) return SFX(bounce)(c, cell, q);
            return SFX(convective)(c, cell, x, y, z, q, sx, sy, sz);
    }
    return SFX(convective)(c, cell, x, y, z, q, sx, sy, sz);
}

static void SFX(stream)(SFX(Ctx) * c) {
    const size_t N = c->cells;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < c->nx; ++x)
        for (int y = 0; y < c->ny; ++y)
            for (int z = 0; z < c->nz; ++z) {
                const size_t cell = SFX(cell_index)(c, x, y, z);
                R fl[27];
                if (c->obstacle[cell]) {
                    for (int q = 0; q < 27; ++q) fl[q] = c->f_post[(size_t)kOPP[q] * N + cell];
                } else {
                    for (int q = 0; q < 27; ++q) {
                        int sx = x - kCX[q], sy = y - kCY[q], sz = z - kCZ[q];
                        SFX(remap)(c, &sx, &sy, &sz, 0);
                        if (sx < 0 || sy < 0 || sz < 0 || sx >= c->nx || sy >= c->ny || sz >= c->nz) {
                            fl[q] = SFX(face_value)(c, cell, x, y, z, q, sx, sy, sz);
                            continue;
                        }
                        const size_t src = SFX(cell_index)(c, sx, sy, sz);
                        fl[q] = c->obstacle[src] ? SFX(bounce)(c, cell, q) : c->f_post[(size_t)q * N + src];
                    }
                }
                for (int q = 0; q < 27; ++q) c->f[(size_t)q * N + cell] = fl[q];
            }
}

/* ---------------- thermal scalar, benchmark first-order path (jni.cpp:2145-2188) ---------------- */
static R SFX(t_neighbor)(const SFX(Ctx) * c, const R* field, int x, int y, int z, R self) {
    SFX(remap)(c, &x, &y, &z, 1);
    if (x < 0 || y < 0 || z < 0 || x >= c->nx || y >= c->ny || z >= c->nz) {
        const AeroLbmBoundaryFaceConfig* fc = SFX(face_for_oob)(c, x, y, z);
        const int kind = fc ? fc->thermal_kind : AERO_LBM_THERMAL_BOUNDARY_INHERIT_GAME;
        switch (kind) {
            case AERO_LBM_THERMAL_BOUNDARY_TEMPERATURE_DIRICHLET:
                return SFX(clampr)((R)fc->temperature, K(-1.0), K(1.0));
            case AERO_LBM_THERMAL_BOUNDARY_HEAT_FLUX_NEUMANN:
                return SFX(clampr)(self + (R)fc->heat_flux, K(-1.0), K(1.0));
            default:
                return self;
        }
    }
    const size_t cell = SFX(cell_index)(c, x, y, z);
    if (c->obstacle[cell]) return self;
    return field[cell];
}

static R SFX(trilinear)(const SFX(Ctx) * c, const R* field, R x, R y, R z, R fb) {
    const int x0 = (int)RFLOOR_(x), y0 = (int)RFLOOR_(y), z0 = (int)RFLOOR_(z);
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const R fx = x - (R)x0, fy = y - (R)y0, fz = z - (R)z0;
    const R c000 = SFX(t_neighbor)(c, field, x0, y0, z0, fb);
    const R c100 = SFX(t_neighbor)(c, field, x1, y0, z0, fb);
    const R c010 = SFX(t_neighbor)(c, field, x0, y1, z0, fb);
    const R c110 = SFX(t_neighbor)(c, field, x1, y1, z0, fb);
    const R c001 = SFX(t_neighbor)(c, field, x0, y0, z1, fb);
    const R c101 = SFX(t_neighbor)(c, field, x1, y0, z1, fb);
    const R c011 = SFX(t_neighbor)(c, field, x0, y1, z1, fb);
    const R c111 = SFX(t_neighbor)(c, field, x1, y1, z1, fb);
    const R c00 = c000 + (c100 - c000) * fx;
    const R c10 = c010 + (c110 - c010) * fx;
    const R c01 = c001 + (c101 - c001) * fx;
    const R c11 = c011 + (c111 - c011) * fx;
    const R c0 = c00 + (c10 - c00) * fy;
    const R c1 = c01 + (c11 - c01) * fy;
    return c0 + (c1 - c0) * fz;
}


/* local_temperature_bounds (jni.cpp:1883-1907): min / max of the cell and its six
 * face neighbours (solid or outside points read as the cell itself, or the face value) */
static void SFX(local_bounds)(const SFX(Ctx) * c, const R* field, int x, int y, int z, R* lo, R* hi) {
    const R center = field[SFX(cell_index)(c, x, y, z)];
    const R nb[6] = {SFX(t_neighbor)(c, field, x + 1, y, z, center), SFX(t_neighbor)(c, field, x - 1, y, z, center),
                     SFX(t_neighbor)(c, field, x, y + 1, z, center), SFX(t_neighbor)(c, field, x, y - 1, z, center),
                     SFX(t_neighbor)(c, field, x, y, z + 1, center), SFX(t_neighbor)(c, field, x, y, z - 1, center)};
    R mn = center, mx = center;
    for (int i = 0; i < 6; ++i) {
        mn = SFX(rmin)(mn, nb[i]);
        mx = SFX(rmax)(mx, nb[i]);
    }
    *lo = mn;
    *hi = mx;
}

/* velocity_component_or_self (jni.cpp:1952-1969): a neighbour's collision velocity,
 * or the cell's own outside the domain (no periodic wrap) or in a solid */
static R SFX(vel_or_self)(const SFX(Ctx) * c, const R* u, int x, int y, int z, R self) {
    if (x < 0 || y < 0 || z < 0 || x >= c->nx || y >= c->ny || z >= c->nz) return self;
    const size_t n = SFX(cell_index)(c, x, y, z);
    if (c->obstacle[n]) return self;
    return u[n];
}

/* obstacle_adjacent (jni.cpp:1297-1316): a face neighbour inside the domain is solid */
static int SFX(obstacle_adjacent)(const SFX(Ctx) * c, int x, int y, int z) {
    static const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int k = 0; k < 6; ++k) {
        const int a = x + off[k][0], b = y + off[k][1], d = z + off[k][2];
        if (a < 0 || b < 0 || d < 0 || a >= c->nx || b >= c->ny || d >= c->nz) continue;
        if (c->obstacle[SFX(cell_index)(c, a, b, d)]) return 1;
    }
    return 0;
}

/* thermal diffusivity of a cell (jni.cpp:1971-2048): benchmark mode the resolved
 * constant; runtime mode alpha0 + nu_t / Pr_t with the Smagorinsky viscosity of the
 * central-difference strain of the collision velocities (0 next to a solid) */
static R SFX(local_alpha)(const SFX(Ctx) * c, int x, int y, int z, size_t cell) {
    if (!c->runtime) return c->alpha;
    R nu_t = K(0.0);
    if (!SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_SGS) && !SFX(obstacle_adjacent)(c, x, y, z)) {
        const R* U[3] = {c->ux, c->uy, c->uz};
        R g[3][3]; /* g[i][j] = d u_i / d x_j */
        for (int i = 0; i < 3; ++i) {
            const R uc = U[i][cell];
            g[i][0] = K(0.5) * (SFX(vel_or_self)(c, U[i], x + 1, y, z, uc) - SFX(vel_or_self)(c, U[i], x - 1, y, z, uc));
            g[i][1] = K(0.5) * (SFX(vel_or_self)(c, U[i], x, y + 1, z, uc) - SFX(vel_or_self)(c, U[i], x, y - 1, z, uc));
            g[i][2] = K(0.5) * (SFX(vel_or_self)(c, U[i], x, y, z + 1, uc) - SFX(vel_or_self)(c, U[i], x, y, z - 1, uc));
        }
        const R sxx = g[0][0], syy = g[1][1], szz = g[2][2];
        const R sxy = K(0.5) * (g[0][1] + g[1][0]);
        const R sxz = K(0.5) * (g[0][2] + g[2][0]);
        const R syz = K(0.5) * (g[1][2] + g[2][1]);
        const R smag = RSQRT_(SFX(rmax)(K(0.0), K(2.0) * (sxx * sxx + syy * syy + szz * szz) +
                                                   K(4.0) * (sxy * sxy + sxz * sxz + syz * syz)));
        nu_t = c->sgs_c2 * smag;
        nu_t = SFX(rmin)(nu_t, c->nut_cap_ratio * c->nu_rt0);
        nu_t = SFX(rmax)(K(0.0), nu_t);
    }
    return c->alpha + nu_t / SFX(rmax)(K(1.0e-6), K(0.85));
}

/* update_temperature_field, BFECC branch (jni.cpp:2050-2139): advect forward
 * (T -> S), advect S backward (-> B), correct N = T + (T - B) / 2, advect N forward,
 * clamp to the bounds of T around the cell, then diffusion and source of T; the
 * step dt is the update stride (1 heated cavity, 2 runtime mode, jni.cpp:901-909) */
static void SFX(update_temperature_bfecc)(SFX(Ctx) * c) {
    const R dt = (R)c->thermal_stride;
    const int src_on = !SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE);
    const R* T = c->T;
    R* S = c->T_scr;
    R* Nf = c->T_next;
#pragma omp parallel for schedule(static)
    for (size_t cell = 0; cell < c->cells; ++cell) {
        if (c->obstacle[cell]) { S[cell] = K(0.0); continue; }
        int x, y, z;
        SFX(decode)(c, cell, &x, &y, &z);
        S[cell] = SFX(trilinear)(c, T, (R)x - dt * c->ux[cell], (R)y - dt * c->uy[cell], (R)z - dt * c->uz[cell],
                                 T[cell]);
    }
#pragma omp parallel for schedule(static)
    for (size_t cell = 0; cell < c->cells; ++cell) {
        if (c->obstacle[cell]) { Nf[cell] = K(0.0); continue; }
        int x, y, z;
        SFX(decode)(c, cell, &x, &y, &z);
        const R tc = T[cell];
        const R back = SFX(trilinear)(c, S, (R)x + dt * c->ux[cell], (R)y + dt * c->uy[cell], (R)z + dt * c->uz[cell],
                                      S[cell]);
        Nf[cell] = SFX(clampr)(tc + K(0.5) * (tc - back), K(-1.0), K(1.0));
    }
#pragma omp parallel for schedule(static)
    for (size_t cell = 0; cell < c->cells; ++cell) {
        if (c->obstacle[cell]) { S[cell] = K(0.0); continue; }
        int x, y, z;
        SFX(decode)(c, cell, &x, &y, &z);
        const R tc = T[cell];
        R adv = SFX(trilinear)(c, Nf, (R)x - dt * c->ux[cell], (R)y - dt * c->uy[cell], (R)z - dt * c->uz[cell],
                               Nf[cell]);
        R lo, hi;
        SFX(local_bounds)(c, T, x, y, z, &lo, &hi);
        adv = SFX(clampr)(adv, lo, hi);
        const R tsum = SFX(t_neighbor)(c, T, x + 1, y, z, tc) + SFX(t_neighbor)(c, T, x - 1, y, z, tc) +
                       SFX(t_neighbor)(c, T, x, y + 1, z, tc) + SFX(t_neighbor)(c, T, x, y - 1, z, tc) +
                       SFX(t_neighbor)(c, T, x, y, z + 1, tc) + SFX(t_neighbor)(c, T, x, y, z - 1, tc);
        const R lap = tsum - K(6.0) * tc;
        R src = K(0.0);
        if (src_on && c->in_ch > 9) src = SFX(clampr)(c->src[cell], K(-0.006), K(0.006));
        const R alpha = SFX(local_alpha)(c, x, y, z, cell);
        S[cell] = SFX(clampr)(adv + dt * (alpha * lap + src - K(0.0) * adv), K(-1.0), K(1.0));
    }
    c->T_scr = c->T;
    c->T = S;
    memcpy(c->T_next, c->T, c->cells * sizeof(R));
}

static void SFX(update_temperature)(SFX(Ctx) * c) {
    if (c->bfecc) {
        SFX(update_temperature_bfecc)(c);
        return;
    }
    const R dt = (R)1; /* benchmark stride 1 (jni.cpp:901-909) */
    const int src_on = !SFX(flag)(c, AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE);
#pragma omp parallel for schedule(static)
    for (size_t cell = 0; cell < c->cells; ++cell) {
        if (c->obstacle[cell]) { c->T_next[cell] = K(0.0); continue; }
        const int yz = c->ny * c->nz;
        const int x = (int)(cell / (size_t)yz);
        const int rem = (int)(cell - (size_t)x * (size_t)yz);
        const int y = rem / c->nz;
        const int z = rem - y * c->nz;
        const R tc = c->T[cell];
        const R adv = SFX(trilinear)(c, c->T, (R)x - dt * c->ux[cell], (R)y - dt * c->uy[cell], (R)z - dt * c->uz[cell], tc);
        const R tsum = SFX(t_neighbor)(c, c->T, x + 1, y, z, tc) + SFX(t_neighbor)(c, c->T, x - 1, y, z, tc) +
                       SFX(t_neighbor)(c, c->T, x, y + 1, z, tc) + SFX(t_neighbor)(c, c->T, x, y - 1, z, tc) +
                       SFX(t_neighbor)(c, c->T, x, y, z + 1, tc) + SFX(t_neighbor)(c, c->T, x, y, z - 1, tc);
        const R lap = tsum - K(6.0) * tc;
        R src = K(0.0);
        if (src_on && c->in_ch > 9) src = SFX(clampr)(c->src[cell], K(-0.006), K(0.006));
        const R tn = adv + dt * (c->alpha * lap + src - K(0.0) * adv);
        c->T_next[cell] = SFX(clampr)(tn, K(-1.0), K(1.0));
    }
    R* tmp = c->T;
    c->T = c->T_next;
    c->T_next = tmp;
}

/* apply_runtime_temperature_reference (jni.cpp:8037-8064), run before the update on
 * every step after the first: solid cells to 0, cells within 8 of a face nudged
 * toward the payload temperature (runtime mode; the nudge is 0 in benchmark mode) */
static void SFX(temperature_reference)(SFX(Ctx) * c) {
    if (c->in_ch <= 10) return;
    for (size_t i = 0; i < c->cells; ++i) {
        if (c->obstacle[i]) {
            c->T[i] = K(0.0);
            c->T_next[i] = K(0.0);
            continue;
        }
        int x, y, z;
        SFX(decode)(c, i, &x, &y, &z);
        const R tn = SFX(state_nudge)(c, x, y, z);
        if (tn <= K(0.0)) continue;
        const R rt = SFX(clampr)(c->ref_T[i], K(-1.0), K(1.0));
        const R keep = K(1.0) - tn;
        c->T[i] = SFX(clampr)(c->T[i] * keep + rt * tn, K(-1.0), K(1.0));
        c->T_next[i] = c->T[i];
    }
}

/* ---------------- ingest / init (jni.cpp:1671-1694, 1752-1781) ---------------- */
static void SFX(ingest)(SFX(Ctx) * c, const float* payload) {
    const int ch = c->in_ch;
    for (size_t cell = 0; cell < c->cells; ++cell) {
        const float* p = payload + cell * (size_t)ch;
        c->obstacle[cell] = p[0] > 0.5f ? 1 : 0;
        c->fan_mask[cell] = SFX(clampr)((R)p[1], K(0.0), K(1.0));
        c->fan_ux[cell] = SFX(finite_or)((R)p[2], K(0.0));
        c->fan_uy[cell] = SFX(finite_or)((R)p[3], K(0.0));
        c->fan_uz[cell] = SFX(finite_or)((R)p[4], K(0.0));
        c->ref_ux[cell] = SFX(finite_or)((R)p[5], K(0.0));
        c->ref_uy[cell] = SFX(finite_or)((R)p[6], K(0.0));
        c->ref_uz[cell] = SFX(finite_or)((R)p[7], K(0.0));
        c->ref_p[cell] = SFX(clampr)((R)p[8], K(-0.03), K(0.03));
        const R rt = ch > 10 ? (R)p[10] : K(0.0);
        c->ref_T[cell] = SFX(clampr)(rt, K(-1.0), K(1.0));
        const R ts = ch > 9 ? (R)p[9] : K(0.0);
        c->src[cell] = SFX(clampr)(ts, K(-0.006), K(0.006));
    }
}

static void SFX(initialize)(SFX(Ctx) * c) {
    for (size_t cell = 0; cell < c->cells; ++cell) {
        if (c->obstacle[cell]) {
            c->rho[cell] = K(1.0);
            c->ux[cell] = c->uy[cell] = c->uz[cell] = K(0.0);
        } else {
            c->rho[cell] = SFX(clampr)(K(1.0) + c->ref_p[cell], K(0.97), K(1.03));
            c->ux[cell] = c->ref_ux[cell];
            c->uy[cell] = c->ref_uy[cell];
            c->uz[cell] = c->ref_uz[cell];
        }
        for (int q = 0; q < 27; ++q) {
            const R e = SFX(feq)(c, q, c->rho[cell], c->ux[cell], c->uy[cell], c->uz[cell]);
            c->f[(size_t)q * c->cells + cell] = e;
            c->f_post[(size_t)q * c->cells + cell] = e;
        }
        if (!c->runtime && RFABS_(c->T[cell]) < K(1e-8)) {
            c->T[cell] = SFX(clampr)((R)c->cfg.initial_temperature, K(-1.0), K(1.0));
        } else if (!c->obstacle[cell]) {
            c->T[cell] = SFX(clampr)(c->ref_T[cell], K(-1.0), K(1.0));
        }
        c->T[cell] = SFX(clampr)(c->T[cell], K(-1.0), K(1.0));
        c->T_next[cell] = c->T[cell];
    }
    c->initialized = 1;
}

/* ---------------- outputs (jni.cpp:2535-2581, 9378-9448) ---------------- */
static void SFX(cell_output)(const SFX(Ctx) * c, size_t cell, R scale, R out[4]) {
    out[0] = out[1] = out[2] = out[3] = K(0.0);
    if (c->obstacle[cell]) return;
    R rho = K(0.0), ux = K(0.0), uy = K(0.0), uz = K(0.0);
    for (int q = 0; q < 27; ++q) {
        const R fq = c->f[(size_t)q * c->cells + cell];
        if (!RISFINITE_(fq)) return;
        rho += fq;
        ux += fq * (R)kCX[q];
        uy += fq * (R)kCY[q];
        uz += fq * (R)kCZ[q];
    }
    if (!RISFINITE_(rho) || !RISFINITE_(ux) || !RISFINITE_(uy) || !RISFINITE_(uz)) return;
    rho = SFX(clampr)(rho, K(0.97), K(1.03));
    if (!RISFINITE_(rho) || rho <= K(1e-6)) return;
    ux /= rho; uy /= rho; uz /= rho;
    if (!RISFINITE_(ux) || !RISFINITE_(uy) || !RISFINITE_(uz)) ux = uy = uz = K(0.0);
    if (RFABS_(rho - K(1.0)) < K(1e-6)) rho = K(1.0);
    if (RFABS_(ux) < K(1e-7)) ux = K(0.0);
    if (RFABS_(uy) < K(1e-7)) uy = K(0.0);
    if (RFABS_(uz) < K(1e-7)) uz = K(0.0);
    out[0] = ux * scale;
    out[1] = uy * scale;
    out[2] = uz * scale;
    out[3] = SFX(clampr)(rho - K(1.0), K(-0.03), K(0.03));
}

/* ================= public oracle API ================= */
ORACLE_API void* SFX(oracle_create)(int nx, int ny, int nz, int in_ch, int out_ch,
                                    const AeroLbmBenchmarkConfig* cfg, const AeroLbmSolverOverrides* ov) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || in_ch < 9 || out_ch < 4 || !cfg) return NULL;
    SFX(Ctx)* c = (SFX(Ctx)*)calloc(1, sizeof(SFX(Ctx)));
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->cells = (size_t)nx * (size_t)ny * (size_t)nz;
    c->in_ch = in_ch; c->out_ch = out_ch;
    c->cfg = *cfg;
    oracle_sanitize_config(&c->cfg);
    SFX(init_weights)(c->w);
    if (!SFX(resolve)(c, ov)) { free(c); return NULL; }
    const size_t N = c->cells;
    c->f = (R*)calloc(27 * N, sizeof(R));
    c->f_post = (R*)calloc(27 * N, sizeof(R));
    c->rho = (R*)calloc(N, sizeof(R));
    c->ux = (R*)calloc(N, sizeof(R)); c->uy = (R*)calloc(N, sizeof(R)); c->uz = (R*)calloc(N, sizeof(R));
    c->ref_ux = (R*)calloc(N, sizeof(R)); c->ref_uy = (R*)calloc(N, sizeof(R)); c->ref_uz = (R*)calloc(N, sizeof(R));
    c->ref_p = (R*)calloc(N, sizeof(R)); c->ref_T = (R*)calloc(N, sizeof(R)); c->src = (R*)calloc(N, sizeof(R));
    c->obstacle = (unsigned char*)calloc(N, 1);
    c->fan_mask = (R*)calloc(N, sizeof(R)); c->fan_ux = (R*)calloc(N, sizeof(R));
    c->fan_uy = (R*)calloc(N, sizeof(R)); c->fan_uz = (R*)calloc(N, sizeof(R));
    c->T = (R*)calloc(N, sizeof(R)); c->T_next = (R*)calloc(N, sizeof(R)); c->T_scr = (R*)calloc(N, sizeof(R));
    for (size_t i = 0; i < N; ++i) c->rho[i] = K(1.0);
    return c;
}

ORACLE_API void SFX(oracle_destroy)(void* h) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    if (!c) return;
    free(c->f); free(c->f_post); free(c->rho); free(c->ux); free(c->uy); free(c->uz);
    free(c->ref_ux); free(c->ref_uy); free(c->ref_uz); free(c->ref_p); free(c->ref_T); free(c->src);
    free(c->fan_mask); free(c->fan_ux); free(c->fan_uy); free(c->fan_uz);
    free(c->obstacle); free(c->T); free(c->T_next); free(c->T_scr); free(c->shell_layer); free(c->shell_target); free(c);
}

/* jni.cpp:9582-9601 assign_temperature_state */
ORACLE_API void SFX(oracle_set_temperature)(void* h, const float* t) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    for (size_t i = 0; i < c->cells; ++i) {
        c->T[i] = SFX(clampr)((R)t[i], K(-1.0), K(1.0));
        c->T_next[i] = c->T[i];
    }
}

/* jni.cpp:6156-6415 and 4725-4778 (patch_d3q27_f16_static_cells, the patch and reset
 * kernels): the entries set the obstacle flags in order; every cell an entry changed
 * has its populations (the streamed f the next step collides) reset after the last
 * entry, to zero when it ends solid and to the rest equilibrium feq(1, 0) when it
 * ends fluid. Fan direction codes drive a force of the reference's FP16 step only. */
ORACLE_API void SFX(oracle_patch_cells)(void* h, int count, const int* cells, const unsigned char* solid) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    const size_t N = c->cells;
    unsigned char* reset = (unsigned char*)calloc(N, 1);
    for (int i = 0; i < count; ++i) {
        const int cell = cells[i];
        if (cell < 0 || (size_t)cell >= N) continue;
        const unsigned char s = solid[i] ? 1 : 0;
        if (c->obstacle[cell] != s) reset[cell] = 1;
        c->obstacle[cell] = s;
    }
    for (size_t cell = 0; cell < N; ++cell) {
        if (!reset[cell]) continue;
        for (int q = 0; q < 27; ++q)
            c->f[(size_t)q * N + cell] = c->obstacle[cell] ? K(0.0) : SFX(feq)(c, q, K(1.0), K(0.0), K(0.0), K(0.0));
    }
    free(reset);
}

/* jni.cpp:8067-8088 shift_scalar_field, 8090-8131 shift_population_field: new(x) =
 * old(x + d), fill outside */
static void SFX(shift_plane)(R* p, int nx, int ny, int nz, int dx, int dy, int dz, R fill) {
    const size_t N = (size_t)nx * ny * nz;
    R* tmp = (R*)malloc(N * sizeof(R));
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int ox = x + dx, oy = y + dy, oz = z + dz;
                const int in = ox >= 0 && oy >= 0 && oz >= 0 && ox < nx && oy < ny && oz < nz;
                tmp[((size_t)x * ny + y) * nz + z] = in ? p[((size_t)ox * ny + oy) * nz + oz] : fill;
            }
    memcpy(p, tmp, N * sizeof(R));
    free(tmp);
}

/* jni.cpp:8432-8476 shift_context_cpu_state */
ORACLE_API int SFX(oracle_shift)(void* h, int dx, int dy, int dz) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t N = c->cells;
    if (dx == 0 && dy == 0 && dz == 0) return 1;
    if (abs(dx) >= nx || abs(dy) >= ny || abs(dz) >= nz) return 0;
    for (int q = 0; q < 27; ++q) {
        const R eq = SFX(feq)(c, q, K(1.0), K(0.0), K(0.0), K(0.0));
        SFX(shift_plane)(c->f + (size_t)q * N, nx, ny, nz, dx, dy, dz, eq);
        SFX(shift_plane)(c->f_post + (size_t)q * N, nx, ny, nz, dx, dy, dz, eq);
    }
    SFX(shift_plane)(c->rho, nx, ny, nz, dx, dy, dz, K(1.0));
    R* zero_fields[] = {c->ux, c->uy, c->uz, c->ref_ux, c->ref_uy, c->ref_uz, c->ref_p, c->ref_T,
                        c->fan_mask, c->fan_ux, c->fan_uy, c->fan_uz, c->src, c->T, c->T_next, c->T_scr};
    for (size_t k = 0; k < sizeof(zero_fields) / sizeof(zero_fields[0]); ++k)
        SFX(shift_plane)(zero_fields[k], nx, ny, nz, dx, dy, dz, K(0.0));
    unsigned char* ob = (unsigned char*)malloc(N);
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int ox = x + dx, oy = y + dy, oz = z + dz;
                const int in = ox >= 0 && oy >= 0 && oz >= 0 && ox < nx && oy < ny && oz < nz;
                ob[((size_t)x * ny + y) * nz + z] = in ? c->obstacle[((size_t)ox * ny + oy) * nz + oz] : 0;
            }
    memcpy(c->obstacle, ob, N);
    free(ob);
    c->force[0] = c->force[1] = c->force[2] = 0.0f;
    return 1;
}

/* jni.cpp:2198-2203 temperature_was_updated_last_step */
static int SFX(t_updated_last)(const SFX(Ctx) * c) {
    const uint64_t stride = (uint64_t)(c->thermal_stride > 1 ? c->thermal_stride : 1);
    if (stride <= 1 || c->step_counter == 0) return 1;
    return ((c->step_counter - 1) % stride) == 0;
}

/* one axis of the halo boxes, jni.cpp:8201-8263 halo_exchange_slab_bounds */
static int SFX(halo_axis)(int off, int n, int halo, int core, int* a_src, int* b_dst, int* b_src, int* a_dst,
                          int* size) {
    if (off == -1) { *a_src = halo; *b_dst = n - halo; *b_src = core; *a_dst = 0; *size = halo; return 1; }
    if (off == 0) { *a_src = halo; *b_dst = halo; *b_src = halo; *a_dst = halo; *size = core; return 1; }
    if (off == 1) { *a_src = core; *b_dst = 0; *b_src = halo; *a_dst = n - halo; *size = halo; return 1; }
    return 0;
}

static void SFX(copy_box)(const R* src, R* dst, int n, int planes, size_t stride, const int s0[3], const int t0[3],
                          const int size[3]) {
    for (int p = 0; p < planes; ++p)
        for (int i = 0; i < size[0]; ++i)
            for (int j = 0; j < size[1]; ++j)
                for (int k = 0; k < size[2]; ++k) {
                    const size_t from = ((size_t)(s0[0] + i) * n + (s0[1] + j)) * n + (s0[2] + k);
                    const size_t to = ((size_t)(t0[0] + i) * n + (t0[1] + j)) * n + (t0[2] + k);
                    dst[(size_t)p * stride + to] = src[(size_t)p * stride + from];
                }
}

/* jni.cpp:8266-8426 exchange_context_halo_cpu_layers: populations f of the boundary
 * boxes both ways, the temperature too when either context updated it last step
 * (then the next-temperature field takes the temperature) */
ORACLE_API int SFX(oracle_exchange_halo)(void* ha, void* hb, int ox, int oy, int oz, int halo_layers) {
    SFX(Ctx)* a = (SFX(Ctx)*)ha;
    SFX(Ctx)* b = (SFX(Ctx)*)hb;
    const int n = a->nx;
    if (a->nx != b->nx || a->ny != b->ny || a->nz != b->nz || a->nx != a->ny || a->ny != a->nz) return 0;
    int halo = halo_layers < n / 2 - 1 ? halo_layers : n / 2 - 1;
    if (halo < 1) halo = 1;
    const int core = n - 2 * halo;
    if (core <= 0) return 0;
    int as[3], bd[3], bs[3], ad[3], size[3];
    const int off[3] = {ox, oy, oz};
    for (int k = 0; k < 3; ++k)
        if (!SFX(halo_axis)(off[k], n, halo, core, &as[k], &bd[k], &bs[k], &ad[k], &size[k])) return 0;
    const size_t N = a->cells;
    SFX(copy_box)(a->f, b->f, n, 27, N, as, bd, size);
    SFX(copy_box)(b->f, a->f, n, 27, N, bs, ad, size);
    if (SFX(t_updated_last)(a) || SFX(t_updated_last)(b)) {
        SFX(copy_box)(a->T, b->T, n, 1, N, as, bd, size);
        SFX(copy_box)(b->T, a->T, n, 1, N, bs, ad, size);
        memcpy(a->T_next, a->T, N * sizeof(R));
        memcpy(b->T_next, b->T, N * sizeof(R));
    }
    return 1;
}

ORACLE_API void SFX(oracle_get_temperature)(void* h, float* t) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    for (size_t i = 0; i < c->cells; ++i) t[i] = (float)c->T[i];
}

/* cylinder of the crossflow preset (jni.cpp:190-228): geometry in binary32 as the
 * reference defines it, diameter 0.1/2.2 of the length (at least 2.5 cells, at
 * most the height - 4), centre at (0.2/2.2 of the length, 0.2/0.41 of the height) */
static int SFX(cylinder_solid)(const SFX(Ctx) * c, int x, int y) {
    const float len = (float)(c->nx - 1 > 1 ? c->nx - 1 : 1);
    const float h = (float)(c->ny - 1 > 1 ? c->ny - 1 : 1);
    float d = 0.1f / 2.2f * len;
    d = d < 2.5f ? 2.5f : d;
    const float dmax = (h - 4.0f) < 2.5f ? 2.5f : (h - 4.0f);
    d = dmax < d ? dmax : d;
    const float r = 0.5f * d;
    const float cx = 0.2f / 2.2f * len;
    const float cy = 0.0f + 0.2f / 0.41f * h;
    const float dx = (float)x - cx, dy = (float)y - cy;
    return dx * dx + dy * dy <= r * r;
}

/* compute_benchmark_force (jni.cpp:2452-2492): momentum exchange 2 f*_q c_q over the
 * in-plane links (c_z = 0) from fluid cells into solid cylinder cells, summed in
 * double in cell and q order, divided by the depth */
static void SFX(benchmark_force)(SFX(Ctx) * c) {
    c->force[0] = c->force[1] = c->force[2] = 0.0f;
    if (!c->cylinder) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int x = 0; x < c->nx; ++x)
        for (int y = 0; y < c->ny; ++y)
            for (int z = 0; z < c->nz; ++z) {
                const size_t cell = SFX(cell_index)(c, x, y, z);
                if (c->obstacle[cell]) continue;
                for (int q = 1; q < 27; ++q) {
                    if (kCZ[q] != 0) continue;
                    const int ax = x + kCX[q], ay = y + kCY[q], az = z + kCZ[q];
                    if (ax < 0 || ay < 0 || az < 0 || ax >= c->nx || ay >= c->ny || az >= c->nz) continue;
                    if (!c->obstacle[SFX(cell_index)(c, ax, ay, az)]) continue;
                    if (!SFX(cylinder_solid)(c, ax, ay)) continue;
                    const double fq = (double)c->f_post[(size_t)q * c->cells + cell];
                    fx += 2.0 * fq * (double)kCX[q];
                    fy += 2.0 * fq * (double)kCY[q];
                    fz += 2.0 * fq * (double)kCZ[q];
                }
            }
    const double inv_depth = c->nz > 0 ? (1.0 / (double)c->nz) : 1.0;
    c->force[0] = (float)(fx * inv_depth);
    c->force[1] = (float)(fy * inv_depth);
    c->force[2] = (float)(fz * inv_depth);
}

ORACLE_API void SFX(oracle_get_force)(void* h, float* out3) {
    const SFX(Ctx)* c = (const SFX(Ctx)*)h;
    for (int k = 0; k < 3; ++k) out3[k] = c->force[k];
}

/* run_cpu_step (jni.cpp:8480-8507). `packet` may be NULL after the first step,
 * meaning "same payload as last step" (the cached-step semantics). */
ORACLE_API int SFX(oracle_step)(void* h, const float* packet, float* out, float out_scale) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    if (packet) SFX(ingest)(c, packet);
    else if (!c->initialized) return 0;
    if (!c->initialized) SFX(initialize)(c);
    else SFX(temperature_reference)(c);
    /* should_update_temperature (jni.cpp:2191-2196) */
    if (c->thermal_stride <= 1 || c->step_counter % (uint64_t)c->thermal_stride == 0) SFX(update_temperature)(c);
    SFX(collide)(c);
    SFX(benchmark_force)(c);
    SFX(stream)(c);
    if (out) {
        for (size_t cell = 0; cell < c->cells; ++cell) {
            R o[4];
            SFX(cell_output)(c, cell, (R)out_scale, o);
            for (int k = 0; k < 4; ++k) out[cell * (size_t)c->out_ch + k] = (float)o[k];
            for (int k = 4; k < c->out_ch; ++k) out[cell * (size_t)c->out_ch + k] = 0.0f;
        }
    }
    c->step_counter += 1;
    return 1;
}

ORACLE_API int SFX(oracle_run)(void* h, const float* packet, int steps) {
    for (int s = 0; s < steps; ++s)
        if (!SFX(oracle_step)(h, s == 0 ? packet : NULL, NULL, 1.0f)) return 0;
    return 1;
}

/* extract_flow_atlas on the CPU path (jni.cpp:9378-9448); stride 1 = get_flow_state */
ORACLE_API void SFX(oracle_extract_atlas)(void* h, int stride, float* out) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    const int sx = (c->nx + stride - 1) / stride, sy = (c->ny + stride - 1) / stride, sz = (c->nz + stride - 1) / stride;
    size_t dst = 0;
    for (int ax = 0; ax < sx; ++ax) {
        const int gx = (ax * stride < c->nx - 1) ? ax * stride : c->nx - 1;
        for (int ay = 0; ay < sy; ++ay) {
            const int gy = (ay * stride < c->ny - 1) ? ay * stride : c->ny - 1;
            for (int az = 0; az < sz; ++az) {
                const int gz = (az * stride < c->nz - 1) ? az * stride : c->nz - 1;
                R o[4];
                SFX(cell_output)(c, SFX(cell_index)(c, gx, gy, gz), K(1.0), o);
                for (int k = 0; k < 4; ++k) out[dst + k] = (float)o[k];
                dst += 4;
            }
        }
    }
}

/* raw state access for tests: post-stream populations f and macro fields */
ORACLE_API void SFX(oracle_get_f)(void* h, R* f) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    memcpy(f, c->f, 27 * c->cells * sizeof(R));
}
/* post-collision populations of the last step (before streaming) */
ORACLE_API void SFX(oracle_get_f_post)(void* h, R* f) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    memcpy(f, c->f_post, 27 * c->cells * sizeof(R));
}
ORACLE_API void SFX(oracle_set_f)(void* h, const R* f) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    memcpy(c->f, f, 27 * c->cells * sizeof(R));
}
ORACLE_API void SFX(oracle_get_velocity)(void* h, R* u) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    for (size_t i = 0; i < c->cells; ++i) {
        u[3 * i] = c->ux[i]; u[3 * i + 1] = c->uy[i]; u[3 * i + 2] = c->uz[i];
    }
}
ORACLE_API void SFX(oracle_get_temperature_r)(void* h, R* t) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    memcpy(t, c->T, c->cells * sizeof(R));
}
ORACLE_API void SFX(oracle_get_params)(void* h, R* out6) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    out6[0] = c->tau_shear_base; out6[1] = c->tau_normal_base; out6[2] = c->nu_base;
    out6[3] = c->alpha; out6[4] = c->beta; out6[5] = c->sgs_c2;
}

/* Single-cell collision for unit pins: f (27 populations), T, returns f_post.
 * Uses a 1x1x1 context's parameters (ctx `h`). */
ORACLE_API void SFX(oracle_collide_cell)(void* h, const R* f_in, R temperature, R* f_out) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    /* operate on cell 0 of the context storage */
    for (int q = 0; q < 27; ++q) c->f[(size_t)q * c->cells] = f_in[q];
    c->T[0] = temperature;
    const unsigned char ob = c->obstacle[0];
    c->obstacle[0] = 0;
    /* collide all cells; only cell 0 matters to the caller */
    SFX(collide)(c);
    c->obstacle[0] = ob;
    for (int q = 0; q < 27; ++q) f_out[q] = c->f_post[(size_t)q * c->cells];
}

/* feq with the oracle's weights (for table/equilibrium pins against the reference core) */
ORACLE_API R SFX(oracle_feq)(int q, R rho, R ux, R uy, R uz) {
    SFX(Ctx) c;
    SFX(init_weights)(c.w);
    return SFX(feq)(&c, q, rho, ux, uy, uz);
}
ORACLE_API void SFX(oracle_weights)(R* w) { SFX(init_weights)(w); }

/* boundary-reference shell: layer per cell (0 = none) and 4 decoded targets per cell;
 * layers <= 0 removes the shell */
ORACLE_API void SFX(oracle_set_shell)(void* h, const unsigned char* layer, const float* targets, int layers,
                                      float relax) {
    SFX(Ctx)* c = (SFX(Ctx)*)h;
    free(c->shell_layer);
    free(c->shell_target);
    c->shell_layer = NULL;
    c->shell_target = NULL;
    if (layers <= 0) return;
    c->shell_layer = (unsigned char*)malloc(c->cells);
    c->shell_target = (R*)malloc(sizeof(R) * 4 * c->cells);
    memcpy(c->shell_layer, layer, c->cells);
    for (size_t i = 0; i < 4 * c->cells; ++i) c->shell_target[i] = (R)targets[i];
    c->shell_layers = layers;
    c->shell_relax = (R)relax;
}
