// TEST INFRASTRUCTURE ONLY: a C-callable shim over the reference's own
// hydro/thermal core sources (native/src/aero_lbm_hydro_core.cpp,
// native/src/aero_lbm_thermal_core.cpp and their headers), compiled where they
// lie under the reference tree by oracle/Makefile into oracle/_ref/. These two
// translation units need nothing the image lacks; the rest of the reference
// solver (aero_lbm_jni.cpp, aero_lbm_mesoscale.cpp) includes <jni.h> and is
// therefore not built here. Tests use this to pin the oracle's lattice tables,
// equilibrium and transport helpers against the reference itself.
#include "aero_lbm_hydro_core.h"
#include "aero_lbm_thermal_core.h"

namespace hc = aero_lbm::hydro_core;
namespace tc = aero_lbm::thermal_core;

extern "C" {

__attribute__((visibility("default"))) void ref_d3q27_tables(int* cx, int* cy, int* cz, int* opp, float* w) {
    for (int q = 0; q < hc::kQ; ++q) {
        cx[q] = hc::kCx[q];
        cy[q] = hc::kCy[q];
        cz[q] = hc::kCz[q];
        opp[q] = hc::kOpp[q];
        w[q] = hc::kW[q];
    }
}

__attribute__((visibility("default"))) float ref_feq_d3q27(int q, float rho, float ux, float uy, float uz) {
    return hc::equilibrium_d3q27(q, rho, ux, uy, uz);
}

__attribute__((visibility("default"))) float ref_feq_d3q7(int q, float t, float ux, float uy, float uz) {
    return tc::equilibrium_d3q7(q, t, ux, uy, uz);
}

__attribute__((visibility("default"))) void ref_derive_transport(
    int nx, int ny, int nz, float dx, float dt, float nu, float pr, float prt, float* out5) {
    const hc::TransportSpec spec{nx, ny, nz, dx, dt, nu, pr, prt};
    const auto h = hc::derive_hydro_transport(spec);
    const auto t = tc::derive_thermal_transport(spec);
    out5[0] = h.velocity_scale_m_s_per_lattice;
    out5[1] = h.nu_molecular_lattice;
    out5[2] = t.alpha_molecular_lattice;
    out5[3] = h.tau_shear_molecular;
    out5[4] = t.tau_thermal_molecular;
}

__attribute__((visibility("default"))) float ref_boussinesq_beta(float g, float bt, float ts, float dt, float dx) {
    return tc::runtime_boussinesq_beta(g, bt, ts, dt, dx);
}

__attribute__((visibility("default"))) float ref_lattice_diffusivity(float d, float dx, float dt) {
    return hc::physical_diffusivity_to_lattice(d, dx, dt);
}

}
