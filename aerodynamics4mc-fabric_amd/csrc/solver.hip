// Wind-tunnel solver C ABI (ref: native/include/aero_solver_capi.h,
// native/src/aero_solver_capi.cpp): a thin host layer over this library's
// aero_lbm_* calls. A solver holds the grid, a solid mask and an 11-channel
// payload packet; each step configures the benchmark boundary for the requested
// wind-tunnel mode and advances one lattice context. After the first upload the
// packet stays on the device and steps run from it (the reference does the same
// on its GPU runtime).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aero_solver_capi.h"

namespace {

// payload channels (ref: native/src/aero_lbm_jni.cpp:64-74)
constexpr int kPacketChannels = 11;
constexpr int kChObstacle = 0, kChUx = 5, kChUy = 6, kChUz = 7, kChPressure = 8;
constexpr int kFlowChannels = AERO_SOLVER_FLOW_CHANNELS;
// defaults and lattice-viscosity bounds (ref: aero_solver_capi.cpp:30-35)
constexpr float kDxDefault = 1.0f, kDtDefault = 0.05f;
constexpr float kRhoDefault = 1.225f, kNuDefault = 1.5e-5f;
constexpr float kNuLatticeMin = 1.0e-7f, kNuLatticeMax = 0.12f;

struct Tunnel {
    long long key = 0;  // lattice context key (the handle, as the reference)
    AeroGridDesc grid{};
    int cells = 0;
    bool custom_flow = false;    // a caller's flow state seeds the packet
    bool packet_dirty = true;    // the packet must be rebuilt before the next step
    bool uploaded = false;       // the packet is on the device
    AeroBoundaryDesc packet_bc{};  // boundary the packet's flow channels were built for
    std::vector<uint8_t> solid;
    std::vector<float> packet;
};

std::mutex g_mu;
std::unordered_map<long long, std::unique_ptr<Tunnel>> g_tunnels;
long long g_next = 1;
std::string g_error;

void fail(const std::string& msg) { g_error = msg; }

float finite_or(float v, float fallback) { return std::isfinite(v) ? v : fallback; }

// m/s -> lattice units
float lattice(const AeroGridDesc& g, float v) { return finite_or(v, 0.0f) / (g.dx / g.dt); }

bool check_grid(const AeroGridDesc& g, int* cells) {
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0) {
        fail("invalid grid dimensions");
        return false;
    }
    if (!std::isfinite(g.dx) || !(g.dx > 0.0f) || !std::isfinite(g.dt) || !(g.dt > 0.0f)) {
        fail("invalid grid dx/dt");
        return false;
    }
    const long long n = (long long)g.nx * g.ny * g.nz;
    if (n > (long long)std::numeric_limits<int>::max()) {
        fail("grid cell count overflows int");
        return false;
    }
    *cells = (int)n;
    return true;
}

Tunnel* find(long long h) {
    auto it = g_tunnels.find(h);
    if (it == g_tunnels.end() || !it->second) {
        fail("invalid solver handle");
        return nullptr;
    }
    return it->second.get();
}

void set_cell_flow(Tunnel& t, size_t cell, float vx, float vy, float vz, float p) {
    float* c = t.packet.data() + cell * kPacketChannels;
    c[kChUx] = lattice(t.grid, vx);
    c[kChUy] = lattice(t.grid, vy);
    c[kChUz] = lattice(t.grid, vz);
    c[kChPressure] = finite_or(p, 0.0f);
}

// obstacle channel from the mask, fan / source / temperature channels zero; the
// flow channels take the inlet state when overwrite_flow (solid cells: zero)
void fill_packet(Tunnel& t, const AeroBoundaryDesc& bc, bool overwrite_flow) {
    for (size_t cell = 0; cell < (size_t)t.cells; ++cell) {
        float* c = t.packet.data() + cell * kPacketChannels;
        const bool solid = t.solid[cell] != 0;
        for (int k = 0; k < kPacketChannels; ++k)
            if (k < kChUx || k > kChPressure) c[k] = 0.0f;
        c[kChObstacle] = solid ? 1.0f : 0.0f;
        if (solid) set_cell_flow(t, cell, 0.0f, 0.0f, 0.0f, 0.0f);
        else if (overwrite_flow) set_cell_flow(t, cell, bc.inlet_vx, bc.inlet_vy, bc.inlet_vz, bc.outlet_pressure);
    }
}

bool same_packet_boundary(const AeroBoundaryDesc& a, const AeroBoundaryDesc& b) {
    return a.mode == b.mode && a.inlet_vx == b.inlet_vx && a.inlet_vy == b.inlet_vy && a.inlet_vz == b.inlet_vz &&
           a.outlet_pressure == b.outlet_pressure;
}

AeroLbmBoundaryFaceConfig face(int hydro) {
    AeroLbmBoundaryFaceConfig f{};
    f.hydrodynamic_kind = hydro;
    f.thermal_kind = AERO_LBM_THERMAL_BOUNDARY_DISABLED;
    return f;
}

// benchmark configuration of a wind-tunnel step (ref: aero_solver_capi.cpp:209-258):
// physics switched off but the collision and the face rules; Reynolds number from
// the inlet speed and the lattice viscosity, so the library derives that viscosity
bool configure(const Tunnel& t, const AeroBoundaryDesc& bc) {
    AeroLbmBenchmarkConfig cfg{};
    aero_lbm_benchmark_default_config(&cfg);
    cfg.enabled = 1;
    cfg.preset = AERO_LBM_BENCHMARK_PRESET_NONE;
    cfg.flags |= AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING | AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_NOISE |
                 AERO_LBM_BENCHMARK_FLAG_DISABLE_SPONGE | AERO_LBM_BENCHMARK_FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND |
                 AERO_LBM_BENCHMARK_FLAG_DISABLE_SGS | AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE |
                 AERO_LBM_BENCHMARK_FLAG_DISABLE_BUOYANCY;
    cfg.flags &= ~(uint32_t)AERO_LBM_BENCHMARK_FLAG_DISABLE_CONVECTIVE_OUTFLOW;
    cfg.reference_density = std::max(1.0e-6f, finite_or(bc.density, kRhoDefault));
    cfg.reference_temperature = 0.0f;
    cfg.reference_length = 1.0f;
    const float u[3] = {lattice(t.grid, bc.inlet_vx), lattice(t.grid, bc.inlet_vy), lattice(t.grid, bc.inlet_vz)};
    for (int k = 0; k < 3; ++k) cfg.initial_velocity[k] = u[k];
    cfg.initial_pressure = finite_or(bc.outlet_pressure, 0.0f);
    const float speed = std::max(std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), 1.0e-6f);
    // a non-positive viscosity falls back to the default value unscaled, as the reference
    const float visc = finite_or(bc.viscosity, kNuDefault);
    const float nu = std::clamp(visc > 0.0f ? visc * t.grid.dt / (t.grid.dx * t.grid.dx) : kNuDefault, kNuLatticeMin,
                                kNuLatticeMax);
    cfg.reynolds_number = std::max(1.0e-6f, speed / nu);
    cfg.mach_number = std::clamp(speed / std::sqrt(1.0f / 3.0f), 1.0e-4f, 0.30f);
    AeroLbmBoundaryFaceConfig* faces[6] = {&cfg.x_min, &cfg.x_max, &cfg.y_min, &cfg.y_max, &cfg.z_min, &cfg.z_max};
    switch (bc.mode == 0 ? AERO_SOLVER_BOUNDARY_WIND_TUNNEL : bc.mode) {
        case AERO_SOLVER_BOUNDARY_CLOSED:
            for (auto* f : faces) *f = face(AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK);
            break;
        case AERO_SOLVER_BOUNDARY_PERIODIC:
            for (auto* f : faces) *f = face(AERO_LBM_HYDRO_BOUNDARY_PERIODIC);
            break;
        case AERO_SOLVER_BOUNDARY_WIND_TUNNEL_PRESSURE_OUTLET:
            for (auto* f : faces) *f = face(AERO_LBM_HYDRO_BOUNDARY_SYMMETRY);
            cfg.x_min = face(AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET);
            cfg.x_max = face(AERO_LBM_HYDRO_BOUNDARY_PRESSURE_DIRICHLET);
            cfg.x_max.pressure = finite_or(bc.outlet_pressure, 0.0f);
            break;
        case AERO_SOLVER_BOUNDARY_WIND_TUNNEL:
            for (auto* f : faces) *f = face(AERO_LBM_HYDRO_BOUNDARY_SYMMETRY);
            cfg.x_min = face(AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET);
            cfg.x_max = face(AERO_LBM_HYDRO_BOUNDARY_CONVECTIVE_OUTFLOW);
            break;
        default:
            fail("invalid wind-tunnel boundary mode");
            return false;
    }
    for (int k = 0; k < 3; ++k) cfg.x_min.velocity[k] = u[k];
    return aero_lbm_benchmark_set_config(&cfg) != 0;
}

bool device_runtime() {
    const char* info = aero_lbm_runtime_info();
    return info && std::strncmp(info, "hip|", 4) == 0;
}

bool run_steps(Tunnel& t, const AeroBoundaryDesc& bc, int steps, float* out, int out_count) {
    if (steps <= 0) {
        fail("steps must be positive");
        return false;
    }
    if (out && out_count != t.cells * kFlowChannels) {
        fail("invalid output flow buffer");
        return false;
    }
    if (!configure(t, bc)) {
        if (g_error.empty()) fail(std::string("failed to configure boundary: ") + aero_lbm_last_error());
        return false;
    }
    const bool overwrite = !t.custom_flow;
    if (t.packet_dirty || (overwrite && !same_packet_boundary(t.packet_bc, bc))) {
        fill_packet(t, bc, overwrite);
        t.packet_bc = bc;
        t.packet_dirty = false;
        t.uploaded = false;
    }
    const float scale = t.grid.dx / t.grid.dt;  // lattice -> m/s of the output velocity
    const AeroGridDesc& g = t.grid;
    for (int s = 0; s < steps; ++s) {
        float* dst = (out && s + 1 == steps) ? out : nullptr;
        bool ok = t.uploaded && device_runtime() &&
                  aero_lbm_step_rect_cached_scaled(g.nx, g.ny, g.nz, t.key, scale, dst) != 0;
        if (!ok) {
            if (!aero_lbm_step_rect_scaled(t.packet.data(), g.nx, g.ny, g.nz, t.key, scale, dst)) {
                fail(std::string("aero_lbm_step_rect failed: ") + aero_lbm_last_error());
                return false;
            }
            t.uploaded = device_runtime();
        }
    }
    return true;
}

AeroBoundaryDesc effective_boundary(const AeroBoundaryDesc* bc) {
    AeroBoundaryDesc b{};
    aero_solver_default_boundary(&b);
    if (bc) b = *bc;
    return b;
}

}  // namespace

extern "C" {

AERO_LBM_CAPI_EXPORT void aero_solver_default_grid(AeroGridDesc* g) {
    if (!g) return;
    g->nx = g->ny = g->nz = 64;
    g->dx = kDxDefault;
    g->dt = kDtDefault;
}

AERO_LBM_CAPI_EXPORT void aero_solver_default_boundary(AeroBoundaryDesc* b) {
    if (!b) return;
    b->mode = AERO_SOLVER_BOUNDARY_WIND_TUNNEL;
    b->inlet_vx = 5.0f;
    b->inlet_vy = 0.0f;
    b->inlet_vz = 0.0f;
    b->outlet_pressure = 0.0f;
    b->density = kRhoDefault;
    b->viscosity = kNuDefault;
}

AERO_LBM_CAPI_EXPORT int aero_solver_create(int nx, int ny, int nz, float dx, float dt, long long* out_handle) {
    const AeroGridDesc g{nx, ny, nz, dx, dt};
    return aero_solver_create_with_grid(&g, out_handle);
}

AERO_LBM_CAPI_EXPORT int aero_solver_create_with_grid(const AeroGridDesc* grid, long long* out_handle) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_error.clear();
    if (!grid || !out_handle) {
        fail("missing grid or output handle");
        return AERO_SOLVER_STATUS_ERROR;
    }
    int cells = 0;
    if (!check_grid(*grid, &cells)) return AERO_SOLVER_STATUS_ERROR;
    if (!aero_lbm_init_rect(grid->nx, grid->ny, grid->nz, kPacketChannels, kFlowChannels)) {
        fail(std::string("aero_lbm_init_rect failed: ") + aero_lbm_last_error());
        return AERO_SOLVER_STATUS_ERROR;
    }
    // one solver at a time: earlier ones end here, with their lattice contexts
    for (auto& kv : g_tunnels) aero_lbm_release_context(kv.second->key);
    g_tunnels.clear();
    auto t = std::make_unique<Tunnel>();
    const long long h = g_next++;
    if (g_next <= 0) g_next = 1;
    t->key = h;
    t->grid = *grid;
    t->cells = cells;
    t->solid.assign((size_t)cells, 0);
    t->packet.assign((size_t)cells * kPacketChannels, 0.0f);
    const AeroBoundaryDesc bc = effective_boundary(nullptr);
    fill_packet(*t, bc, true);
    t->packet_bc = bc;
    t->packet_dirty = false;
    g_tunnels[h] = std::move(t);
    *out_handle = h;
    return AERO_SOLVER_STATUS_OK;
}

AERO_LBM_CAPI_EXPORT int aero_solver_set_solid_mask(long long handle, const uint8_t* mask, int cell_count) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_error.clear();
    Tunnel* t = find(handle);
    if (!t) return AERO_SOLVER_STATUS_ERROR;
    if (!mask || cell_count != t->cells) {
        fail("invalid solid mask buffer");
        return AERO_SOLVER_STATUS_ERROR;
    }
    t->solid.assign(mask, mask + cell_count);
    t->custom_flow = false;
    t->packet_dirty = true;
    t->uploaded = false;
    aero_lbm_release_context(t->key);
    return AERO_SOLVER_STATUS_OK;
}

AERO_LBM_CAPI_EXPORT int aero_solver_set_flow_state(long long handle, const float* flow, int value_count) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_error.clear();
    Tunnel* t = find(handle);
    if (!t) return AERO_SOLVER_STATUS_ERROR;
    if (!flow || value_count != t->cells * kFlowChannels) {
        fail("invalid flow state buffer");
        return AERO_SOLVER_STATUS_ERROR;
    }
    for (size_t cell = 0; cell < (size_t)t->cells; ++cell) {
        const float* f = flow + cell * kFlowChannels;
        if (t->solid[cell]) set_cell_flow(*t, cell, 0.0f, 0.0f, 0.0f, 0.0f);
        else set_cell_flow(*t, cell, f[0], f[1], f[2], f[3]);
    }
    t->custom_flow = true;
    t->packet_dirty = false;
    t->uploaded = false;
    aero_lbm_release_context(t->key);
    return AERO_SOLVER_STATUS_OK;
}

AERO_LBM_CAPI_EXPORT int aero_solver_step_wind_tunnel(long long handle, const AeroBoundaryDesc* boundary, int steps,
                                                      float* out_flow, int out_value_count) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_error.clear();
    Tunnel* t = find(handle);
    if (!t) return AERO_SOLVER_STATUS_ERROR;
    return run_steps(*t, effective_boundary(boundary), steps, out_flow, out_value_count) ? AERO_SOLVER_STATUS_OK
                                                                                         : AERO_SOLVER_STATUS_ERROR;
}

AERO_LBM_CAPI_EXPORT int aero_solver_advance_wind_tunnel(long long handle, const AeroBoundaryDesc* boundary,
                                                         int steps) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_error.clear();
    Tunnel* t = find(handle);
    if (!t) return AERO_SOLVER_STATUS_ERROR;
    return run_steps(*t, effective_boundary(boundary), steps, nullptr, 0) ? AERO_SOLVER_STATUS_OK
                                                                          : AERO_SOLVER_STATUS_ERROR;
}

AERO_LBM_CAPI_EXPORT int aero_solver_run_wind_tunnel(const AeroStepInput* in, AeroStepOutput* out) {
    if (!in || !out || !out->flow_out) {
        std::lock_guard<std::mutex> lk(g_mu);
        fail("missing step input or output");
        return AERO_SOLVER_STATUS_ERROR;
    }
    long long h = 0;
    if (!aero_solver_create_with_grid(&in->grid, &h)) {
        out->status = AERO_SOLVER_STATUS_ERROR;
        return AERO_SOLVER_STATUS_ERROR;
    }
    const int cells = in->grid.nx * in->grid.ny * in->grid.nz;
    if ((in->solid_mask && !aero_solver_set_solid_mask(h, in->solid_mask, cells)) ||
        (in->prev_flow && !aero_solver_set_flow_state(h, in->prev_flow, cells * kFlowChannels))) {
        aero_solver_destroy(h);
        out->status = AERO_SOLVER_STATUS_ERROR;
        return AERO_SOLVER_STATUS_ERROR;
    }
    const int ok = aero_solver_step_wind_tunnel(h, &in->boundary, in->steps, out->flow_out, cells * kFlowChannels);
    float vmax = 0.0f;
    if (ok) {
        for (size_t c = 0; c < (size_t)cells; ++c) {
            const float* f = out->flow_out + c * kFlowChannels;
            const float s = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
            if (std::isfinite(s)) vmax = std::max(vmax, s);
        }
    }
    out->max_velocity = vmax;
    out->status = ok;
    aero_solver_destroy(h);
    return ok;
}

AERO_LBM_CAPI_EXPORT void aero_solver_destroy(long long handle) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tunnels.find(handle);
    if (it == g_tunnels.end()) return;
    aero_lbm_release_context(it->second->key);
    g_tunnels.erase(it);
}

AERO_LBM_CAPI_EXPORT const char* aero_solver_last_error(void) {
    return g_error.empty() ? aero_lbm_last_error() : g_error.c_str();
}

AERO_LBM_CAPI_EXPORT const char* aero_solver_runtime_info(void) { return aero_lbm_runtime_info(); }

AERO_LBM_CAPI_EXPORT int aero_solver_finish(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_error.clear();
    if (!aero_lbm_finish()) {
        fail(std::string("aero_lbm_finish failed: ") + aero_lbm_last_error());
        return AERO_SOLVER_STATUS_ERROR;
    }
    return AERO_SOLVER_STATUS_OK;
}

}  // extern "C"
