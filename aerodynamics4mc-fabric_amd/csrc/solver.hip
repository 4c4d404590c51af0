// Wind-tunnel solver C ABI (include/aero_solver_capi.h; the reference's host layer
// is native/src/aero_solver_capi.cpp). Every solver keeps its state on the device:
// the solid mask as one byte per cell and the 11-channel payload packet the lattice
// context ingests. Two kernels edit the packet in place, one rebuilding it from the
// mask (obstacle channel, inlet flow) and one writing a caller's flow state into it,
// so the host uploads 1 byte per cell for a mask and 16 for a flow state instead of
// a 44-byte packet per cell. The packet reaches the lattice context by a device copy
// (aero::step_rect_device_packet) the first time a context steps it; later steps run
// from the context's resident copy.
//
// Observable rules kept from the reference layer (aero_solver_capi.cpp:130-160,
// 209-258, 403-470): the benchmark configuration each mode selects, the packet's
// channel contents, when the packet is rebuilt, and that a flow state applies on top
// of the packet's current obstacle channel (which only a rebuild refreshes).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/aero_solver_capi.h"
#include "runtime_internal.h"

namespace {

// payload packet layout (ref: native/src/aero_lbm_jni.cpp:64-74): obstacle, fan
// (4 channels), velocity (3, lattice units), pressure, source, temperature
constexpr int kChannels = 11;
constexpr int kChFlow = 5;  // first of the 4 flow channels (vx, vy, vz, pressure)

constexpr float kDefaultDx = 1.0f, kDefaultDt = 0.05f;
constexpr float kDefaultDensity = 1.225f, kDefaultViscosity = 1.5e-5f;
// lattice viscosity bounds of the Reynolds number (ref: aero_solver_capi.cpp:30-35)
constexpr float kLatticeNuMin = 1.0e-7f, kLatticeNuMax = 0.12f;

// ---------------------------------------------------------------- mode table
// Hydrodynamic rule of each face (x-, x+, y-, y+, z-, z+) per wind-tunnel mode
// (ref: aero_solver_capi.cpp:209-258); every face's thermal rule is "disabled".
struct ModeFaces {
    int mode;
    int hydro[6];
};
constexpr int kVel = AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET, kConv = AERO_LBM_HYDRO_BOUNDARY_CONVECTIVE_OUTFLOW,
              kPres = AERO_LBM_HYDRO_BOUNDARY_PRESSURE_DIRICHLET, kSym = AERO_LBM_HYDRO_BOUNDARY_SYMMETRY,
              kWall = AERO_LBM_HYDRO_BOUNDARY_BOUNCE_BACK, kPer = AERO_LBM_HYDRO_BOUNDARY_PERIODIC;
constexpr ModeFaces kModeFaces[] = {
    {AERO_SOLVER_BOUNDARY_WIND_TUNNEL, {kVel, kConv, kSym, kSym, kSym, kSym}},
    {AERO_SOLVER_BOUNDARY_WIND_TUNNEL_PRESSURE_OUTLET, {kVel, kPres, kSym, kSym, kSym, kSym}},
    {AERO_SOLVER_BOUNDARY_CLOSED, {kWall, kWall, kWall, kWall, kWall, kWall}},
    {AERO_SOLVER_BOUNDARY_PERIODIC, {kPer, kPer, kPer, kPer, kPer, kPer}},
};

const ModeFaces* faces_of(int mode) {
    if (mode == 0) mode = AERO_SOLVER_BOUNDARY_WIND_TUNNEL;  // 0 selects the default mode
    for (const ModeFaces& m : kModeFaces)
        if (m.mode == mode) return &m;
    return nullptr;
}

float finite_or(float v, float fallback) { return std::isfinite(v) ? v : fallback; }

// The benchmark configuration of a wind-tunnel step: every physics switch off but
// the collision, the face rules and convective outflow; Reynolds and Mach numbers
// from the inlet speed and the lattice viscosity, from which the library derives
// the relaxation time. `speed_unit` = dx / dt converts m/s to lattice units.
AeroLbmBenchmarkConfig tunnel_config(const AeroGridDesc& g, const AeroBoundaryDesc& bc, const ModeFaces& mf) {
    AeroLbmBenchmarkConfig cfg{};
    aero_lbm_benchmark_default_config(&cfg);
    cfg.enabled = 1;
    cfg.preset = AERO_LBM_BENCHMARK_PRESET_NONE;
    const uint32_t off = AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_FORCING | AERO_LBM_BENCHMARK_FLAG_DISABLE_FAN_NOISE |
                         AERO_LBM_BENCHMARK_FLAG_DISABLE_SPONGE |
                         AERO_LBM_BENCHMARK_FLAG_DISABLE_OBSTACLE_BOUNCE_BLEND | AERO_LBM_BENCHMARK_FLAG_DISABLE_SGS |
                         AERO_LBM_BENCHMARK_FLAG_DISABLE_INTERNAL_THERMAL_SOURCE |
                         AERO_LBM_BENCHMARK_FLAG_DISABLE_BUOYANCY;
    cfg.flags = (cfg.flags | off) & ~(uint32_t)AERO_LBM_BENCHMARK_FLAG_DISABLE_CONVECTIVE_OUTFLOW;
    const float speed_unit = g.dx / g.dt;
    const float u[3] = {finite_or(bc.inlet_vx, 0.0f) / speed_unit, finite_or(bc.inlet_vy, 0.0f) / speed_unit,
                        finite_or(bc.inlet_vz, 0.0f) / speed_unit};
    cfg.reference_density = std::max(1.0e-6f, finite_or(bc.density, kDefaultDensity));
    cfg.reference_temperature = 0.0f;
    cfg.reference_length = 1.0f;
    std::copy(u, u + 3, cfg.initial_velocity);
    cfg.initial_pressure = finite_or(bc.outlet_pressure, 0.0f);
    const float speed = std::max(std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), 1.0e-6f);
    // a viscosity <= 0 selects the default value, taken as a lattice viscosity
    const float visc = finite_or(bc.viscosity, kDefaultViscosity);
    const float nu_lattice = visc > 0.0f ? visc * g.dt / (g.dx * g.dx) : kDefaultViscosity;
    cfg.reynolds_number = std::max(1.0e-6f, speed / std::clamp(nu_lattice, kLatticeNuMin, kLatticeNuMax));
    cfg.mach_number = std::clamp(speed / std::sqrt(1.0f / 3.0f), 1.0e-4f, 0.30f);
    AeroLbmBoundaryFaceConfig* faces[6] = {&cfg.x_min, &cfg.x_max, &cfg.y_min, &cfg.y_max, &cfg.z_min, &cfg.z_max};
    for (int k = 0; k < 6; ++k) {
        *faces[k] = AeroLbmBoundaryFaceConfig{};
        faces[k]->hydrodynamic_kind = mf.hydro[k];
        faces[k]->thermal_kind = AERO_LBM_THERMAL_BOUNDARY_DISABLED;
        if (mf.hydro[k] == kPres) faces[k]->pressure = finite_or(bc.outlet_pressure, 0.0f);
    }
    std::copy(u, u + 3, cfg.x_min.velocity);  // the inlet velocity, set whatever the x- rule
    return cfg;
}

// ---------------------------------------------------------------- device kernels
// Rebuild: obstacle channel from the mask, fan / source / temperature channels zero,
// flow channels zero in solid cells and, with inlet_flow, the inlet state elsewhere.
__global__ void __launch_bounds__(256) rebuild_packet_kernel(int cells, const uint8_t* __restrict__ solid,
                                                             float* __restrict__ packet, int inlet_flow, float vx,
                                                             float vy, float vz, float p) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    float* c = packet + (size_t)cell * kChannels;
    const bool s = solid[cell] != 0;
    c[0] = s ? 1.0f : 0.0f;
    for (int k = 1; k < kChFlow; ++k) c[k] = 0.0f;
    c[9] = 0.0f;
    c[10] = 0.0f;
    if (s || inlet_flow) {
        c[kChFlow + 0] = s ? 0.0f : vx;
        c[kChFlow + 1] = s ? 0.0f : vy;
        c[kChFlow + 2] = s ? 0.0f : vz;
        c[kChFlow + 3] = s ? 0.0f : p;
    }
}

// A caller's flow state (vx, vy, vz in m/s, pressure) into the flow channels:
// velocities divided by dx / dt, non-finite values as zero, solid cells zero.
__global__ void __launch_bounds__(256) flow_state_kernel(int cells, const uint8_t* __restrict__ solid,
                                                         const float4* __restrict__ flow, float speed_unit,
                                                         float* __restrict__ packet) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= cells) return;
    float* c = packet + (size_t)cell * kChannels + kChFlow;
    const float4 v = flow[cell];
    const bool s = solid[cell] != 0;
    c[0] = s ? 0.0f : (isfinite(v.x) ? v.x : 0.0f) / speed_unit;
    c[1] = s ? 0.0f : (isfinite(v.y) ? v.y : 0.0f) / speed_unit;
    c[2] = s ? 0.0f : (isfinite(v.z) ? v.z : 0.0f) / speed_unit;
    c[3] = s ? 0.0f : (isfinite(v.w) ? v.w : 0.0f);
}

// ---------------------------------------------------------------- solver state
class DeviceBuffer {
   public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (p_) (void)hipFree(p_);
    }
    bool allocate(size_t bytes) { return hipMalloc(&p_, bytes) == hipSuccess; }
    template <class T>
    T* as() const {
        return static_cast<T*>(p_);
    }

   private:
    void* p_ = nullptr;
};

// What the device packet's flow channels hold.
enum class FlowSource { Inlet, Caller };

struct WindTunnel {
    AeroGridDesc grid{};
    int cells = 0;
    DeviceBuffer solid;   // current solid mask, 1 byte per cell
    DeviceBuffer packet;  // payload packet, kChannels floats per cell
    FlowSource flow = FlowSource::Inlet;
    AeroBoundaryDesc inlet{};  // boundary whose inlet state the flow channels hold (FlowSource::Inlet)
    bool obstacle_stale = false;  // the mask changed since the packet's last rebuild
    bool resident = false;        // the lattice context holds this packet
};

struct Registry {
    std::mutex mu;
    std::map<long long, WindTunnel> tunnels;
    long long serial = 0;  // handles count up from 1 and are never reused
    std::string error;
};

Registry& reg() {
    static Registry r;
    return r;
}

bool fail(const std::string& msg) {
    reg().error = msg;
    return false;
}

bool device_ok(hipError_t e, const char* what) {
    return e == hipSuccess || fail(std::string(what) + ": " + hipGetErrorString(e));
}

dim3 cell_grid(int cells) { return dim3((unsigned)((cells + 255) / 256)); }

AeroBoundaryDesc boundary_or_default(const AeroBoundaryDesc* bc) {
    AeroBoundaryDesc b{};
    if (bc) b = *bc;
    else aero_solver_default_boundary(&b);
    return b;
}

// the inlet fields the packet's flow channels depend on
bool same_inlet(const AeroBoundaryDesc& a, const AeroBoundaryDesc& b) {
    return a.mode == b.mode && a.inlet_vx == b.inlet_vx && a.inlet_vy == b.inlet_vy && a.inlet_vz == b.inlet_vz &&
           a.outlet_pressure == b.outlet_pressure;
}

bool rebuild_packet(WindTunnel& t, const AeroBoundaryDesc& bc) {
    const float speed_unit = t.grid.dx / t.grid.dt;
    const bool inlet = t.flow == FlowSource::Inlet;
    hipLaunchKernelGGL(rebuild_packet_kernel, cell_grid(t.cells), dim3(256), 0, aero::rt().stream, t.cells,
                       t.solid.as<uint8_t>(), t.packet.as<float>(), inlet ? 1 : 0,
                       finite_or(bc.inlet_vx, 0.0f) / speed_unit, finite_or(bc.inlet_vy, 0.0f) / speed_unit,
                       finite_or(bc.inlet_vz, 0.0f) / speed_unit, finite_or(bc.outlet_pressure, 0.0f));
    if (!device_ok(hipGetLastError(), "packet rebuild")) return false;
    if (inlet) t.inlet = bc;
    t.obstacle_stale = false;
    t.resident = false;
    return true;
}

WindTunnel* lookup(long long handle) {
    auto it = reg().tunnels.find(handle);
    if (it == reg().tunnels.end()) {
        fail("invalid solver handle");
        return nullptr;
    }
    return &it->second;
}

bool valid_grid(const AeroGridDesc& g, int* cells) {
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0) return fail("invalid grid dimensions");
    const bool dx_ok = std::isfinite(g.dx) && g.dx > 0.0f, dt_ok = std::isfinite(g.dt) && g.dt > 0.0f;
    if (!dx_ok || !dt_ok) return fail("invalid grid dx/dt");
    const long long n = (long long)g.nx * g.ny * g.nz;
    if (n > std::numeric_limits<int>::max()) return fail("grid cell count overflows int");
    *cells = (int)n;
    return true;
}

// steps of a solver: configure the mode, rebuild the packet if the mask or (for an
// inlet-driven flow) the inlet changed, then step the lattice context; the packet
// is copied into the context when the context does not hold it yet
bool advance(long long handle, WindTunnel& t, const AeroBoundaryDesc& bc, int steps, float* out, int out_count) {
    if (steps <= 0) return fail("steps must be positive");
    if (out && out_count != t.cells * AERO_SOLVER_FLOW_CHANNELS) return fail("invalid output flow buffer");
    const ModeFaces* mf = faces_of(bc.mode);
    if (!mf) return fail("invalid wind-tunnel boundary mode");
    const AeroLbmBenchmarkConfig cfg = tunnel_config(t.grid, bc, *mf);
    if (!aero_lbm_benchmark_set_config(&cfg))
        return fail(std::string("failed to configure boundary: ") + aero_lbm_last_error());
    if (t.obstacle_stale || (t.flow == FlowSource::Inlet && !same_inlet(t.inlet, bc))) {
        if (!rebuild_packet(t, bc)) return false;
    }
    const AeroGridDesc& g = t.grid;
    const float speed_unit = g.dx / g.dt;  // lattice -> m/s of the output velocity
    for (int s = 0; s < steps; ++s) {
        const bool ingest = !t.resident || !aero::context_resident(handle, g.nx, g.ny, g.nz);
        float* dst = s + 1 == steps ? out : nullptr;
        if (!aero::step_rect_device_packet(ingest ? t.packet.as<float>() : nullptr, g.nx, g.ny, g.nz, handle,
                                           speed_unit, dst)) {
            t.resident = false;
            return fail(std::string("aero_lbm_step_rect failed: ") + aero_lbm_last_error());
        }
        t.resident = true;
    }
    return true;
}

int status(bool ok) { return ok ? AERO_SOLVER_STATUS_OK : AERO_SOLVER_STATUS_ERROR; }

}  // namespace

extern "C" {

AERO_LBM_CAPI_EXPORT void aero_solver_default_grid(AeroGridDesc* g) {
    if (g) *g = AeroGridDesc{64, 64, 64, kDefaultDx, kDefaultDt};
}

AERO_LBM_CAPI_EXPORT void aero_solver_default_boundary(AeroBoundaryDesc* b) {
    if (b) *b = AeroBoundaryDesc{AERO_SOLVER_BOUNDARY_WIND_TUNNEL, 5.0f, 0.0f, 0.0f, 0.0f, kDefaultDensity,
                                 kDefaultViscosity};
}

AERO_LBM_CAPI_EXPORT int aero_solver_create(int nx, int ny, int nz, float dx, float dt, long long* out_handle) {
    const AeroGridDesc g{nx, ny, nz, dx, dt};
    return aero_solver_create_with_grid(&g, out_handle);
}

AERO_LBM_CAPI_EXPORT int aero_solver_create_with_grid(const AeroGridDesc* grid, long long* out_handle) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    r.error.clear();
    if (!grid || !out_handle) return status(fail("missing grid or output handle"));
    int cells = 0;
    if (!valid_grid(*grid, &cells)) return status(false);
    // the library holds one grid shape: re-initialising it ends every solver
    r.tunnels.clear();
    if (!aero_lbm_init_rect(grid->nx, grid->ny, grid->nz, kChannels, AERO_SOLVER_FLOW_CHANNELS))
        return status(fail(std::string("aero_lbm_init_rect failed: ") + aero_lbm_last_error()));
    const long long handle = ++r.serial;
    WindTunnel& t = r.tunnels[handle];
    t.grid = *grid;
    t.cells = cells;
    hipStream_t st = aero::rt().stream;
    const bool ok = t.solid.allocate((size_t)cells) && t.packet.allocate((size_t)cells * kChannels * sizeof(float)) &&
                    device_ok(hipMemsetAsync(t.solid.as<uint8_t>(), 0, (size_t)cells, st), "mask clear") &&
                    rebuild_packet(t, boundary_or_default(nullptr));
    if (!ok) {
        if (r.error.empty()) fail("device allocation for the solver failed");
        r.tunnels.erase(handle);
        return status(false);
    }
    *out_handle = handle;
    return status(true);
}

AERO_LBM_CAPI_EXPORT int aero_solver_set_solid_mask(long long handle, const uint8_t* mask, int cell_count) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    r.error.clear();
    WindTunnel* t = lookup(handle);
    if (!t) return status(false);
    if (!mask || cell_count != t->cells) return status(fail("invalid solid mask buffer"));
    hipStream_t st = aero::rt().stream;
    if (!device_ok(hipMemcpyAsync(t->solid.as<uint8_t>(), mask, (size_t)cell_count, hipMemcpyHostToDevice, st),
                   "mask upload") ||
        !device_ok(hipStreamSynchronize(st), "mask upload"))
        return status(false);
    // a new mask restarts from the inlet state at the next step's rebuild
    t->flow = FlowSource::Inlet;
    t->obstacle_stale = true;
    t->resident = false;
    aero_lbm_release_context(handle);
    return status(true);
}

AERO_LBM_CAPI_EXPORT int aero_solver_set_flow_state(long long handle, const float* flow, int value_count) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    r.error.clear();
    WindTunnel* t = lookup(handle);
    if (!t) return status(false);
    if (!flow || value_count != t->cells * AERO_SOLVER_FLOW_CHANNELS) return status(fail("invalid flow state buffer"));
    DeviceBuffer staged;
    const size_t bytes = (size_t)value_count * sizeof(float);
    hipStream_t st = aero::rt().stream;
    if (!staged.allocate(bytes)) return status(fail("device allocation for the flow state failed"));
    if (!device_ok(hipMemcpyAsync(staged.as<float>(), flow, bytes, hipMemcpyHostToDevice, st), "flow upload"))
        return status(false);
    hipLaunchKernelGGL(flow_state_kernel, cell_grid(t->cells), dim3(256), 0, st, t->cells, t->solid.as<uint8_t>(),
                       staged.as<float4>(), t->grid.dx / t->grid.dt, t->packet.as<float>());
    if (!device_ok(hipGetLastError(), "flow state") || !device_ok(hipStreamSynchronize(st), "flow state"))
        return status(false);
    // the packet keeps its obstacle channel: a pending mask change is dropped with
    // the rebuild it was waiting for, as in the reference layer
    t->flow = FlowSource::Caller;
    t->obstacle_stale = false;
    t->resident = false;
    aero_lbm_release_context(handle);
    return status(true);
}

AERO_LBM_CAPI_EXPORT int aero_solver_step_wind_tunnel(long long handle, const AeroBoundaryDesc* boundary, int steps,
                                                      float* out_flow, int out_value_count) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    r.error.clear();
    WindTunnel* t = lookup(handle);
    return status(t && advance(handle, *t, boundary_or_default(boundary), steps, out_flow, out_value_count));
}

AERO_LBM_CAPI_EXPORT int aero_solver_advance_wind_tunnel(long long handle, const AeroBoundaryDesc* boundary,
                                                         int steps) {
    return aero_solver_step_wind_tunnel(handle, boundary, steps, nullptr, 0);
}

AERO_LBM_CAPI_EXPORT int aero_solver_run_wind_tunnel(const AeroStepInput* in, AeroStepOutput* out) {
    if (!in || !out || !out->flow_out) {
        std::lock_guard<std::mutex> lk(reg().mu);
        return status(fail("missing step input or output"));
    }
    long long h = 0;
    bool ok = aero_solver_create_with_grid(&in->grid, &h) != 0;
    const int cells = ok ? in->grid.nx * in->grid.ny * in->grid.nz : 0;
    const int values = cells * AERO_SOLVER_FLOW_CHANNELS;
    ok = ok && (!in->solid_mask || aero_solver_set_solid_mask(h, in->solid_mask, cells)) &&
         (!in->prev_flow || aero_solver_set_flow_state(h, in->prev_flow, values)) &&
         aero_solver_step_wind_tunnel(h, &in->boundary, in->steps, out->flow_out, values);
    float vmax = 0.0f;
    for (int c = 0; ok && c < cells; ++c) {
        const float* f = out->flow_out + (size_t)c * AERO_SOLVER_FLOW_CHANNELS;
        const float speed = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        if (std::isfinite(speed) && speed > vmax) vmax = speed;
    }
    out->max_velocity = vmax;
    out->status = status(ok);
    if (h) aero_solver_destroy(h);
    return status(ok);
}

AERO_LBM_CAPI_EXPORT void aero_solver_destroy(long long handle) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    if (r.tunnels.erase(handle)) aero_lbm_release_context(handle);
}

AERO_LBM_CAPI_EXPORT const char* aero_solver_last_error(void) {
    const std::string& e = reg().error;
    return e.empty() ? aero_lbm_last_error() : e.c_str();
}

AERO_LBM_CAPI_EXPORT const char* aero_solver_runtime_info(void) { return aero_lbm_runtime_info(); }

AERO_LBM_CAPI_EXPORT int aero_solver_finish(void) {
    Registry& r = reg();
    std::lock_guard<std::mutex> lk(r.mu);
    r.error.clear();
    if (aero_lbm_finish()) return status(true);
    return status(fail(std::string("aero_lbm_finish failed: ") + aero_lbm_last_error()));
}

}  // extern "C"
