// Native brick runtime: the world of cubic bricks that the client keeps around the
// player (ref: native/src/aero_lbm_simulation_bridge.cpp:128-210 for the state,
// 1004-1171 for one world step). Each active brick owns one solver context of this
// library (negative keys counting down from -1); the runtime seeds new bricks, keeps
// the static constraints, builds the 11-channel packet when a brick (re)starts,
// drives the cached in-place step and publishes quantised flow atlases.
//
// The runtime is host code that calls the library's own C ABI, the way the
// reference's bridge calls its solver; every per-cell field it hands to the device
// travels once per (re)start, not per step.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/aero_lbm_capi.h"
#include "runtime_internal.h"

namespace aero {
namespace bricks {

// ---- constants (ref: simulation_bridge.cpp:230-262, 480-523) ----
constexpr int kPacketCh = 11;
constexpr int kFlowCh = AERO_LBM_SIMULATION_FLOW_STATE_CHANNELS;
constexpr int kAtlasCh = AERO_LBM_SIMULATION_PACKED_ATLAS_CHANNELS;
constexpr int kGhostLayers = 1;
constexpr int kRefLayers = 8;
constexpr float kRefRelax = 0.10f;
constexpr uint64_t kRetentionEpochs = 16;
constexpr float kAtlasVelocityRange = 5.6f;
constexpr float kAtlasPressureRange = 0.03f;
constexpr float kFanSpeed = 4.0f;            // m/s of a client fan
constexpr int kFanLength = 5;                // cells a fan pushes along its axis
constexpr int kFanRadius = 1;                // half width of the fan's 3x3 disk
constexpr int kHeatPlumeHeight = 4;          // cells above a heat source
constexpr float kHeatCoupling = 0.85f;       // share of a source's power 1 cell up
constexpr float kSecondsPerStep = 0.05f;
constexpr float kAirDensity = 1.225f, kAirHeat = 1005.0f, kCellVolume = 1.0f, kTempScale = 20.0f;
constexpr float kSourceMax = 0.006f;
constexpr uint8_t kKindFanXNeg = 32;         // surface kinds 32..37: fans toward x-, x+, y-, y+, z-, z+
constexpr int kDeltaCellPatch = 8, kDeltaBrickUpload = 9, kDeltaSourcePatch = 10;
// face-neighbour order x-, x+, y-, y+, z-, z+
constexpr int kNbr[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
// the client's face order for open-face masks: y-, y+, z-, z+, x-, x+ (ref: 478-481)
constexpr int kJavaFace[6][3] = {{0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}, {-1, 0, 0}, {1, 0, 0}};
// fan direction codes 1..6 = x-, x+, y-, y+, z-, z+ (ref: 482-486)
constexpr int kFanDir[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

struct Coord {
    int x = 0, y = 0, z = 0;
};
inline bool operator==(const Coord& a, const Coord& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
struct CoordHash {
    size_t operator()(const Coord& c) const noexcept {
        uint64_t h = (uint64_t)(uint32_t)c.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)c.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)c.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

struct StaticData {
    int n = 0;
    std::vector<uint8_t> obstacle, surface_kind, fan_dir_src, sky, direct;
    std::vector<uint16_t> open_face;
    std::vector<float> emitter, emitter_src;
};
struct DynamicData {
    int n = 0;
    std::vector<float> flow, air, surface;
};
struct Brick {
    bool active_hint = false, active = false;
    bool geometry_dirty = false, forcing_dirty = false, pending_reinit = false;
    uint64_t last_hint_epoch = 0, last_active_epoch = 0;
    long long context = 0;
    bool compact = false;  // the context steps cached with the compact boundary
    bool stale = false;    // dyn is older than the context's state
    float cv[3] = {0.0f, 0.0f, 0.0f};
    std::shared_ptr<StaticData> stat;
    std::shared_ptr<DynamicData> dyn;
    std::shared_ptr<const DynamicData> bref;
    std::vector<float> packet;
    std::vector<int> patched_cells;
};
using BrickMap = std::unordered_map<Coord, Brick, CoordHash>;
using CoordSet = std::unordered_set<Coord, CoordHash>;

// Packed atlases of a world's in-place bricks, published together on the device at
// the first request after a world step and then handed out brick by brick.
struct AtlasCache {
    uint64_t epoch = ~0ull;
    int stride = 0;
    float vscale = 0.0f;
    std::unordered_map<Coord, std::pair<long long, size_t>, CoordHash> index;  // context key, offset in data
    std::vector<int16_t> data;
};

struct World {
    int size = 0;
    float dx = 1.0f, dt = 0.05f;
    int mode = AERO_LBM_REALTIME_SOLVER_CLASSIC_D3Q27;
    bool exact_hints = false;
    bool batching = true;  // cached in-place bricks step in shared launches
    uint64_t epoch = 0;
    uint64_t batched_steps = 0;  // world steps whose cached bricks took one batched launch set
    double host_ms_total = 0, host_ms_prepare = 0, host_ms_batch = 0;  // host time inside world steps
    uint64_t cached_steps = 0, packet_steps = 0, dynamic_upload_resets = 0;
    uint64_t patch_attempts = 0, patch_steps = 0, patch_failures = 0, patch_cells = 0;
    BrickMap bricks;
    CoordSet closure;
    std::vector<AeroLbmWorldDelta> pending;
    AtlasCache atlas_cache;
};
struct PendingUpload {
    long long world = 0;
    Coord coord;
    std::shared_ptr<StaticData> stat;
};
struct Service {
    int focus[4] = {0, 0, 0, 0};
    bool l2_ready = false;
    int l2[5] = {0, 0, 0, 0, 0};
    std::vector<AeroLbmWorldDelta> deltas;
    std::unordered_map<long long, World> worlds;
    long long next_key = -1;
    int next_upload = 1;
    std::unordered_map<int, PendingUpload> uploads;
    std::unordered_map<long long, std::vector<int16_t>> atlases;  // callers' packed flow atlases
};

std::mutex g_mu;
Service g_storage;
Service* g_service = nullptr;
long long g_service_key = 0;
std::string g_error;

void set_err(const std::string& s) { g_error = s; }
Service* lookup(long long key) { return (g_service && key == g_service_key) ? g_service : nullptr; }
size_t cube(int n) { return (size_t)n * n * n; }
int cell_of(int n, int x, int y, int z) { return (x * n + y) * n + z; }
bool cell_count(int nx, int ny, int nz, int* out) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return false;
    const long long c = (long long)nx * ny * nz;
    if (c > std::numeric_limits<int>::max()) return false;
    *out = (int)c;
    return true;
}

// ---- predicates and counts (ref: 661-830) ----
bool has_static(const World& w, const Brick& b) { return b.stat && b.stat->n == w.size; }
bool solver_active(const World& w, const Brick& b) { return b.active && has_static(w, b); }
bool dyn_valid(const World& w, const DynamicData* d) {
    if (!d || d->n != w.size) return false;
    const size_t c = cube(w.size);
    return d->flow.size() == c * kFlowCh && d->air.size() == c && d->surface.size() == c;
}
bool resident(const Brick& b) {
    return b.active_hint || b.active || b.geometry_dirty || b.forcing_dirty || b.pending_reinit;
}
bool experimental(const World& w) {
    switch (w.mode) {
        case AERO_LBM_REALTIME_SOLVER_COMPACT_EXPERIMENTAL:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_FP16_INPLACE_EXPERIMENTAL:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_BF16_INPLACE:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_S16_INPLACE:
        case AERO_LBM_REALTIME_SOLVER_D3Q27_FP32_INPLACE:
            return true;
        case AERO_LBM_REALTIME_SOLVER_AUTO: {
            // ref: 362-384: AERO_LBM_BRICK_WORLD_COMPACT, else AERO_LBM_SIMULATION_COMPACT
            auto flag = [](const char* v) {
                return std::strcmp(v, "0") && std::strcmp(v, "false") && std::strcmp(v, "FALSE") &&
                       std::strcmp(v, "off") && std::strcmp(v, "OFF");
            };
            if (const char* v = std::getenv("AERO_LBM_BRICK_WORLD_COMPACT")) return flag(v);
            if (const char* v = std::getenv("AERO_LBM_SIMULATION_COMPACT")) return flag(v);
            return false;
        }
        default:
            return false;
    }
}
bool compact_ready(const World& w, const Brick& b) {
    return experimental(w) && b.compact && b.context != 0 && !b.forcing_dirty && !b.geometry_dirty &&
           !b.pending_reinit && dyn_valid(w, b.dyn.get());
}

void release_context(Brick& b) {
    if (b.context != 0) {
        aero_lbm_release_context(b.context);
        b.context = 0;
    }
    b.compact = false;
    b.stale = false;
    b.patched_cells.clear();
}

// ref: 897-914
void recompute_active(World& w) {
    for (auto& e : w.bricks) {
        const bool hinted = w.closure.count(e.first) != 0;
        Brick& b = e.second;
        b.active = w.exact_hints ? (hinted && has_static(w, b))
                                 : ((hinted && has_static(w, b)) || b.geometry_dirty || b.forcing_dirty ||
                                    b.pending_reinit);
        if (b.active) b.last_active_epoch = w.epoch;
    }
}

// ref: 926-950
void prune(World& w) {
    for (auto it = w.bricks.begin(); it != w.bricks.end();) {
        Brick& b = it->second;
        if (w.exact_hints && !w.closure.count(it->first)) {
            release_context(b);
            it = w.bricks.erase(it);
            continue;
        }
        const bool stale = !b.active && !b.geometry_dirty && !b.forcing_dirty && !b.pending_reinit &&
                           w.epoch > b.last_active_epoch && (w.epoch - b.last_active_epoch) > kRetentionEpochs;
        if (!stale) {
            ++it;
            continue;
        }
        release_context(b);
        w.closure.erase(it->first);
        it = w.bricks.erase(it);
    }
}

// ref: 1865-1883
void clear_unbacked(World& w) {
    for (auto& e : w.bricks) {
        Brick& b = e.second;
        if (b.active_hint || has_static(w, b)) continue;
        release_context(b);
        b.geometry_dirty = b.forcing_dirty = b.pending_reinit = b.active = false;
        b.dyn.reset();
        b.bref.reset();
        b.packet.clear();
    }
}

void ensure_dynamic(const World& w, Brick& b) {
    if (dyn_valid(w, b.dyn.get())) return;
    auto d = std::make_shared<DynamicData>();
    d->n = w.size;
    d->flow.assign(cube(w.size) * kFlowCh, 0.0f);
    d->air.assign(cube(w.size), 0.0f);
    d->surface.assign(cube(w.size), 0.0f);
    b.dyn = std::move(d);
}

// solid cells carry no flow or temperature; fluid cells without an open face,
// surface or emitter carry no surface temperature (ref: 1204-1240)
void apply_constraints(const World& w, const StaticData& s, DynamicData& d) {
    if (!dyn_valid(w, &d) || s.n != w.size) return;
    const size_t cells = d.air.size();
    for (size_t c = 0; c < cells; ++c) {
        if (s.obstacle[c]) {
            std::fill_n(d.flow.begin() + (ptrdiff_t)(c * kFlowCh), kFlowCh, 0.0f);
            d.air[c] = 0.0f;
            d.surface[c] = 0.0f;
        } else if (s.open_face[c] == 0 && s.surface_kind[c] == 0 && !(s.emitter[c] > 0.0f)) {
            d.surface[c] = 0.0f;
        }
    }
}
void apply_constraints(const World& w, Brick& b) {
    if (b.stat && b.dyn) apply_constraints(w, *b.stat, *b.dyn);
}

bool nonzero_flow(const DynamicData& d) {
    for (size_t i = 0; i + 2 < d.flow.size(); i += kFlowCh)
        if (std::fabs(d.flow[i]) > 1e-8f || std::fabs(d.flow[i + 1]) > 1e-8f || std::fabs(d.flow[i + 2]) > 1e-8f)
            return true;
    return false;
}

struct Mean {
    float flow[kFlowCh] = {0, 0, 0, 0};
    float air = 0.0f, surface = 0.0f;
};
// plain FP32 running sums in cell order (ref: 1264-1285)
Mean brick_mean(const DynamicData& d) {
    Mean m;
    const size_t cells = d.air.size();
    for (size_t c = 0; c < cells; ++c) {
        for (int k = 0; k < kFlowCh; ++k) m.flow[k] += d.flow[c * kFlowCh + k];
        m.air += d.air[c];
        m.surface += d.surface[c];
    }
    const float inv = 1.0f / (float)cells;
    for (float& v : m.flow) v *= inv;
    m.air *= inv;
    m.surface *= inv;
    return m;
}

using Snapshots = std::unordered_map<Coord, std::shared_ptr<const DynamicData>, CoordHash>;

// a new brick starts from the mean of its face neighbours that carry flow (ref: 1287-1334)
bool seed_from_neighbours(const World& w, const Coord& c, const Snapshots& snaps, DynamicData& d) {
    Mean acc;
    int count = 0;
    for (const auto& o : kNbr) {
        auto it = snaps.find(Coord{c.x + o[0], c.y + o[1], c.z + o[2]});
        if (it == snaps.end() || !dyn_valid(w, it->second.get()) || !nonzero_flow(*it->second)) continue;
        const Mean m = brick_mean(*it->second);
        for (int k = 0; k < kFlowCh; ++k) acc.flow[k] += m.flow[k];
        acc.air += m.air;
        acc.surface += m.surface;
        ++count;
    }
    if (count <= 0) return false;
    const float inv = 1.0f / (float)count;
    for (float& v : acc.flow) v *= inv;
    acc.air *= inv;
    acc.surface *= inv;
    const size_t cells = d.air.size();
    for (size_t i = 0; i < cells; ++i) {
        for (int k = 0; k < kFlowCh; ++k) d.flow[i * kFlowCh + k] = acc.flow[k];
        d.air[i] = acc.air;
        d.surface[i] = acc.surface;
    }
    return true;
}

// ... or from its boundary reference when that carries flow (ref: 1336-1352)
bool seed_from_reference(const World& w, const Brick& b, DynamicData& d) {
    const DynamicData* r = b.bref.get();
    if (!dyn_valid(w, r) || !dyn_valid(w, &d) || !nonzero_flow(*r)) return false;
    d.flow = r->flow;
    d.air = r->air;
    d.surface = r->surface;
    return true;
}

// the outer layer of the face toward a neighbour takes the neighbour's adjacent
// layer (ref: 1354-1409)
void copy_ghost(const World& w, const DynamicData& nb, int axis, bool positive, DynamicData& d) {
    if (!dyn_valid(w, &d) || !dyn_valid(w, &nb) || !nonzero_flow(nb)) return;
    const int n = w.size;
    const int layers = std::max(1, std::min(kGhostLayers, n));
    for (int l = 0; l < layers; ++l) {
        const int dst = positive ? n - 1 - l : l;
        const int src = positive ? l : n - 1 - l;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) {
                int dc, sc;
                if (axis == 0) {
                    dc = cell_of(n, dst, a, b);
                    sc = cell_of(n, src, a, b);
                } else if (axis == 1) {
                    dc = cell_of(n, a, dst, b);
                    sc = cell_of(n, a, src, b);
                } else {
                    dc = cell_of(n, a, b, dst);
                    sc = cell_of(n, a, b, src);
                }
                std::copy_n(nb.flow.begin() + (ptrdiff_t)sc * kFlowCh, kFlowCh, d.flow.begin() + (ptrdiff_t)dc * kFlowCh);
                d.air[dc] = nb.air[sc];
                d.surface[dc] = nb.surface[sc];
            }
    }
}

// ref: 1441-1468
bool face_exposed(const World& w, const Coord& c, int axis, bool positive) {
    Coord n = c;
    (axis == 0 ? n.x : axis == 1 ? n.y : n.z) += positive ? 1 : -1;
    auto it = w.bricks.find(n);
    return it == w.bricks.end() || !solver_active(w, it->second);
}
int exposed_mask(const World& w, const Coord& c) {
    int m = 0;
    for (int f = 0; f < 6; ++f)
        if (face_exposed(w, c, f >> 1, f & 1)) m |= 1 << f;
    return m;
}

// ref: 1470-1495
void update_shell(const World& w, const Coord& c, const Brick& b) {
    if (b.context == 0) return;
    const int mask = exposed_mask(w, c);
    if (mask == 0 || !dyn_valid(w, b.bref.get())) {
        aero_lbm_clear_d3q27_f16_boundary_shell_rect(b.context);
        return;
    }
    (void)aero_lbm_set_d3q27_f16_boundary_shell_rect(w.size, w.size, w.size, b.context, b.bref->flow.data(), kFlowCh,
                                                      mask, kRefLayers, kRefRelax);
}

// fan velocity of a client fan surface kind (ref: 524-537)
bool fan_velocity(uint8_t kind, float v[3]) {
    v[0] = v[1] = v[2] = 0.0f;
    if (kind < kKindFanXNeg || kind > kKindFanXNeg + 5) return false;
    const int d = kind - kKindFanXNeg;  // 0..5 = x-, x+, y-, y+, z-, z+
    v[d >> 1] = (d & 1) ? kFanSpeed : -kFanSpeed;
    return true;
}

// emitter power (W) -> thermal source per step, clamped (ref: 3152-3158)
float power_to_source(float watts) {
    const float s = watts * kSecondsPerStep / (kAirDensity * kAirHeat * kCellVolume * kTempScale);
    return std::min(std::max(s, -kSourceMax), kSourceMax);
}

// the 11-channel packet of a brick: obstacle, fan mask + velocity, reference state
// (the boundary reference within 8 layers of exposed faces, else the brick's own
// state), emitter source (ref: 1497-1596)
bool build_packet(const World& w, const Coord& c, const Brick& b, const DynamicData& d, std::vector<float>& pk) {
    if (!b.stat || !dyn_valid(w, &d) || b.stat->n != w.size) {
        set_err("brick_step_packet: missing or invalid brick state at (" + std::to_string(c.x) + "," +
                std::to_string(c.y) + "," + std::to_string(c.z) + ")");
        return false;
    }
    const StaticData& s = *b.stat;
    const int n = w.size;
    pk.assign(cube(n) * kPacketCh, 0.0f);
    const bool ref_ok = dyn_valid(w, b.bref.get());
    bool exp[6];
    for (int f = 0; f < 6; ++f) exp[f] = ref_ok && face_exposed(w, c, f >> 1, f & 1);
    const int layers = std::max(1, std::min(kRefLayers, n));
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
            for (int z = 0; z < n; ++z) {
                const size_t cell = (size_t)cell_of(n, x, y, z);
                float* p = pk.data() + cell * kPacketCh;
                if (s.obstacle[cell]) {
                    p[0] = 1.0f;
                    continue;
                }
                float fv[3];
                if (fan_velocity(s.surface_kind[cell], fv)) {
                    p[1] = 1.0f;
                    p[2] = fv[0];
                    p[3] = fv[1];
                    p[4] = fv[2];
                }
                const bool use_ref = ref_ok && ((exp[0] && x < layers) || (exp[1] && x >= n - layers) ||
                                                (exp[2] && y < layers) || (exp[3] && y >= n - layers) ||
                                                (exp[4] && z < layers) || (exp[5] && z >= n - layers));
                const DynamicData& src = use_ref ? *b.bref : d;
                for (int k = 0; k < 4; ++k) p[5 + k] = src.flow[cell * kFlowCh + k];
                p[10] = src.air[cell];
                if (s.emitter[cell] > 0.0f) p[9] = power_to_source(s.emitter[cell]);
            }
    return true;
}

struct Summary {
    bool supported = false;
    float v[3] = {0, 0, 0};
    int fluid = 0, fans = 0;
};
// mean reference velocity over the finite fluid cells, in FP64 (ref: 317-358)
Summary summarize(const std::vector<float>& pk) {
    Summary s;
    if (pk.empty() || pk.size() % kPacketCh) return s;
    double sum[3] = {0, 0, 0};
    for (size_t i = 0; i < pk.size(); i += kPacketCh) {
        if (pk[i] > 0.5f) continue;
        if (pk[i + 1] > 0.5f) s.fans++;
        const float vx = pk[i + 5], vy = pk[i + 6], vz = pk[i + 7];
        if (!std::isfinite(vx) || !std::isfinite(vy) || !std::isfinite(vz)) continue;
        sum[0] += vx;
        sum[1] += vy;
        sum[2] += vz;
        s.fluid++;
    }
    if (s.fluid > 0)
        for (int k = 0; k < 3; ++k) s.v[k] = (float)(sum[k] / s.fluid);
    s.supported = s.fluid > 0 && s.fans == 0;
    return s;
}

// velocity inlet x-, convective outflow x+, symmetry elsewhere (ref: 452-475)
bool configure_compact(const float v[3], int n) {
    AeroLbmBenchmarkConfig cfg{};
    aero_lbm_benchmark_default_config(&cfg);
    cfg.enabled = 1;
    cfg.preset = AERO_LBM_BENCHMARK_PRESET_NONE;
    cfg.reference_length = (float)std::max(1, n);
    for (int k = 0; k < 3; ++k) cfg.initial_velocity[k] = v[k];
    AeroLbmBoundaryFaceConfig* faces[6] = {&cfg.x_min, &cfg.x_max, &cfg.y_min, &cfg.y_max, &cfg.z_min, &cfg.z_max};
    for (int f = 0; f < 6; ++f) {
        *faces[f] = AeroLbmBoundaryFaceConfig{};
        faces[f]->hydrodynamic_kind = AERO_LBM_HYDRO_BOUNDARY_SYMMETRY;
        faces[f]->thermal_kind = AERO_LBM_THERMAL_BOUNDARY_DISABLED;
    }
    cfg.x_min.hydrodynamic_kind = AERO_LBM_HYDRO_BOUNDARY_VELOCITY_DIRICHLET;
    for (int k = 0; k < 3; ++k) cfg.x_min.velocity[k] = v[k];
    cfg.x_max.hydrodynamic_kind = AERO_LBM_HYDRO_BOUNDARY_CONVECTIVE_OUTFLOW;
    return aero_lbm_benchmark_set_config(&cfg) != 0;
}

struct SolverModeScope {
    explicit SolverModeScope(int mode) { aero_lbm_set_realtime_solver_mode(mode); }
    ~SolverModeScope() { aero_lbm_set_realtime_solver_mode(AERO_LBM_REALTIME_SOLVER_AUTO); }
};

bool sync_dynamic(const World& w, Brick& b) {
    if (!b.dyn || b.context == 0 || !dyn_valid(w, b.dyn.get())) return false;
    const int n = w.size;
    if (!aero_lbm_get_flow_state_rect(n, n, n, b.context, b.dyn->flow.data())) return false;
    if (!aero_lbm_get_temperature_state_rect(n, n, n, b.context, b.dyn->air.data())) return false;
    b.stale = false;
    return true;
}

// fan direction code (1-6 = x-, x+, y-, y+, z-, z+) of a client fan surface kind
// (ref: simulation_bridge.cpp:539-549)
uint8_t fan_code_of_kind(uint8_t kind) {
    return (kind >= kKindFanXNeg && kind <= kKindFanXNeg + 5) ? (uint8_t)(kind - kKindFanXNeg + 1) : 0;
}

// the cells patched by world deltas, applied to the cached in-place context in place
// of a packet rebuild (ref: simulation_bridge.cpp:1617-1676)
bool apply_static_patches(World& w, Brick& b) {
    if (b.patched_cells.empty() || b.context == 0 || !has_static(w, b)) return false;
    w.patch_attempts++;
    const StaticData& s = *b.stat;
    const int n = w.size;
    std::vector<int> cells;
    std::vector<uint8_t> solid, fan;
    std::unordered_set<int> seen;
    for (int cell : b.patched_cells) {
        if (cell < 0 || (size_t)cell >= cube(n) || !seen.insert(cell).second) continue;
        const uint8_t so = s.obstacle[(size_t)cell] ? 1 : 0;
        cells.push_back(cell);
        solid.push_back(so);
        fan.push_back(so ? 0 : fan_code_of_kind(s.surface_kind[(size_t)cell]));
    }
    if (!cells.empty() && !aero_lbm_patch_d3q27_f16_static_cells_rect(n, n, n, b.context, (int)cells.size(),
                                                                      cells.data(), solid.data(), fan.data())) {
        w.patch_failures++;
        return false;
    }
    w.patch_steps += cells.empty() ? 0 : 1;
    w.patch_cells += cells.size();
    b.patched_cells.clear();
    b.forcing_dirty = false;
    b.geometry_dirty = false;
    b.pending_reinit = false;
    b.stale = true;
    return true;
}

// ref: 1678-1704: a cached in-place context takes pending cell patches in place and
// steps on; anything else re-enters through the packet path
bool step_cached(World& w, Brick& b) {
    if (!experimental(w) || !b.compact || b.context == 0 || b.geometry_dirty || b.pending_reinit) return false;
    if (b.forcing_dirty && !apply_static_patches(w, b)) return false;
    const int n = w.size;
    bool ok = configure_compact(b.cv, n) && aero_lbm_step_rect_cached(n, n, n, b.context, nullptr) != 0;
    aero_lbm_benchmark_reset_config();
    if (ok) {
        b.stale = true;
        w.cached_steps++;
    }
    return ok;
}

// The host half of step_cached for a brick whose step joins the world step's
// batched launch: parameters resolved under the brick's compact configuration,
// launch deferred. false: the brick steps on its own (step_actual).
bool defer_cached(World& w, Brick& b, std::vector<Context*>& batch) {
    if (!experimental(w) || !b.compact || b.context == 0 || b.geometry_dirty || b.pending_reinit || b.forcing_dirty)
        return false;
    if (!context_batchable(b.context)) return false;
    SolverModeScope scope(w.mode);
    const int n = w.size;
    Context* c = configure_compact(b.cv, n) ? prepare_cached_step(b.context, n, n, n) : nullptr;
    aero_lbm_benchmark_reset_config();
    if (!c) return false;
    batch.push_back(c);
    b.stale = true;
    w.cached_steps++;
    return true;
}

long long next_context_key(Service& s) { return s.next_key--; }

// one brick step: cached when the compact context is ready, else a packet step
// that (re)starts or refreshes the context (ref: 1706-1832)
bool step_actual(Service& svc, World& w, const Coord& c, Brick& b, DynamicData& d) {
    SolverModeScope scope(w.mode);
    const int n = w.size;
    if (step_cached(w, b)) return true;
    std::shared_ptr<DynamicData> zero_fallback;
    if (b.forcing_dirty && nonzero_flow(d)) zero_fallback = std::make_shared<DynamicData>(d);
    w.packet_steps++;
    if (!build_packet(w, c, b, d, b.packet)) return false;
    const bool had_context = b.context != 0;
    if (b.context == 0) b.context = next_context_key(svc);
    bool ok = false, compact_ok = false;
    if (experimental(w)) {
        const Summary s = summarize(b.packet);
        const bool configured = configure_compact(s.v, n);
        if (s.supported && configured && had_context && b.compact)
            ok = aero_lbm_step_rect_cached(n, n, n, b.context, nullptr) != 0;
        if (!ok && configured) ok = aero_lbm_step_rect(b.packet.data(), n, n, n, b.context, nullptr) != 0;
        aero_lbm_benchmark_reset_config();
        if (ok && aero_lbm_context_realtime_cached_initialized(b.context)) {
            compact_ok = true;
            b.compact = true;
            for (int k = 0; k < 3; ++k) b.cv[k] = s.v[k];
            b.stale = true;
        }
    } else {
        ok = aero_lbm_step_rect(b.packet.data(), n, n, n, b.context, nullptr) != 0;
        b.compact = false;
    }
    if (!ok) {
        set_err("brick_step_actual: aero_lbm_step_rect failed at (" + std::to_string(c.x) + "," +
                std::to_string(c.y) + "," + std::to_string(c.z) + "): " + aero_lbm_last_error());
        release_context(b);
        return false;
    }
    if (compact_ok) {
        update_shell(w, c, b);
        b.patched_cells.clear();
        return true;
    }
    if (!sync_dynamic(w, b)) {
        set_err("brick_step_actual: sync from context failed at (" + std::to_string(c.x) + "," +
                std::to_string(c.y) + "," + std::to_string(c.z) + "): " + aero_lbm_last_error());
        release_context(b);
        return false;
    }
    if (zero_fallback && !nonzero_flow(*b.dyn)) {
        b.dyn = std::move(zero_fallback);
        apply_constraints(w, b);
        release_context(b);
        return true;
    }
    apply_constraints(w, b);
    update_shell(w, c, b);
    b.patched_cells.clear();
    return true;
}

bool ensure_l2(Service& s, int n) {
    const int want[5] = {n, n, n, kPacketCh, kFlowCh};
    if (s.l2_ready && std::equal(want, want + 5, s.l2)) return true;
    if (!aero_lbm_init_rect(n, n, n, kPacketCh, kFlowCh)) {
        set_err(std::string("simulation_l2_runtime_init failed: ") + aero_lbm_last_error());
        return false;
    }
    std::copy(want, want + 5, s.l2);
    s.l2_ready = true;
    return true;
}

// ---- static regions from client sources (ref: 1897-2097) ----
bool local_solid(const StaticData& s, int x, int y, int z) {
    const int n = s.n;
    if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) return true;
    return s.obstacle[(size_t)cell_of(n, x, y, z)] != 0;
}

// open-face masks from the obstacle field, fan disks from the fan sources and heat
// plumes from the emitter sources
void derive_from_sources(StaticData& s) {
    const int n = s.n;
    const size_t cells = cube(n);
    s.surface_kind.assign(cells, 0);
    s.emitter.assign(cells, 0.0f);
    s.open_face.assign(cells, 0);
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
            for (int z = 0; z < n; ++z) {
                if (local_solid(s, x, y, z)) continue;
                uint16_t m = 0;
                for (int f = 0; f < 6; ++f)
                    if (!local_solid(s, x + kJavaFace[f][0], y + kJavaFace[f][1], z + kJavaFace[f][2]))
                        m = (uint16_t)(m | (1u << f));
                s.open_face[(size_t)cell_of(n, x, y, z)] = m;
            }
    // each fan pushes along its axis for kFanLength cells over a 3x3 disk; a disk cell
    // counts when it is fluid and the lane from the fan to it is clear
    for (int fx = 0; fx < n; ++fx)
        for (int fy = 0; fy < n; ++fy)
            for (int fz = 0; fz < n; ++fz) {
                const int dir = std::min<int>(s.fan_dir_src[(size_t)cell_of(n, fx, fy, fz)], 6);
                if (dir == 0) continue;
                const int* d = kFanDir[dir];
                const int axis = dir <= 2 ? 0 : (dir <= 4 ? 1 : 2);
                for (int dist = 1; dist <= kFanLength; ++dist)
                    for (int a = -kFanRadius; a <= kFanRadius; ++a)
                        for (int b = -kFanRadius; b <= kFanRadius; ++b) {
                            int lx = fx, ly = fy, lz = fz;  // lane start beside the fan
                            if (axis == 0) {
                                ly += a;
                                lz += b;
                            } else if (axis == 1) {
                                lx += a;
                                lz += b;
                            } else {
                                lx += a;
                                ly += b;
                            }
                            const int tx = lx + d[0] * dist, ty = ly + d[1] * dist, tz = lz + d[2] * dist;
                            if (local_solid(s, tx, ty, tz)) continue;
                            bool clear = true;
                            for (int k = 1; k < dist && clear; ++k)
                                clear = !local_solid(s, lx + d[0] * k, ly + d[1] * k, lz + d[2] * k);
                            if (!clear) continue;
                            s.surface_kind[(size_t)cell_of(n, tx, ty, tz)] = (uint8_t)(kKindFanXNeg + dir - 1);
                        }
            }
    // heat rises: a source heats its own cell and up to kHeatPlumeHeight cells above
    // with power * 0.85 / d^2 while the column stays clear
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
            for (int z = 0; z < n; ++z) {
                const float p = s.emitter_src[(size_t)cell_of(n, x, y, z)];
                if (!(p > 0.0f) || !std::isfinite(p)) continue;
                for (int dist = 0; dist <= kHeatPlumeHeight; ++dist) {
                    const int ty = y + dist;
                    if (ty >= n || local_solid(s, x, ty, z)) continue;
                    bool clear = true;
                    for (int k = y + 1; k < ty && clear; ++k) clear = !local_solid(s, x, k, z);
                    if (!clear) continue;
                    s.emitter[(size_t)cell_of(n, x, ty, z)] +=
                        dist == 0 ? p : p * kHeatCoupling / (float)(dist * dist);
                }
            }
}

// ref: 2660-2685
void install_static(World& w, Brick& b, std::shared_ptr<StaticData> s) {
    const bool live = b.context != 0 || dyn_valid(w, b.dyn.get());
    b.stat = std::move(s);
    b.patched_cells.clear();
    if (live) {
        b.forcing_dirty = true;
        b.geometry_dirty = false;
        b.pending_reinit = false;
        apply_constraints(w, b);
    } else {
        b.geometry_dirty = true;
        b.pending_reinit = true;
    }
    b.active = true;
    b.last_active_epoch = w.epoch;
}

int floor_div(int v, int d) {
    const int q = v / d, r = v % d;
    return (r != 0 && v < 0) ? q - 1 : q;
}

void mark(World& w, const Coord& c, bool geometry, bool forcing, bool reinit) {
    Brick& b = w.bricks[c];
    b.geometry_dirty |= geometry;
    b.forcing_dirty |= forcing;
    b.pending_reinit |= reinit;
    b.active = true;
    b.last_active_epoch = w.epoch;
}

// ref: 2687-2771 (delta 10, client source patches, is not applied by this build)
void apply_delta(World& w, const AeroLbmWorldDelta& d) {
    const Coord c{floor_div(d.x, w.size), floor_div(d.y, w.size), floor_div(d.z, w.size)};
    switch (d.type) {
        case kDeltaCellPatch: {
            auto it = w.bricks.find(c);
            if (it == w.bricks.end() || !has_static(w, it->second)) break;
            Brick& b = it->second;
            const int lx = d.x - c.x * w.size, ly = d.y - c.y * w.size, lz = d.z - c.z * w.size;
            const size_t cell = (size_t)cell_of(w.size, lx, ly, lz);
            if (b.stat.use_count() > 1) b.stat = std::make_shared<StaticData>(*b.stat);
            StaticData& s = *b.stat;
            const uint8_t ob = (d.data1 & 1) ? 1 : 0;
            const uint8_t kind = (uint8_t)((d.data1 >> 8) & 0xFF);
            const uint16_t open = (uint16_t)(d.data2 & 0xFF);
            const float power = std::isfinite(d.value0) ? d.value0 : 0.0f;
            const bool changed = s.obstacle[cell] != ob || s.surface_kind[cell] != kind ||
                                 s.open_face[cell] != open || s.emitter[cell] != power;
            s.obstacle[cell] = ob;
            s.surface_kind[cell] = kind;
            s.open_face[cell] = open;
            s.emitter[cell] = power;
            std::fill_n(s.sky.begin() + (ptrdiff_t)cell * 6, 6, 0);
            std::fill_n(s.direct.begin() + (ptrdiff_t)cell * 6, 6, 0);
            apply_constraints(w, b);
            b.forcing_dirty |= changed;
            if (changed) b.patched_cells.push_back((int)cell);
            b.active = true;
            b.last_active_epoch = w.epoch;
            break;
        }
        case 1:
            mark(w, c, true, false, false);
            break;
        case 2:
        case 3:
            mark(w, c, true, false, true);
            for (const auto& o : kNbr) {
                const Coord nb{c.x + o[0], c.y + o[1], c.z + o[2]};
                if (w.bricks.count(nb)) mark(w, nb, false, false, true);
            }
            break;
        case 4:
        case 5:
            mark(w, c, false, true, false);
            break;
        default:
            break;
    }
}

// ref: 2773-2809
void apply_pending(Service& svc, long long key, World& w) {
    if (w.pending.empty()) return;
    std::vector<AeroLbmWorldDelta> pending;
    pending.swap(w.pending);
    bool touched = false;
    for (const AeroLbmWorldDelta& d : pending) {
        if (d.type == kDeltaBrickUpload) {
            auto it = svc.uploads.find(d.data1);
            if (it == svc.uploads.end()) continue;
            if (it->second.world == key && it->second.stat) {
                install_static(w, w.bricks[it->second.coord], it->second.stat);
                touched = true;
            }
            svc.uploads.erase(it);
            continue;
        }
        apply_delta(w, d);
        touched = true;
    }
    if (touched) {
        recompute_active(w);
        prune(w);
    }
}

// ref: 1411-1439 (the result is not checked, as there)
bool exchange_face_halo(const World& w, const Brick& a, const Brick& b, int ox, int oy, int oz) {
    if (w.size <= 2 * kGhostLayers || a.context == 0 || b.context == 0 || !dyn_valid(w, a.dyn.get()) ||
        !dyn_valid(w, b.dyn.get()) || !nonzero_flow(*a.dyn) || !nonzero_flow(*b.dyn))
        return false;
    return aero_lbm_exchange_halo_layers_rect(w.size, w.size, w.size, kGhostLayers, a.context, b.context, ox, oy, oz) != 0;
}

// one or more world steps (ref: 1004-1171)
bool step_world(Service& svc, long long key, World& w, int steps) {
    if (steps <= 0) return false;
    using clk = std::chrono::steady_clock;
    const auto t_world = clk::now();
    auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
    if (!ensure_l2(svc, w.size)) return false;
    bool all_ok = true;
    apply_pending(svc, key, w);
    clear_unbacked(w);
    for (int s = 0; s < steps; ++s) {
        w.epoch++;
        recompute_active(w);
        CoordSet needs_seed;
        for (auto& e : w.bricks) {
            Brick& b = e.second;
            if (!solver_active(w, b)) continue;
            const bool fast = compact_ready(w, b);
            if (b.context != 0 && b.forcing_dirty && dyn_valid(w, b.dyn.get())) sync_dynamic(w, b);
            const bool had_dyn = dyn_valid(w, b.dyn.get());
            const bool had_flow = fast || (had_dyn && nonzero_flow(*b.dyn));
            if ((b.pending_reinit || b.geometry_dirty || !had_flow) && b.context != 0) release_context(b);
            ensure_dynamic(w, b);
            if (!fast && b.stat && b.dyn) apply_constraints(w, b);
            if (!had_flow) needs_seed.insert(e.first);
        }
        Snapshots snaps;
        for (const auto& e : w.bricks)
            if (solver_active(w, e.second) && dyn_valid(w, e.second.dyn.get())) snaps.emplace(e.first, e.second.dyn);
        // the population halo between face-adjacent contexts (ref: 1056-1089): the
        // classic mode's exact FP32 contexts exchange one ghost layer with each
        // positive-face neighbour; the in-place contexts keep their populations in the
        // A-A buffer, which the reference's exchange (jni.cpp:7817-7818) does not
        // touch, so in-place worlds skip the calls
        if (!experimental(w))
            for (auto& e : w.bricks) {
                Brick& b = e.second;
                if (!solver_active(w, b) || b.context == 0) continue;
                for (int f = 1; f < 6; f += 2) {
                    auto it = w.bricks.find(Coord{e.first.x + kNbr[f][0], e.first.y + kNbr[f][1], e.first.z + kNbr[f][2]});
                    if (it == w.bricks.end() || !solver_active(w, it->second) || it->second.context == 0) continue;
                    exchange_face_halo(w, b, it->second, kNbr[f][0], kNbr[f][1], kNbr[f][2]);
                }
            }
        // Cached in-place bricks are independent within a world step, so their steps
        // are collected and launched together after the loop.
        std::vector<Context*> batch;
        for (auto& e : w.bricks) {
            Brick& b = e.second;
            if (!solver_active(w, b) || !b.dyn) continue;
            if (compact_ready(w, b) && !needs_seed.count(e.first)) {
                if (w.batching) {
                    const auto t0 = clk::now();
                    const bool deferred = defer_cached(w, b, batch);
                    w.host_ms_prepare += ms_since(t0);
                    if (deferred) continue;
                }
                if (!step_actual(svc, w, e.first, b, *b.dyn)) all_ok = false;
                continue;
            }
            ensure_dynamic(w, b);
            auto next = std::make_shared<DynamicData>(*b.dyn);
            const bool seed = needs_seed.count(e.first) != 0;
            const bool bootstrap = seed || b.context == 0;
            bool seeded = !seed;
            if (seed) seeded = seed_from_reference(w, b, *next) || seed_from_neighbours(w, e.first, snaps, *next);
            if (bootstrap) {
                for (int f = 0; f < 6; ++f) {
                    auto it = snaps.find(Coord{e.first.x + kNbr[f][0], e.first.y + kNbr[f][1], e.first.z + kNbr[f][2]});
                    if (it == snaps.end() || !it->second || !nonzero_flow(*it->second)) continue;
                    copy_ghost(w, *it->second, f >> 1, f & 1, *next);
                }
            }
            if (seed && !seeded && !nonzero_flow(*next)) continue;
            Brick trial = b;
            trial.dyn = next;
            apply_constraints(w, trial);
            if (!step_actual(svc, w, e.first, trial, *trial.dyn)) {
                all_ok = false;
                b.dyn = std::move(trial.dyn);
            } else {
                b.context = trial.context;
                b.compact = trial.compact;
                b.stale = trial.stale;
                std::copy(trial.cv, trial.cv + 3, b.cv);
                b.packet = std::move(trial.packet);
                b.patched_cells = std::move(trial.patched_cells);
                b.dyn = std::move(trial.dyn);
            }
            b.pending_reinit = b.geometry_dirty = b.forcing_dirty = false;
        }
        if (!batch.empty()) {
            const auto t0 = clk::now();
            const bool stepped = batch_step_contexts(batch);
            w.host_ms_batch += ms_since(t0);
            if (stepped) {
                w.batched_steps++;
            } else {
                set_err(std::string("step_brick_world_runtime: batched step failed: ") + aero_lbm_last_error());
                all_ok = false;
            }
        }
        recompute_active(w);
        prune(w);
    }
    w.host_ms_total += ms_since(t_world);
    return all_ok;
}

int16_t quantize(float v, float range) {
    if (!(range > 0.0f) || !std::isfinite(v)) return 0;
    const float t = std::min(std::max(v / range, -1.0f), 1.0f);
    return (int16_t)std::lround(t * 32767.0f);
}

World* find_world(Service* s, long long key, const char* who) {
    if (!s) {
        set_err(std::string(who) + ": missing service");
        return nullptr;
    }
    auto it = s->worlds.find(key);
    if (it == s->worlds.end()) {
        set_err(std::string(who) + ": missing brick runtime");
        return nullptr;
    }
    return &it->second;
}

int set_hints(long long skey, long long wkey, int size, const int* coords, int count, bool closure,
              const char* who) {
    Service* s = lookup(skey);
    if (!s) {
        set_err(std::string(who) + ": missing service");
        return 0;
    }
    if (count < 0 || (count > 0 && !coords)) {
        set_err(std::string(who) + ": invalid hint payload");
        return 0;
    }
    World* w = find_world(s, wkey, who);
    if (!w) return 0;
    if (w->size != size) {
        set_err(std::string(who) + ": brick size mismatch");
        return 0;
    }
    w->exact_hints = !closure;
    for (auto& e : w->bricks) e.second.active_hint = false;
    w->closure.clear();
    w->epoch++;
    for (int i = 0; i < count; ++i) {
        const Coord c{coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
        Brick& b = w->bricks[c];
        b.active_hint = true;
        b.last_hint_epoch = b.last_active_epoch = w->epoch;
        w->closure.insert(c);
        if (closure)
            for (const auto& o : kNbr) w->closure.insert(Coord{c.x + o[0], c.y + o[1], c.z + o[2]});
    }
    for (const Coord& c : w->closure) w->bricks.try_emplace(c);
    recompute_active(*w);
    prune(*w);
    return 1;
}

std::shared_ptr<DynamicData> make_dynamic(int n, const float* flow, const float* air, const float* surface) {
    auto d = std::make_shared<DynamicData>();
    d->n = n;
    const size_t cells = cube(n);
    d->flow.assign(flow, flow + cells * kFlowCh);
    d->air.assign(air, air + cells);
    d->surface.assign(surface, surface + cells);
    return d;
}

}  // namespace bricks
}  // namespace aero

using namespace aero::bricks;

// ---------------------------------------------------------------- C ABI
// (ref: native/include/aero_lbm_capi.h:303-696, native/src/aero_lbm_simulation_bridge.cpp:4562-5535)

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_create_service(long long* out_key) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!out_key) {
        set_err("simulation_create_service: out_service_key is null");
        return 0;
    }
    if (!g_service) {
        g_storage = Service{};
        g_service = &g_storage;
        g_service_key = 1;
    }
    *out_key = g_service_key;
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT void aero_lbm_simulation_release_service(long long key) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_service && key == g_service_key) {
        for (auto& w : g_service->worlds)
            for (auto& b : w.second.bricks) release_context(b.second);
        g_service = nullptr;
        g_service_key = 0;
    }
    if (!g_service) aero_lbm_shutdown();
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_focus(long long key, int bx, int by, int bz, int radius) {
    std::lock_guard<std::mutex> lk(g_mu);
    Service* s = lookup(key);
    if (!s) {
        set_err("simulation_set_focus: missing service");
        return 0;
    }
    s->focus[0] = bx;
    s->focus[1] = by;
    s->focus[2] = bz;
    s->focus[3] = radius;
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_submit_world_deltas(long long key,
                                                                            const AeroLbmWorldDelta* deltas,
                                                                            int count) {
    std::lock_guard<std::mutex> lk(g_mu);
    Service* s = lookup(key);
    if (!s) {
        set_err("simulation_submit_world_deltas: missing service");
        return 0;
    }
    if (count < 0) {
        set_err("simulation_submit_world_deltas: negative delta_count");
        return 0;
    }
    if (count > 0 && !deltas) {
        set_err("simulation_submit_world_deltas: deltas is null");
        return 0;
    }
    s->deltas.insert(s->deltas.end(), deltas, deltas + count);
    for (int i = 0; i < count; ++i) {
        // ref: 2811-2825: data0 names the world (0 -> 1); type 6 drops the world
        long long wkey = deltas[i].data0 == 0 ? 1 : deltas[i].data0;
        if (deltas[i].type == 6) {
            auto it = s->worlds.find(wkey);
            if (it != s->worlds.end()) {
                for (auto& b : it->second.bricks) release_context(b.second);
                s->worlds.erase(it);
            }
            continue;
        }
        auto it = s->worlds.find(wkey);
        if (it != s->worlds.end()) it->second.pending.push_back(deltas[i]);
    }
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_ensure_brick_world_runtime(long long skey, long long wkey,
                                                                                   int size, float dx, float dt) {
    std::lock_guard<std::mutex> lk(g_mu);
    Service* s = lookup(skey);
    if (!s) {
        set_err("simulation_ensure_brick_world_runtime: missing service");
        return 0;
    }
    if (wkey == 0) {
        set_err("simulation_brick_world_runtime: invalid world key");
        return 0;
    }
    if (size <= 0 || !std::isfinite(dx) || dx <= 0.0f || !std::isfinite(dt) || dt <= 0.0f) {
        set_err("simulation_brick_world_runtime: invalid runtime config");
        return 0;
    }
    auto res = s->worlds.try_emplace(wkey);
    World& w = res.first->second;
    if (res.second || w.size != size || w.dx != dx || w.dt != dt) {
        for (auto& b : w.bricks) release_context(b.second);
        w = World{};
        w.size = size;
        w.dx = dx;
        w.dt = dt;
    }
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_solver_mode(long long skey, long long wkey,
                                                                                    int mode) {
    std::lock_guard<std::mutex> lk(g_mu);
    World* w = find_world(lookup(skey), wkey, "simulation_set_brick_world_solver_mode");
    if (!w) return 0;
    const int m = (mode >= AERO_LBM_REALTIME_SOLVER_AUTO && mode <= AERO_LBM_REALTIME_SOLVER_D3Q27_S16_INPLACE)
                      ? mode
                      : AERO_LBM_REALTIME_SOLVER_CLASSIC_D3Q27;
    if (m != w->mode) {
        // contexts were created for the previous storage: restart them
        for (auto& b : w->bricks) {
            if (b.second.context != 0 && b.second.stale) sync_dynamic(*w, b.second);
            release_context(b.second);
        }
        w->mode = m;
    }
    return 1;
}

// extension: batched stepping of a world's cached in-place bricks (on by default);
// off, every brick steps through its own cached-step call as the reference does
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_batching(long long skey, long long wkey,
                                                                                int enabled) {
    std::lock_guard<std::mutex> lk(g_mu);
    World* w = find_world(lookup(skey), wkey, "simulation_set_brick_world_batching");
    if (!w) return 0;
    w->batching = enabled != 0;
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_active_hints(long long skey, long long wkey,
                                                                                     int size, const int* coords,
                                                                                     int count) {
    std::lock_guard<std::mutex> lk(g_mu);
    return set_hints(skey, wkey, size, coords, count, true, "simulation_set_brick_world_active_hints");
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_brick_world_exact_active_hints(long long skey,
                                                                                           long long wkey, int size,
                                                                                           const int* coords,
                                                                                           int count) {
    std::lock_guard<std::mutex> lk(g_mu);
    return set_hints(skey, wkey, size, coords, count, false, "simulation_set_brick_world_exact_active_hints");
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_get_brick_world_runtime_status(long long skey, long long wkey,
                                                                                      int* out, int count) {
    std::lock_guard<std::mutex> lk(g_mu);
    Service* s = lookup(skey);
    if (!s) {
        set_err("simulation_get_brick_world_runtime_status: missing service");
        return 0;
    }
    if (!out || count < AERO_LBM_SIMULATION_BRICK_RUNTIME_STATUS_FIELDS) {
        set_err("simulation_get_brick_world_runtime_status: invalid output buffer");
        return 0;
    }
    World* w = find_world(s, wkey, "simulation_get_brick_world_runtime_status");
    if (!w) return 0;
    int hint = 0, active = 0, geo = 0, forcing = 0, reinit = 0;
    for (const auto& e : w->bricks) {
        hint += e.second.active_hint;
        active += e.second.active;
        geo += e.second.geometry_dirty;
        forcing += e.second.forcing_dirty;
        reinit += e.second.pending_reinit;
    }
    out[0] = w->size;
    out[1] = (int)w->bricks.size();
    out[2] = hint;
    out[3] = active;
    out[4] = geo;
    out[5] = forcing;
    out[6] = reinit;
    out[7] = (int)std::min<uint64_t>(w->epoch, (uint64_t)std::numeric_limits<int>::max());
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_step_brick_world_runtime(long long skey, long long wkey,
                                                                                 int steps) {
    std::lock_guard<std::mutex> lk(g_mu);
    Service* s = lookup(skey);
    if (!s) {
        set_err("simulation_step_brick_world_runtime: missing service");
        return 0;
    }
    if (steps <= 0) {
        set_err("simulation_step_brick_world_runtime: invalid step count");
        return 0;
    }
    World* w = find_world(s, wkey, "simulation_step_brick_world_runtime");
    if (!w) return 0;
    return step_world(*s, wkey, *w, steps) ? 1 : 0;
}

static int count_bricks(long long skey, long long wkey, bool want_resident, const char* who) {
    std::lock_guard<std::mutex> lk(g_mu);
    World* w = find_world(lookup(skey), wkey, who);
    if (!w) return -1;
    int n = 0;
    for (const auto& e : w->bricks) n += want_resident ? resident(e.second) : e.second.active;
    return n;
}

static int copy_bricks(long long skey, long long wkey, int* out, int capacity, bool want_resident, const char* who) {
    std::lock_guard<std::mutex> lk(g_mu);
    World* w = find_world(lookup(skey), wkey, who);
    if (!w) return -1;
    if (capacity > 0 && !out) return -1;
    int written = 0;
    for (const auto& e : w->bricks) {
        if (want_resident ? !resident(e.second) : !e.second.active) continue;
        if (written >= capacity) return -1;
        out[3 * written] = e.first.x;
        out[3 * written + 1] = e.first.y;
        out[3 * written + 2] = e.first.z;
        ++written;
    }
    return written;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_get_brick_world_resident_brick_count(long long skey,
                                                                                            long long wkey) {
    return count_bricks(skey, wkey, true, "simulation_get_brick_world_resident_brick_count");
}
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_resident_bricks(long long skey, long long wkey,
                                                                                        int* out, int capacity) {
    return copy_bricks(skey, wkey, out, capacity, true, "simulation_copy_brick_world_resident_bricks");
}
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_get_brick_world_active_brick_count(long long skey,
                                                                                          long long wkey) {
    return count_bricks(skey, wkey, false, "simulation_get_brick_world_active_brick_count");
}
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_active_bricks(long long skey, long long wkey,
                                                                                      int* out, int capacity) {
    return copy_bricks(skey, wkey, out, capacity, false, "simulation_copy_brick_world_active_bricks");
}

static std::shared_ptr<StaticData> make_static(int n, const uint8_t* obstacle, const uint8_t* kind,
                                               const uint16_t* open, const float* emitter, const uint8_t* fan_src,
                                               const float* emitter_src, const uint8_t* sky, const uint8_t* direct) {
    auto s = std::make_shared<StaticData>();
    s->n = n;
    const size_t cells = cube(n);
    s->obstacle.assign(obstacle, obstacle + cells);
    s->surface_kind.assign(kind, kind + cells);
    s->open_face.assign(open, open + cells);
    s->emitter.assign(emitter, emitter + cells);
    if (fan_src)
        s->fan_dir_src.assign(fan_src, fan_src + cells);
    else
        s->fan_dir_src.assign(cells, 0);
    if (emitter_src)
        s->emitter_src.assign(emitter_src, emitter_src + cells);
    else
        s->emitter_src.assign(cells, 0.0f);
    if (fan_src || emitter_src) derive_from_sources(*s);
    s->sky.assign(sky, sky + cells * 6);
    s->direct.assign(direct, direct + cells * 6);
    return s;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_static_brick_with_sources(
    long long skey, long long wkey, int size, int bx, int by, int bz, const uint8_t* obstacle,
    const uint8_t* surface_kind, const uint16_t* open_face_mask, const float* emitter_power_watts,
    const uint8_t* source_fan_dir, const float* source_emitter_power_watts, const uint8_t* face_sky_exposure,
    const uint8_t* face_direct_exposure) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cells = 0;
    if (!cell_count(size, size, size, &cells)) {
        set_err("simulation_upload_brick_world_static_brick: invalid brick size");
        return 0;
    }
    if (!obstacle || !surface_kind || !open_face_mask || !emitter_power_watts || !face_sky_exposure ||
        !face_direct_exposure) {
        set_err("simulation_upload_brick_world_static_brick: null brick buffers");
        return 0;
    }
    World* w = find_world(lookup(skey), wkey, "simulation_upload_brick_world_static_brick");
    if (!w) return 0;
    if (w->size != size) {
        set_err("simulation_upload_brick_world_static_brick: brick size mismatch");
        return 0;
    }
    install_static(*w, w->bricks[Coord{bx, by, bz}],
                   make_static(size, obstacle, surface_kind, open_face_mask, emitter_power_watts, source_fan_dir,
                               source_emitter_power_watts, face_sky_exposure, face_direct_exposure));
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_static_brick(
    long long skey, long long wkey, int size, int bx, int by, int bz, const uint8_t* obstacle,
    const uint8_t* surface_kind, const uint16_t* open_face_mask, const float* emitter_power_watts,
    const uint8_t* face_sky_exposure, const uint8_t* face_direct_exposure) {
    return aero_lbm_simulation_upload_brick_world_static_brick_with_sources(
        skey, wkey, size, bx, by, bz, obstacle, surface_kind, open_face_mask, emitter_power_watts, nullptr, nullptr,
        face_sky_exposure, face_direct_exposure);
}

// queued upload: installed at the next world step by a delta of type 9 whose data1
// is the returned id (ref: 5033-5112)
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_queue_brick_world_static_brick_upload(
    long long skey, long long wkey, int size, int bx, int by, int bz, const uint8_t* obstacle,
    const uint8_t* surface_kind, const uint16_t* open_face_mask, const float* emitter_power_watts,
    const uint8_t* face_sky_exposure, const uint8_t* face_direct_exposure) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cells = 0;
    if (!cell_count(size, size, size, &cells) || !obstacle || !surface_kind || !open_face_mask ||
        !emitter_power_watts || !face_sky_exposure || !face_direct_exposure) {
        set_err("simulation_queue_brick_world_static_brick_upload: invalid arguments");
        return 0;
    }
    Service* s = lookup(skey);
    World* w = find_world(s, wkey, "simulation_queue_brick_world_static_brick_upload");
    if (!w) return 0;
    if (w->size != size) {
        set_err("simulation_queue_brick_world_static_brick_upload: brick size mismatch");
        return 0;
    }
    const int id = s->next_upload++;
    if (s->next_upload <= 0) s->next_upload = 1;
    s->uploads[id] = PendingUpload{wkey, Coord{bx, by, bz},
                                   make_static(size, obstacle, surface_kind, open_face_mask, emitter_power_watts,
                                               nullptr, nullptr, face_sky_exposure, face_direct_exposure)};
    return id;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_dynamic_brick(
    long long skey, long long wkey, int size, int bx, int by, int bz, const float* flow, const float* air,
    const float* surface) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cells = 0;
    if (!cell_count(size, size, size, &cells) || !flow || !air || !surface) {
        set_err("simulation_upload_brick_world_dynamic_brick: invalid arguments");
        return 0;
    }
    World* w = find_world(lookup(skey), wkey, "simulation_upload_brick_world_dynamic_brick");
    if (!w) return 0;
    if (w->size != size) {
        set_err("simulation_upload_brick_world_dynamic_brick: brick size mismatch");
        return 0;
    }
    Brick& b = w->bricks[Coord{bx, by, bz}];
    if (b.context != 0) {
        release_context(b);
        w->dynamic_upload_resets++;
    }
    b.dyn = make_dynamic(size, flow, air, surface);
    b.compact = false;
    b.stale = false;
    b.active = true;
    b.last_active_epoch = w->epoch;
    b.pending_reinit = false;
    b.forcing_dirty = false;
    apply_constraints(*w, b);
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_upload_brick_world_boundary_reference_brick(
    long long skey, long long wkey, int size, int bx, int by, int bz, const float* flow, const float* air,
    const float* surface) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cells = 0;
    if (!cell_count(size, size, size, &cells) || !flow || !air || !surface) {
        set_err("simulation_upload_brick_world_boundary_reference_brick: invalid arguments");
        return 0;
    }
    World* w = find_world(lookup(skey), wkey, "simulation_upload_brick_world_boundary_reference_brick");
    if (!w) return 0;
    if (w->size != size) {
        set_err("simulation_upload_brick_world_boundary_reference_brick: brick size mismatch");
        return 0;
    }
    const Coord c{bx, by, bz};
    Brick& b = w->bricks[c];
    b.bref = make_dynamic(size, flow, air, surface);
    b.active = true;
    b.last_active_epoch = w->epoch;
    update_shell(*w, c, b);
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_dynamic_brick(
    long long skey, long long wkey, int size, int bx, int by, int bz, float* out_flow, float* out_air,
    float* out_surface) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cells = 0;
    if (!cell_count(size, size, size, &cells) || !out_flow || !out_air || !out_surface) {
        set_err("simulation_copy_brick_world_dynamic_brick: invalid arguments");
        return 0;
    }
    World* w = find_world(lookup(skey), wkey, "simulation_copy_brick_world_dynamic_brick");
    if (!w) return 0;
    if (w->size != size) {
        set_err("simulation_copy_brick_world_dynamic_brick: brick size mismatch");
        return 0;
    }
    auto it = w->bricks.find(Coord{bx, by, bz});
    if (it == w->bricks.end() || !it->second.dyn) return 0;
    Brick& b = it->second;
    if (b.stale && b.context != 0 && !sync_dynamic(*w, b)) {
        set_err(std::string("simulation_copy_brick_world_dynamic_brick sync failed: ") + aero_lbm_last_error());
        return 0;
    }
    if (!dyn_valid(*w, b.dyn.get())) {
        set_err("simulation_copy_brick_world_dynamic_brick: invalid stored brick dynamic state");
        return 0;
    }
    std::copy(b.dyn->flow.begin(), b.dyn->flow.end(), out_flow);
    std::copy(b.dyn->air.begin(), b.dyn->air.end(), out_air);
    std::copy(b.dyn->surface.begin(), b.dyn->surface.end(), out_surface);
    return 1;
}

// int16 atlas of a brick: every stride-th cell (clamped to the last), velocity in
// m/s over +-5.6, pressure over +-0.03 (ref: 5301-5410, quantisation 3122-3128)
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_copy_brick_world_packed_flow_atlas(
    long long skey, long long wkey, int size, int bx, int by, int bz, int stride, int16_t* out, int value_count) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cells = 0;
    if (!cell_count(size, size, size, &cells) || stride <= 0 || !out) {
        set_err("simulation_copy_brick_world_packed_flow_atlas: invalid arguments");
        return 0;
    }
    const int sn = (size + stride - 1) / stride;
    int acells = 0;
    if (!cell_count(sn, sn, sn, &acells) || value_count != acells * kAtlasCh) {
        set_err("simulation_copy_brick_world_packed_flow_atlas: output size mismatch");
        return 0;
    }
    World* w = find_world(lookup(skey), wkey, "simulation_copy_brick_world_packed_flow_atlas");
    if (!w) return 0;
    if (w->size != size) {
        set_err("simulation_copy_brick_world_packed_flow_atlas: brick size mismatch");
        return 0;
    }
    auto it = w->bricks.find(Coord{bx, by, bz});
    if (it == w->bricks.end() || !it->second.dyn) return 0;
    Brick& b = it->second;
    const float vscale = w->dt > 0.0f ? w->dx / w->dt : 1.0f;
    // an in-place context publishes on the device: int16 quantisation there, a quarter
    // of the float atlas read back. With batching, the first request after a world
    // step publishes every in-place brick of the world in one launch and one readback.
    if (b.context != 0 && experimental(*w) && aero::context_batchable(b.context)) {
        AtlasCache& ac = w->atlas_cache;
        const Coord at{bx, by, bz};
        auto cached = [&]() -> const int16_t* {
            if (ac.epoch != w->epoch || ac.stride != stride || ac.vscale != vscale) return nullptr;
            auto hit = ac.index.find(at);
            return (hit == ac.index.end() || hit->second.first != b.context) ? nullptr : ac.data.data() + hit->second.second;
        };
        const int16_t* src = cached();
        if (!src && w->batching) {
            std::vector<long long> keys;
            std::vector<Coord> coords;
            for (const auto& e : w->bricks)
                if (e.second.context != 0 && aero::context_batchable(e.second.context)) {
                    keys.push_back(e.second.context);
                    coords.push_back(e.first);
                }
            std::vector<size_t> offsets;
            ac.index.clear();
            ac.epoch = ~0ull;
            if (aero::publish_contexts_atlas(keys, stride, vscale, ac.data, offsets)) {
                for (size_t i = 0; i < keys.size(); ++i) ac.index[coords[i]] = {keys[i], offsets[i]};
                ac.epoch = w->epoch;
                ac.stride = stride;
                ac.vscale = vscale;
                src = cached();
            }
        }
        if (src) {
            std::copy_n(src, (size_t)value_count, out);
            return 1;
        }
        if (aero::publish_context_atlas(b.context, stride, vscale, out, value_count)) return 1;
    }
    if (b.context != 0) {
        std::vector<float> atlas((size_t)acells * kFlowCh);
        if (!aero_lbm_extract_flow_atlas_rect(size, size, size, b.context, stride, atlas.data(), (int)atlas.size())) {
            set_err(std::string("simulation_copy_brick_world_packed_flow_atlas extract failed: ") +
                    aero_lbm_last_error());
            return 0;
        }
        for (int i = 0; i < acells; ++i) {
            for (int k = 0; k < 3; ++k) out[4 * i + k] = quantize(atlas[4 * i + k] * vscale, kAtlasVelocityRange);
            out[4 * i + 3] = quantize(atlas[4 * i + 3], kAtlasPressureRange);
        }
        return 1;
    }
    if (!dyn_valid(*w, b.dyn.get())) {
        set_err("simulation_copy_brick_world_packed_flow_atlas: invalid stored brick dynamic state");
        return 0;
    }
    size_t dst = 0;
    for (int x = 0; x < sn; ++x)
        for (int y = 0; y < sn; ++y)
            for (int z = 0; z < sn; ++z) {
                const int g = cell_of(size, std::min(size - 1, x * stride), std::min(size - 1, y * stride),
                                      std::min(size - 1, z * stride));
                const float* f = b.dyn->flow.data() + (size_t)g * kFlowCh;
                for (int k = 0; k < 3; ++k) out[dst + k] = quantize(f[k] * vscale, kAtlasVelocityRange);
                out[dst + 3] = quantize(f[3], kAtlasPressureRange);
                dst += kAtlasCh;
            }
    return 1;
}

// ref: simulation_bridge.cpp:6711-6762
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_set_packed_flow_atlas(long long service_key, long long atlas_key,
                                                                             const int16_t* values, int value_count) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (value_count <= 0 || !values) {
        set_err("simulation_set_packed_flow_atlas: invalid atlas buffer");
        return 0;
    }
    Service* svc = lookup(service_key);
    if (!svc) {
        set_err("simulation_set_packed_flow_atlas: missing service");
        return 0;
    }
    svc->atlases[atlas_key].assign(values, values + value_count);
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_simulation_poll_packed_flow_atlas(long long service_key,
                                                                              long long atlas_key, int16_t* out,
                                                                              int value_count) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (value_count <= 0 || !out) {
        set_err("simulation_poll_packed_flow_atlas: invalid output buffer");
        return 0;
    }
    Service* svc = lookup(service_key);
    if (!svc) {
        set_err("simulation_poll_packed_flow_atlas: missing service");
        return 0;
    }
    auto it = svc->atlases.find(atlas_key);
    if (it == svc->atlases.end()) {
        set_err("simulation_poll_packed_flow_atlas: missing atlas");
        return 0;
    }
    if ((int)it->second.size() != value_count) {
        set_err("simulation_poll_packed_flow_atlas: size mismatch");
        return 0;
    }
    std::copy(it->second.begin(), it->second.end(), out);
    return 1;
}

// ref: simulation_bridge.cpp:6866-6990 (the text's fields are this build's own)
extern "C" AERO_LBM_CAPI_EXPORT const char* aero_lbm_simulation_runtime_info() {
    static thread_local std::string text;
    std::lock_guard<std::mutex> lk(g_mu);
    size_t worlds = 0, bricks = 0, active = 0, contexts = 0, compact = 0, atlases = 0;
    uint64_t cached = 0, packet = 0, resets = 0, p_try = 0, p_ok = 0, p_fail = 0, p_cells = 0, epoch = 0;
    uint64_t batched = 0;
    double h_total = 0, h_prep = 0, h_batch = 0;
    int mode = -1;
    bool mixed = false;
    if (g_service) {
        atlases = g_service->atlases.size();
        worlds = g_service->worlds.size();
        for (const auto& wv : g_service->worlds) {
            const World& w = wv.second;
            mixed |= mode >= 0 && mode != w.mode;
            mode = w.mode;
            bricks += w.bricks.size();
            for (const auto& bv : w.bricks) {
                active += bv.second.active ? 1 : 0;
                contexts += bv.second.context != 0 ? 1 : 0;
                compact += bv.second.compact ? 1 : 0;
            }
            cached += w.cached_steps;
            batched += w.batched_steps;
            h_total += w.host_ms_total;
            h_prep += w.host_ms_prepare;
            h_batch += w.host_ms_batch;
            packet += w.packet_steps;
            resets += w.dynamic_upload_resets;
            p_try += w.patch_attempts;
            p_ok += w.patch_steps;
            p_fail += w.patch_failures;
            p_cells += w.patch_cells;
            epoch = std::max(epoch, w.epoch);
        }
    }
    text = "service=" + std::to_string(g_service ? 1 : 0) + " worlds=" + std::to_string(worlds) +
           " solver_mode=" + (mixed ? std::string("mixed") : std::to_string(mode)) +
           " bricks=" + std::to_string(bricks) + " active=" + std::to_string(active) +
           " contexts=" + std::to_string(contexts) + " cached_contexts=" + std::to_string(compact) +
           " atlases=" + std::to_string(atlases) + " cached_steps=" + std::to_string(cached) +
           " batched_world_steps=" + std::to_string(batched) + " world_host_ms=" + std::to_string(h_total) +
           " world_host_ms_prepare=" + std::to_string(h_prep) + " world_host_ms_batch=" + std::to_string(h_batch) +
           " packet_steps=" + std::to_string(packet) + " dynamic_upload_resets=" + std::to_string(resets) +
           " patch_attempts=" + std::to_string(p_try) + " patch_steps=" + std::to_string(p_ok) +
           " patch_failures=" + std::to_string(p_fail) + " patch_cells=" + std::to_string(p_cells) +
           " epoch=" + std::to_string(epoch) + " lbm=" + (aero_lbm_runtime_info() ? aero_lbm_runtime_info() : "");
    return text.c_str();
}

extern "C" AERO_LBM_CAPI_EXPORT const char* aero_lbm_simulation_last_error() {
    std::lock_guard<std::mutex> lk(g_mu);
    static thread_local std::string copy;
    copy = g_error;
    return copy.c_str();
}
