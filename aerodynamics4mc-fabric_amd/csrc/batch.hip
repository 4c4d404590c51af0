// Batched stepping of many in-place contexts: every brick of a batch advances in
// the same launches (rare-cell gather, rare-cell update and bulk kernel per step),
// driven by device tables of per-brick parameters and per-block owners, and the
// cycle "K steps + int16 atlas publish of every brick" is captured once in a HIP
// graph and replayed. Bricks of any size mix in one launch; each keeps the
// parameters of its last individual step (its own inflow, faces and forcing).
//
// The A-A layout, face-store index and temperature buffer of a context alternate
// every step, so a cycle of an even number of steps returns every context to its
// starting parity: the tables of the K steps are built once and the graph can be
// replayed any number of times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aero_lbm_capi.h"
#include "fast_kernels.h"
#include "runtime_internal.h"

namespace aero {

namespace {

struct Member {
    long long key;
    Context* ctx;
    int layout, store_cur, t_cur;  // parity at creation, restored after every cycle
    size_t atlas_offset;           // int16 values before this brick in the atlas buffer
    int atlas_values;
    uint64_t gen;                  // the context's generation the tables were built from
};

struct Batch {
    std::vector<Member> members;
    FastStorage storage = FastStorage::F16;
    int steps = 2;                // steps per publish cycle (even)
    int stride = 4;               // atlas sample stride
    int policy = 0;               // store cache policy of the bulk kernel
    bool shell = false;
    std::vector<int> order;       // launch order of the members (autotuned)
    BatchBrick* tables = nullptr; // steps x members entries
    uint16_t* owners = nullptr;
    std::vector<BatchGrids> grids;  // per step (same owners for every step)
    int16_t* atlas_dev = nullptr;
    int16_t* atlas_host = nullptr;  // pinned
    size_t atlas_values = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<float> scales;
    std::string tuning;
    ~Batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (tables) (void)hipFree(tables);
        if (owners) (void)hipFree(owners);
        if (atlas_dev) (void)hipFree(atlas_dev);
        if (atlas_host) (void)hipHostFree(atlas_host);
    }
};

std::unordered_map<long long, std::unique_ptr<Batch>> g_batches;
long long g_next_batch = 1;

FastStorage kind_of(Storage s) {
    return s == Storage::F16    ? FastStorage::F16
           : s == Storage::BF16 ? FastStorage::BF16
           : s == Storage::S16  ? FastStorage::S16
                                : FastStorage::F32;
}
void* aa_buf(Context& c) { return c.storage == Storage::F32AA ? (void*)c.f32[0] : (void*)c.h16; }
int flip(int layout) { return layout == 1 ? 2 : 1; }
size_t elem_bytes(Storage s) { return s == Storage::F32AA ? 4 : 2; }

bool members_alive(const Batch& b) {
    for (const Member& m : b.members) {
        if (find_context(m.key) != m.ctx) {
            set_error("batch: a member context was released or re-created; create the batch again");
            return false;
        }
        if (m.ctx->generation != m.gen || m.ctx->aa_layout != m.layout || m.ctx->store_cur != m.store_cur ||
            m.ctx->t_cur != m.t_cur) {
            set_error("batch: a member context was stepped or changed outside the batch; create the batch again");
            return false;
        }
    }
    return true;
}

// Table entry of context c for a step that reads face store sc and temperature
// buffer tc, with the context's cached parameters. Every byte is set (padding
// zeroed), so two entries compare equal with memcmp exactly when they describe the
// same step.
void fill_entry(BatchBrick& e, Context& c, int sc, int tc) {
    std::memset(&e, 0, sizeof(e));
    std::memcpy(&e.P, &c.params_host, sizeof(StepParams));
    const CellFields C = cell_fields(c);
    std::memcpy(&e.C, &C, sizeof(CellFields));
    e.mask = c.mask;
    e.f = aa_buf(c);
    e.t_in = c.T[tc];
    e.t_out = c.T[1 - tc];
    e.u_state = c.u_state;
    e.store_in = c.face_store[sc];
    e.store_out = c.face_store[1 - sc];
    e.rare_list = c.rare_list;
    e.rare_count = c.rare_count;
    e.rare_cap = c.rare_cap;
    e.rare_vals = c.rare_vals;
}

// tables and owner maps for the current member order
bool build_tables(Batch& b) {
    const int nm = (int)b.members.size();
    const int K = b.steps;
    std::vector<BatchBrick> host((size_t)K * nm);
    std::vector<uint16_t> bulk, gather, update, publish;
    for (int oi = 0; oi < nm; ++oi) {
        const int mi = b.order[oi];
        const Member& m = b.members[mi];
        Context& c = *m.ctx;
        const StepParams& P = c.params_host;
        const int tiles = fast_tile_blocks(P);
        const int gblocks = (c.rare_count + 255) / 256, ublocks = (c.rare_count + 127) / 128;
        const int sn = (c.nx + b.stride - 1) / b.stride, sy = (c.ny + b.stride - 1) / b.stride,
                  sz = (c.nz + b.stride - 1) / b.stride;
        const int pblocks = (sn * sy * sz + 255) / 256;
        const int bulk_begin = (int)bulk.size(), gather_begin = (int)gather.size(),
                  update_begin = (int)update.size(), publish_begin = (int)publish.size();
        bulk.insert(bulk.end(), tiles, (uint16_t)oi);
        gather.insert(gather.end(), gblocks, (uint16_t)oi);
        update.insert(update.end(), ublocks, (uint16_t)oi);
        publish.insert(publish.end(), pblocks, (uint16_t)oi);
        for (int j = 0; j < K; ++j) {
            BatchBrick& e = host[(size_t)j * nm + oi];
            const int layout_before = (j & 1) ? flip(m.layout) : m.layout;
            const int sc = m.store_cur ^ (j & 1);
            const int tc = P.thermal_on ? (m.t_cur ^ (j & 1)) : m.t_cur;
            fill_entry(e, c, sc, tc);
            e.bulk_begin = bulk_begin;
            e.gather_begin = gather_begin;
            e.update_begin = update_begin;
            e.publish_begin = publish_begin;
            e.publish_layout = (layout_before == 2) ? 1 : 2;  // even step -> layout 1, odd -> 2
            e.publish_stride = b.stride;
            e.publish_scale = b.scales[mi];
            e.publish_out = b.atlas_dev + m.atlas_offset;
        }
    }
    if (bulk.size() > 65535u * 64u || nm > 65535) {
        set_error("batch: too many bricks");
        return false;
    }
    if (b.tables) (void)hipFree(b.tables);
    if (b.owners) (void)hipFree(b.owners);
    b.tables = nullptr;
    b.owners = nullptr;
    const size_t nown = bulk.size() + gather.size() + update.size() + publish.size();
    hipStream_t st = rt().stream;
    if (!hip_ok(hipMalloc((void**)&b.tables, host.size() * sizeof(BatchBrick)), "hipMalloc") ||
        !hip_ok(hipMalloc((void**)&b.owners, std::max<size_t>(1, nown) * sizeof(uint16_t)), "hipMalloc"))
        return false;
    std::vector<uint16_t> all;
    all.reserve(nown);
    all.insert(all.end(), bulk.begin(), bulk.end());
    all.insert(all.end(), gather.begin(), gather.end());
    all.insert(all.end(), update.begin(), update.end());
    all.insert(all.end(), publish.begin(), publish.end());
    if (!hip_ok(hipMemcpyAsync(b.tables, host.data(), host.size() * sizeof(BatchBrick), hipMemcpyHostToDevice, st),
                "upload batch tables") ||
        !hip_ok(hipMemcpyAsync(b.owners, all.data(), all.size() * sizeof(uint16_t), hipMemcpyHostToDevice, st),
                "upload batch owners") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return false;
    BatchGrids g{};
    g.bulk_owner = b.owners;
    g.gather_owner = b.owners + bulk.size();
    g.update_owner = g.gather_owner + gather.size();
    g.publish_owner = g.update_owner + update.size();
    g.bulk_blocks = (int)bulk.size();
    g.gather_blocks = (int)gather.size();
    g.update_blocks = (int)update.size();
    g.publish_blocks = (int)publish.size();
    g.yface = 0;
    for (const BatchBrick& e : host)
        if (e.P.yface_bulk[0] || e.P.yface_bulk[1]) g.yface = 1;
    b.grids.assign(K, g);
    return true;
}

// one cycle: K batched steps, the atlas publish of every brick, its copy to pinned host memory
void enqueue_cycle(Batch& b, hipStream_t st) {
    const int nm = (int)b.members.size();
    for (int j = 0; j < b.steps; ++j) {
        const int layout_before = (j & 1) ? flip(b.members[0].layout) : b.members[0].layout;
        const int mode = layout_before == 2 ? 1 : 2;
        launch_batch_step(b.storage, mode, b.policy, b.shell, b.tables + (size_t)j * nm, b.grids[j], st);
    }
    launch_batch_publish(b.storage, b.tables + (size_t)(b.steps - 1) * nm, b.grids.back(), st);
    (void)hipMemcpyAsync(b.atlas_host, b.atlas_dev, b.atlas_values * sizeof(int16_t), hipMemcpyDeviceToHost, st);
}

bool capture(Batch& b) {
    hipStream_t st = rt().stream;
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    if (!hip_ok(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) return false;
    enqueue_cycle(b, st);
    if (!hip_ok(hipStreamEndCapture(st, &b.graph), "hipStreamEndCapture")) return false;
    return hip_ok(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

bool run_cycles(Batch& b, int cycles, bool graph, float* out_ms) {
    hipStream_t st = rt().stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (out_ms && (!hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate")))
        return false;
    if (e0) (void)hipEventRecord(e0, st);
    for (int i = 0; i < cycles; ++i) {
        if (graph)
            (void)hipGraphLaunch(b.exec, st);
        else
            enqueue_cycle(b, st);
    }
    if (e1) (void)hipEventRecord(e1, st);
    bool ok = hip_ok(hipGetLastError(), "batch run") && hip_ok(hipStreamSynchronize(st), "sync");
    if (ok && out_ms) ok = hip_ok(hipEventElapsedTime(out_ms, e0, e1), "hipEventElapsedTime");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (ok)
        for (Member& m : b.members) m.ctx->steps += (uint64_t)cycles * b.steps;
    return ok;
}

// Device copies of every buffer a step writes (populations, temperature pair,
// carried velocity, face stores), so timing runs leave the simulation unchanged.
struct Snapshot {
    std::vector<std::pair<void*, void*>> saved;  // (live buffer, copy)
    std::vector<size_t> bytes;
    std::vector<uint64_t> steps;
    ~Snapshot() {
        for (auto& p : saved) (void)hipFree(p.second);
    }
    bool add(void* live, size_t n) {
        if (!live || n == 0) return true;
        void* copy = nullptr;
        if (!hip_ok(hipMalloc(&copy, n), "hipMalloc") ||
            !hip_ok(hipMemcpyAsync(copy, live, n, hipMemcpyDeviceToDevice, rt().stream), "snapshot"))
            return false;
        saved.emplace_back(live, copy);
        bytes.push_back(n);
        return true;
    }
    bool take(const Batch& b) {
        for (const Member& m : b.members) {
            Context& c = *m.ctx;
            steps.push_back(c.steps);
            if (!add(aa_buf(c), 27 * elem_bytes(c.storage) * (size_t)c.params_host.plane) ||
                !add(c.T[0], c.cells * 4) || !add(c.T[1], c.cells * 4) || !add(c.u_state, 3 * c.cells * 4) ||
                !add(c.face_store[0], c.face_store_floats * 4) || !add(c.face_store[1], c.face_store_floats * 4))
                return false;
        }
        return true;
    }
    bool restore(Batch& b) {
        for (size_t i = 0; i < saved.size(); ++i)
            if (!hip_ok(hipMemcpyAsync(saved[i].first, saved[i].second, bytes[i], hipMemcpyDeviceToDevice,
                                       rt().stream), "restore"))
                return false;
        for (size_t i = 0; i < b.members.size(); ++i) b.members[i].ctx->steps = steps[i];
        return hip_ok(hipStreamSynchronize(rt().stream), "sync");
    }
};

// The autotuner times a few cycles of each launch layout and keeps the fastest:
// brick order in the grids (largest bricks first, smallest first, as given) and
// the bulk kernel's store cache policy (default, or nontemporal streaming stores).
// The bricks' state is saved before and restored after the timing runs.
bool autotune(Batch& b, size_t working_set) {
    const int nm = (int)b.members.size();
    std::vector<int> asis(nm), big(nm), small(nm);
    std::iota(asis.begin(), asis.end(), 0);
    big = asis;
    std::stable_sort(big.begin(), big.end(), [&](int a, int c) { return b.members[a].ctx->cells > b.members[c].ctx->cells; });
    small = big;
    std::reverse(small.begin(), small.end());
    const std::vector<int>* orders[3] = {&big, &small, &asis};
    const char* order_names[3] = {"largest-first", "smallest-first", "given"};
    const int pols[2] = {fast_store_policy(working_set), fast_store_policy(working_set) == 0 ? 2 : 0};
    Snapshot snap;
    if (!snap.take(b)) return false;
    float best = 1e30f;
    int best_o = 0, best_p = pols[0];
    for (int o = 0; o < 3; ++o) {
        b.order = *orders[o];
        if (!build_tables(b)) return false;
        for (int p = 0; p < 2; ++p) {
            b.policy = pols[p];
            float ms = 0.0f;
            if (!run_cycles(b, 1, false, nullptr)) return false;  // warm
            if (!run_cycles(b, 2, false, &ms)) return false;
            if (ms < best) {
                best = ms;
                best_o = o;
                best_p = pols[p];
            }
        }
    }
    if (!snap.restore(b)) return false;
    b.order = *orders[best_o];
    b.policy = best_p;
    b.tuning = std::string("order=") + order_names[best_o] + " store=" + (best_p == 0 ? "default" : "nontemporal") +
               " ms_per_cycle=" + std::to_string(best / 2.0f);
    return build_tables(b);
}

// ---------------------------------------------------------------- one step of resident contexts
// The brick-world runtime steps its cached in-place bricks through here, one world
// step at a time: the contexts of one storage and A-A parity form a group that
// takes one rare-gather, one rare-update and one bulk launch. A group keeps the
// device copy of its table and block owners and uploads them again only when an
// entry changed (parameters, buffers, rare-cell list, membership).
struct StepGroup {
    std::vector<BatchBrick> host;  // the entries as last uploaded
    std::vector<uint16_t> owners_host;
    BatchBrick* table = nullptr;
    uint16_t* owners = nullptr;
    size_t table_cap = 0, owners_cap = 0;
    BatchGrids grids{};
    ~StepGroup() {
        if (table) (void)hipFree(table);
        if (owners) (void)hipFree(owners);
    }
};

// key: storage kind, step mode (1 even, 2 odd) and shell instantiation
std::unordered_map<int, std::unique_ptr<StepGroup>> g_step_groups;

bool batchable(const Context& c) {
    const StepParams& P = c.params_host;
    return (storage_16bit(c.storage) || c.storage == Storage::F32AA) && c.initialized && c.has_payload &&
           c.params_valid && (c.aa_layout == 1 || c.aa_layout == 2) && !P.runtime_mode && !P.thermal_bfecc &&
           !P.cylinder && !P.fan_on;
}

bool step_group(const std::vector<Context*>& ctxs, FastStorage kind, int mode, bool shell) {
    const int nm = (int)ctxs.size();
    std::vector<BatchBrick> fresh((size_t)nm);
    std::vector<uint16_t> bulk, gather, update;
    size_t working_set = 0;
    for (int i = 0; i < nm; ++i) {
        Context& c = *ctxs[i];
        BatchBrick& e = fresh[(size_t)i];
        fill_entry(e, c, c.store_cur, c.t_cur);
        e.bulk_begin = (int)bulk.size();
        e.gather_begin = (int)gather.size();
        e.update_begin = (int)update.size();
        bulk.insert(bulk.end(), fast_tile_blocks(c.params_host), (uint16_t)i);
        gather.insert(gather.end(), (c.rare_count + 255) / 256, (uint16_t)i);
        update.insert(update.end(), (c.rare_count + 127) / 128, (uint16_t)i);
        working_set += 27 * elem_bytes(c.storage) * (size_t)c.params_host.plane + c.cells * 16;
    }
    if (bulk.size() > 65535u * 64u || nm > 65535) {
        set_error("batched world step: too many bricks");
        return false;
    }
    const int key = ((int)kind * 4 + mode) * 2 + (shell ? 1 : 0);
    std::unique_ptr<StepGroup>& slot = g_step_groups[key];
    if (!slot) slot = std::make_unique<StepGroup>();
    StepGroup& g = *slot;
    hipStream_t st = rt().stream;
    std::vector<uint16_t> owners;
    owners.reserve(bulk.size() + gather.size() + update.size());
    owners.insert(owners.end(), bulk.begin(), bulk.end());
    owners.insert(owners.end(), gather.begin(), gather.end());
    owners.insert(owners.end(), update.begin(), update.end());
    const bool same = g.host.size() == fresh.size() && g.owners_host == owners &&
                      std::memcmp(g.host.data(), fresh.data(), fresh.size() * sizeof(BatchBrick)) == 0;
    if (!same) {
        // the previous upload may still read the host copies: wait before replacing them
        if (!hip_ok(hipStreamSynchronize(st), "sync")) return false;
        g.host.swap(fresh);
        g.owners_host.swap(owners);
        if (g.host.size() > g.table_cap) {
            if (g.table) (void)hipFree(g.table);
            g.table = nullptr;
            g.table_cap = 0;
            if (!hip_ok(hipMalloc((void**)&g.table, g.host.size() * sizeof(BatchBrick)), "hipMalloc")) return false;
            g.table_cap = g.host.size();
        }
        if (g.owners_host.size() > g.owners_cap) {
            if (g.owners) (void)hipFree(g.owners);
            g.owners = nullptr;
            g.owners_cap = 0;
            if (!hip_ok(hipMalloc((void**)&g.owners, g.owners_host.size() * sizeof(uint16_t)), "hipMalloc"))
                return false;
            g.owners_cap = g.owners_host.size();
        }
        if (!hip_ok(hipMemcpyAsync(g.table, g.host.data(), g.host.size() * sizeof(BatchBrick), hipMemcpyHostToDevice,
                                   st), "upload step table") ||
            !hip_ok(hipMemcpyAsync(g.owners, g.owners_host.data(), g.owners_host.size() * sizeof(uint16_t),
                                   hipMemcpyHostToDevice, st), "upload step owners"))
            return false;
        g.grids.bulk_owner = g.owners;
        g.grids.gather_owner = g.owners + bulk.size();
        g.grids.update_owner = g.grids.gather_owner + gather.size();
        g.grids.publish_owner = nullptr;
        g.grids.bulk_blocks = (int)bulk.size();
        g.grids.gather_blocks = (int)gather.size();
        g.grids.update_blocks = (int)update.size();
        g.grids.publish_blocks = 0;
    }
    launch_batch_step(kind, mode, fast_store_policy(working_set), shell, g.table, g.grids, st);
    if (!hip_ok(hipGetLastError(), "batched world step")) return false;
    // the state changes of an individual A-A step (storage_step)
    for (Context* c : ctxs) {
        c->store_cur ^= 1;
        c->aa_layout = mode == 1 ? 1 : 2;
        c->post_collision = true;
        if (c->params_host.thermal_on) c->t_cur ^= 1;
        c->steps += 1;
    }
    return true;
}

}  // namespace

bool context_batchable(long long key) {
    std::lock_guard<std::mutex> lk(rt().mu);
    const Context* c = find_context(key);
    return c && batchable(*c);
}

Context* prepare_cached_step(long long key, int nx, int ny, int nz) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    Context* c = get_context(key, nx, ny, nz, false);
    if (!c || !c->has_payload) {
        set_error("batched world step: context " + std::to_string(key) + " has no resident payload");
        return nullptr;
    }
    // the host half of step_context with the resident payload
    StepParams P;
    if (!prepare_params(*c, P)) return nullptr;
    update_thermal_activity(*c, P);
    P.thermal_on = c->thermal_on ? 1 : 0;
    step_switches(*c, P);
    if (!sync_params(*c, P) || !storage_prepare_step(*c, P)) return nullptr;
    if (!batchable(*c)) {
        set_error("batched world step: context " + std::to_string(key) + " cannot join a batched step");
        return nullptr;
    }
    return c;
}

bool batch_step_contexts(const std::vector<Context*>& ctxs) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    // groups in first-appearance order: (storage, mode, shell)
    std::vector<std::pair<int, std::vector<Context*>>> groups;
    for (Context* c : ctxs) {
        if (!c || !batchable(*c)) {
            set_error("batched world step: a context cannot join a batched step");
            return false;
        }
        const int mode = c->aa_layout == 2 ? 1 : 2;  // after an odd step the next one is even
        const int key = ((int)kind_of(c->storage) * 4 + mode) * 2 + (c->shell_on ? 1 : 0);
        auto it = std::find_if(groups.begin(), groups.end(), [&](const auto& g) { return g.first == key; });
        if (it == groups.end()) {
            groups.emplace_back(key, std::vector<Context*>{});
            it = groups.end() - 1;
        }
        it->second.push_back(c);
    }
    for (auto& g : groups) {
        Context& c0 = *g.second.front();
        const int mode = c0.aa_layout == 2 ? 1 : 2;
        if (!step_group(g.second, kind_of(c0.storage), mode, c0.shell_on)) return false;
    }
    return true;
}

bool publish_contexts_atlas(const std::vector<long long>& keys, int stride, float vscale, std::vector<int16_t>& out,
                            std::vector<size_t>& offsets) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (stride <= 0 || keys.empty()) {
        set_error("packed atlas publish: invalid arguments");
        return false;
    }
    const int nm = (int)keys.size();
    std::vector<BatchBrick> host((size_t)nm);
    std::vector<uint16_t> owners;
    offsets.assign((size_t)nm, 0);
    size_t values = 0;
    for (int i = 0; i < nm; ++i) {
        Context* c = find_context(keys[(size_t)i]);
        if (!c || !c->initialized || !c->params_valid || !(storage_16bit(c->storage) || c->storage == Storage::F32AA)) {
            set_error("packed atlas publish: context missing or not in an in-place storage");
            return false;
        }
        const int sx = (c->nx + stride - 1) / stride, sy = (c->ny + stride - 1) / stride,
                  sz = (c->nz + stride - 1) / stride;
        const size_t n = (size_t)sx * sy * sz;
        BatchBrick& e = host[(size_t)i];
        fill_entry(e, *c, c->store_cur ^ 1, c->t_cur);  // the face store and layout of the last step
        e.publish_layout = c->aa_layout;
        e.publish_stride = stride;
        e.publish_scale = vscale;
        e.publish_begin = (int)owners.size();
        owners.insert(owners.end(), (n + 255) / 256, (uint16_t)i);
        offsets[(size_t)i] = values;
        values += 4 * n;
        if (i > 0 && kind_of(c->storage) != kind_of(find_context(keys[0])->storage)) {
            set_error("packed atlas publish: contexts of one call share a storage");
            return false;
        }
    }
    if (owners.size() > 0x7fffffffu || nm > 65535) {
        set_error("packed atlas publish: too many bricks");
        return false;
    }
    hipStream_t st = rt().stream;
    BatchBrick* table = nullptr;
    uint16_t* own = nullptr;
    int16_t* dev = nullptr;
    out.resize(values);
    bool ok = hip_ok(hipMallocAsync((void**)&table, host.size() * sizeof(BatchBrick), st), "hipMallocAsync") &&
              hip_ok(hipMallocAsync((void**)&own, owners.size() * sizeof(uint16_t), st), "hipMallocAsync") &&
              hip_ok(hipMallocAsync((void**)&dev, values * sizeof(int16_t), st), "hipMallocAsync");
    if (ok) {
        for (int i = 0; i < nm; ++i) host[(size_t)i].publish_out = dev + offsets[(size_t)i];
        ok = hip_ok(hipMemcpyAsync(table, host.data(), host.size() * sizeof(BatchBrick), hipMemcpyHostToDevice, st),
                    "upload publish table") &&
             hip_ok(hipMemcpyAsync(own, owners.data(), owners.size() * sizeof(uint16_t), hipMemcpyHostToDevice, st),
                    "upload publish owners");
    }
    if (ok) {
        BatchGrids g{};
        g.publish_owner = own;
        g.publish_blocks = (int)owners.size();
        launch_batch_publish(kind_of(find_context(keys[0])->storage), table, g, st);
        ok = hip_ok(hipGetLastError(), "publish") &&
             hip_ok(hipMemcpyAsync(out.data(), dev, values * sizeof(int16_t), hipMemcpyDeviceToHost, st),
                    "atlas readback");
    }
    if (table) (void)hipFreeAsync(table, st);
    if (own) (void)hipFreeAsync(own, st);
    if (dev) (void)hipFreeAsync(dev, st);
    return hip_ok(hipStreamSynchronize(st), "sync") && ok;
}

bool publish_context_atlas(long long key, int stride, float vscale, int16_t* out, int value_count) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    Context* c = find_context(key);
    if (!c || !c->initialized || !c->params_valid || !(storage_16bit(c->storage) || c->storage == Storage::F32AA) ||
        stride <= 0 || !out) {
        set_error("packed atlas publish: context missing or not in an in-place storage");
        return false;
    }
    const int sx = (c->nx + stride - 1) / stride, sy = (c->ny + stride - 1) / stride, sz = (c->nz + stride - 1) / stride;
    const size_t values = (size_t)sx * sy * sz * 4;
    if ((size_t)value_count != values) {
        set_error("packed atlas publish: value_count mismatch");
        return false;
    }
    // one table entry whose "step" is the context's last one: store_out is the face
    // store that step wrote and the layout is the one it left
    BatchBrick e;
    fill_entry(e, *c, c->store_cur ^ 1, c->t_cur);
    e.publish_layout = c->aa_layout;
    e.publish_stride = stride;
    e.publish_scale = vscale;
    const int blocks = (int)((values / 4 + 255) / 256);
    hipStream_t st = rt().stream;
    BatchBrick* table = nullptr;
    uint16_t* owners = nullptr;
    int16_t* dev = nullptr;
    bool ok = hip_ok(hipMallocAsync((void**)&table, sizeof(BatchBrick), st), "hipMallocAsync") &&
              hip_ok(hipMallocAsync((void**)&owners, (size_t)blocks * sizeof(uint16_t), st), "hipMallocAsync") &&
              hip_ok(hipMallocAsync((void**)&dev, values * sizeof(int16_t), st), "hipMallocAsync");
    if (ok) {
        e.publish_out = dev;
        ok = hip_ok(hipMemcpyAsync(table, &e, sizeof(e), hipMemcpyHostToDevice, st), "upload publish entry") &&
             hip_ok(hipMemsetAsync(owners, 0, (size_t)blocks * sizeof(uint16_t), st), "owners");
    }
    if (ok) {
        BatchGrids g{};
        g.publish_owner = owners;
        g.publish_blocks = blocks;
        launch_batch_publish(kind_of(c->storage), table, g, st);
        ok = hip_ok(hipGetLastError(), "publish") &&
             hip_ok(hipMemcpyAsync(out, dev, values * sizeof(int16_t), hipMemcpyDeviceToHost, st), "atlas readback");
    }
    if (table) (void)hipFreeAsync(table, st);
    if (owners) (void)hipFreeAsync(owners, st);
    if (dev) (void)hipFreeAsync(dev, st);
    return hip_ok(hipStreamSynchronize(st), "sync") && ok;
}

}  // namespace aero

using namespace aero;

// ---------------------------------------------------------------- C ABI (extension: no counterpart in the reference)

extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_batch_create(const long long* keys, int count, int steps_per_publish,
                                                          int publish_stride, const float* publish_scales,
                                                          int autotune_launch, long long* out_batch) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    if (!keys || count <= 0 || !out_batch) {
        set_error("batch_create: invalid arguments");
        return 0;
    }
    if (steps_per_publish < 2 || (steps_per_publish & 1)) {
        set_error("batch_create: steps_per_publish must be even and at least 2");
        return 0;
    }
    if (publish_stride <= 0) {
        set_error("batch_create: publish_stride must be positive");
        return 0;
    }
    auto b = std::make_unique<Batch>();
    b->steps = steps_per_publish;
    b->stride = publish_stride;
    size_t working_set = 0;
    for (int i = 0; i < count; ++i) {
        Context* c = find_context(keys[i]);
        if (!c || !c->initialized || !c->params_valid || c->steps == 0) {
            set_error("batch_create: context " + std::to_string(keys[i]) + " missing or never stepped");
            return 0;
        }
        if (!storage_16bit(c->storage) && c->storage != Storage::F32AA) {
            set_error("batch_create: batches take the in-place storages (FP16, BF16, S16, FP32 A-A) only");
            return 0;
        }
        if (i > 0 && c->storage != b->members[0].ctx->storage) {
            set_error("batch_create: all contexts of a batch use the same storage");
            return 0;
        }
        if ((c->aa_layout != 1 && c->aa_layout != 2) || (i > 0 && c->aa_layout != b->members[0].layout)) {
            set_error("batch_create: all contexts of a batch must be at the same step parity");
            return 0;
        }
        for (int k = 0; k < i; ++k)
            if (keys[k] == keys[i]) {
                set_error("batch_create: a context appears twice");
                return 0;
            }
        // the cached parameters with the context-level switches set since its last step
        StepParams P = c->params_host;
        P.thermal_on = c->thermal_on ? 1 : 0;
        P.shell_on = c->shell_on ? 1 : 0;
        P.shell_layers = c->shell_on ? c->shell_layers : 0;
        P.shell_relax = c->shell_on ? c->shell_relax : 0.0f;
        if (!sync_params(*c, P)) return 0;
        Member m{keys[i], c, c->aa_layout, c->store_cur, c->t_cur, 0, 0, c->generation};
        const int sx = (c->nx + publish_stride - 1) / publish_stride, sy = (c->ny + publish_stride - 1) / publish_stride,
                  sz = (c->nz + publish_stride - 1) / publish_stride;
        m.atlas_offset = b->atlas_values;
        m.atlas_values = sx * sy * sz * 4;
        b->atlas_values += (size_t)m.atlas_values;
        b->members.push_back(m);
        b->shell = b->shell || c->shell_on;
        working_set += 27 * elem_bytes(c->storage) * (size_t)c->params_host.plane + c->cells * 16;
        b->scales.push_back(publish_scales ? publish_scales[i] : 1.0f);
    }
    b->storage = kind_of(b->members[0].ctx->storage);
    b->policy = fast_store_policy(working_set);
    if (!hip_ok(hipMalloc((void**)&b->atlas_dev, b->atlas_values * sizeof(int16_t)), "hipMalloc") ||
        !hip_ok(hipHostMalloc((void**)&b->atlas_host, b->atlas_values * sizeof(int16_t)), "hipHostMalloc"))
        return 0;
    b->order.resize(count);
    std::iota(b->order.begin(), b->order.end(), 0);
    if (autotune_launch) {
        if (!autotune(*b, working_set)) return 0;
    } else if (!build_tables(*b)) {
        return 0;
    }
    if (!capture(*b)) return 0;
    const long long id = g_next_batch++;
    g_batches[id] = std::move(b);
    *out_batch = id;
    return 1;
}

// cycles x (steps_per_publish batched steps + atlas publish); use_graph 0 enqueues
// the same launches one by one. out_ms (may be null): HIP-event time on the stream.
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_batch_run(long long batch, int cycles, int use_graph, float* out_ms) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    auto it = g_batches.find(batch);
    if (it == g_batches.end() || cycles <= 0) {
        set_error("batch_run: unknown batch or invalid cycle count");
        return 0;
    }
    Batch& b = *it->second;
    if (!members_alive(b)) return 0;
    return run_cycles(b, cycles, use_graph != 0, out_ms) ? 1 : 0;
}

// The same cycles with one launch sequence per brick (rare gather, rare update,
// bulk per brick and step, then the atlas of each brick read back and quantised on
// the host): the per-brick pattern that batching replaces.
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_batch_run_per_brick(long long batch, int cycles, float* out_ms) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    auto it = g_batches.find(batch);
    if (it == g_batches.end() || cycles <= 0) {
        set_error("batch_run_per_brick: unknown batch or invalid cycle count");
        return 0;
    }
    Batch& b = *it->second;
    if (!members_alive(b)) return 0;
    hipStream_t st = rt().stream;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<float> atlas;
    for (int cyc = 0; cyc < cycles; ++cyc) {
        for (int j = 0; j < b.steps; ++j)
            for (Member& m : b.members)
                if (!storage_step(*m.ctx, m.ctx->params_host)) return 0;
        for (size_t mi = 0; mi < b.members.size(); ++mi) {
            Member& m = b.members[mi];
            Context& c = *m.ctx;
            const int sn = m.atlas_values / 4;
            if (!storage_readout(c, c.params_host, b.stride, 4, 1.0f, c.out_dev)) return 0;
            atlas.resize((size_t)sn * 4);
            if (!hip_ok(hipMemcpyAsync(atlas.data(), c.out_dev, atlas.size() * 4, hipMemcpyDeviceToHost, st),
                        "readback") ||
                !hip_ok(hipStreamSynchronize(st), "sync"))
                return 0;
            int16_t* dst = b.atlas_host + m.atlas_offset;
            auto quant = [](float v, float range) -> int16_t {
                if (!std::isfinite(v)) return 0;
                const float t = std::min(std::max(v / range, -1.0f), 1.0f);
                return (int16_t)std::lround(t * 32767.0f);
            };
            for (int i = 0; i < sn; ++i) {
                for (int k = 0; k < 3; ++k) dst[4 * i + k] = quant(atlas[4 * i + k] * b.scales[mi], 5.6f);
                dst[4 * i + 3] = quant(atlas[4 * i + 3], 0.03f);
            }
        }
    }
    // storage_step advanced each context by an even number of steps: its parity is back
    for (Member& m : b.members) m.ctx->steps += (uint64_t)cycles * b.steps;
    if (!hip_ok(hipStreamSynchronize(st), "sync")) return 0;
    if (out_ms) *out_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 1;
}

// the last published atlas of member `index` (4 int16 per sampled cell)
extern "C" AERO_LBM_CAPI_EXPORT int aero_lbm_batch_copy_atlas(long long batch, int index, int16_t* out,
                                                              int value_count) {
    std::lock_guard<std::mutex> lk(rt().mu);
    clear_error();
    auto it = g_batches.find(batch);
    if (it == g_batches.end() || !out || index < 0 || index >= (int)it->second->members.size()) {
        set_error("batch_copy_atlas: invalid arguments");
        return 0;
    }
    const Member& m = it->second->members[index];
    if (value_count != m.atlas_values) {
        set_error("batch_copy_atlas: value_count mismatch");
        return 0;
    }
    std::copy_n(it->second->atlas_host + m.atlas_offset, m.atlas_values, out);
    return 1;
}

extern "C" AERO_LBM_CAPI_EXPORT const char* aero_lbm_batch_info(long long batch) {
    std::lock_guard<std::mutex> lk(rt().mu);
    static thread_local std::string s;
    auto it = g_batches.find(batch);
    if (it == g_batches.end()) return "";
    const Batch& b = *it->second;
    const BatchGrids& g = b.grids.front();
    s = "bricks=" + std::to_string(b.members.size()) + " steps_per_publish=" + std::to_string(b.steps) +
        " stride=" + std::to_string(b.stride) + " bulk_blocks=" + std::to_string(g.bulk_blocks) +
        " rare_blocks=" + std::to_string(g.gather_blocks) + " publish_blocks=" + std::to_string(g.publish_blocks) +
        " policy=" + (b.policy == 0 ? "default" : "nontemporal") + (b.tuning.empty() ? "" : " tuned:" + b.tuning);
    return s.c_str();
}

extern "C" AERO_LBM_CAPI_EXPORT void aero_lbm_batch_release(long long batch) {
    std::lock_guard<std::mutex> lk(rt().mu);
    g_batches.erase(batch);
}
