// Storage-mode dispatch: which kernels advance and read a context.
//   F32Exact: FP32 A-B, reference operation order (parity path)
//   F32Fast : FP32 A-B, factorised moment transforms (performance path)
//   F16/BF16: equilibrium-shifted 16-bit populations, A-A in-place streaming
#include "fast_kernels.h"
#include "runtime_internal.h"

namespace aero {

static FastStorage fast_kind(Storage s) {
    return s == Storage::F16    ? FastStorage::F16
           : s == Storage::BF16 ? FastStorage::BF16
           : s == Storage::S16  ? FastStorage::S16
                                : FastStorage::F32;
}

// link mask and rare-cell list of the fast kernels
static bool alloc_fast_aux(Context& c) {
    const size_t N = c.cells;
    if (!(c.alloc((void**)&c.mask, N * 4) && c.alloc((void**)&c.rare_list, N * 4) &&
          c.alloc((void**)&c.rare_counter, 4) && c.alloc((void**)&c.rare_scratch, ((N + 255) / 256) * 4) &&
          true))
        return false;
    c.rare_count = 0;
    c.mask_valid = false;
    return true;
}

bool storage_alloc(Context& c) {
    const size_t N = c.cells;
    const size_t Np = (size_t)fast_plane_stride((int)c.cells);  // fast kernels' plane stride
    if (c.storage != Storage::F32Exact) {
        // the fast kernels address each population array through one buffer resource
        const size_t bytes = 27 * Np * (storage_16bit(c.storage) ? 2 : 4);
        if (bytes > (size_t)kFastMaxBufferBytes) {
            set_error("brick too large for the fast kernels (population array above 2 GiB); "
                      "use the exact kernel variant or smaller bricks");
            return false;
        }
    }
    switch (c.storage) {
        case Storage::F32Exact:
            return c.alloc((void**)&c.f32[0], 27 * N * 4) && c.alloc((void**)&c.f32[1], 27 * N * 4);
        case Storage::F32Fast:
            return c.alloc((void**)&c.f32[0], 27 * Np * 4) && c.alloc((void**)&c.f32[1], 27 * Np * 4) &&
                   alloc_fast_aux(c);
        case Storage::F16:
        case Storage::BF16:
        case Storage::S16:
            return c.alloc((void**)&c.h16, 27 * Np * 2) && alloc_fast_aux(c);
        case Storage::F32AA:
            return c.alloc((void**)&c.f32[0], 27 * Np * 4) && alloc_fast_aux(c);
    }
    return false;
}

// population buffer of an A-A context
static void* aa_buffer(Context& c) { return c.storage == Storage::F32AA ? (void*)c.f32[0] : (void*)c.h16; }

size_t context_device_bytes(const Context& c) {
    const size_t N = c.cells;
    const size_t Np = (size_t)fast_plane_stride((int)c.cells);
    size_t b = N * (1 + 4 * 4 + 4 * 2 + 12 + 16) + 64;
    switch (c.storage) {
        case Storage::F32Exact: b += 2 * 27 * N * 4; break;
        case Storage::F32Fast: b += 2 * 27 * Np * 4 + N * 4; break;
        case Storage::F16:
        case Storage::BF16:
        case Storage::S16: b += 27 * Np * 2 + N * 4; break;
        case Storage::F32AA: b += 27 * Np * 4 + N * 4; break;
    }
    b += 2 * c.face_store_floats * 4;
    if (c.payload_dev) b += c.cells * (size_t)rt().cfg.in_ch * 4;
    return b;
}

void mask_signature(const StepParams& P, int* out) {
    int k = 0;
    for (int a = 0; a < 3; ++a) {
        out[k++] = P.hydro_wrap_lo[a];
        out[k++] = P.hydro_wrap_hi[a];
    }
    for (int f = 0; f < 6; ++f) {
        out[k++] = P.face[f].hydro_kind;
        out[k++] = P.face[f].thermal_kind;
    }
    out[k++] = P.convective_disabled;
    out[k++] = P.sponge_on;
    out[k++] = P.face_store_faces;
    out[k++] = P.thermal_wrap_matches;
}

bool storage_rebuild_masks(Context& c, const StepParams& P) {
    if (!c.mask) return true;
    c.generation++;
    hipStream_t st = rt().stream;
    launch_fast_mask(P, c.flags, c.mask, st);
    launch_fast_rare_list(P, c.mask, c.rare_list, c.rare_scratch, c.rare_counter, st);
    int count = 0;
    if (!hip_ok(hipMemcpyAsync(&count, c.rare_counter, 4, hipMemcpyDeviceToHost, st), "rare count") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return false;
    c.rare_count = count;
    if (count > c.rare_cap) {
        // per-step values of the listed cells: 27 incoming populations and the temperature
        if (c.rare_vals) {
            (void)hipFree(c.rare_vals);
            for (auto& p : c.allocations)
                if (p == c.rare_vals) p = nullptr;
            c.rare_vals = nullptr;
            c.rare_cap = 0;
        }
        const int cap = (count + 255) & ~255;
        if (!c.alloc((void**)&c.rare_vals, (size_t)28 * cap * sizeof(float))) return false;
        c.rare_cap = cap;
    }
    mask_signature(P, c.mask_sig);
    c.mask_valid = true;
    return true;
}

static bool ensure_face_store(Context& c, const StepParams& P) {
    if ((size_t)P.face_store_floats <= c.face_store_floats) return true;
    if (!hip_ok(hipStreamSynchronize(rt().stream), "sync")) return false;
    c.generation++;
    for (int i = 0; i < 2; ++i) {
        if (c.face_store[i]) {
            hipFree(c.face_store[i]);
            for (auto& p : c.allocations)
                if (p == c.face_store[i]) p = nullptr;
            c.face_store[i] = nullptr;
        }
        if (!c.alloc((void**)&c.face_store[i], (size_t)P.face_store_floats * 4)) return false;
        if (!hip_ok(hipMemsetAsync(c.face_store[i], 0, (size_t)P.face_store_floats * 4, rt().stream), "memset"))
            return false;
    }
    c.face_store_floats = (size_t)P.face_store_floats;
    return true;
}

bool storage_prepare_step(Context& c, const StepParams& P) {
    return c.storage == Storage::F32Exact || ensure_face_store(c, P);
}

bool storage_step(Context& c, const StepParams& P) {
    hipStream_t st = rt().stream;
    const CellFields C = cell_fields(c);
    if (P.runtime_mode && c.storage != Storage::F32Exact) {
        set_error("runtime (non-benchmark) mode runs on the exact FP32 kernels of the classic D3Q27 solver mode; "
                  "this context was created for another storage: release it or set a benchmark configuration");
        return false;
    }
    if ((P.thermal_bfecc || P.cylinder) && c.storage != Storage::F32Exact) {
        set_error("heated-cavity and cylinder presets (BFECC temperature, inlet profile and drag) run on the "
                  "exact FP32 kernels: select AERO_LBM_KERNEL_EXACT with the classic D3Q27 solver mode");
        return false;
    }
    if (c.storage == Storage::F32Exact) {
        if (!c.initialized) {
            launch_exact_init(P, C, c.f32[c.f_cur], c.T[c.t_cur], c.u_state, rt().bench.initial_temperature, st);
            c.initialized = true;
            c.post_collision = false;
        }
        // runtime mode, after the first step: the payload temperature nudged in at the faces
        // (ref: jni.cpp:8499-8503)
        else if (P.runtime_mode && P.thermal_on && P.in_ch > 10) {
            launch_exact_t_reference(P, C, c.T[c.t_cur], st);
        }
        if (P.thermal_on && P.thermal_bfecc && P.thermal_update) {
            if (!c.t_bfecc && !c.alloc((void**)&c.t_bfecc, c.cells * 4)) return false;
            launch_exact_bfecc(P, C, c.T[c.t_cur], c.T[1 - c.t_cur], c.t_bfecc, c.u_state, st);
        }
        const int local = c.post_collision ? 0 : 1;
        launch_exact_step(P, C, c.f32[c.f_cur], c.f32[1 - c.f_cur], c.T[c.t_cur], c.T[1 - c.t_cur], c.u_state, local,
                          st);
        c.f_cur ^= 1;
        if (P.thermal_on) c.t_cur ^= 1;
        c.post_collision = true;
        c.force_valid = false;
        if (P.cylinder) {
            const size_t blocks = (c.cells + 255) / 256;
            if (!c.force_part && !c.alloc((void**)&c.force_part, (3 * blocks + 2) * sizeof(double))) return false;
            launch_exact_force(P, C, c.f32[c.f_cur], c.force_part, st);
            c.force_valid = true;
        }
        return hip_ok(hipGetLastError(), "exact step");
    }
    const FastStorage fk = fast_kind(c.storage);
    if (!ensure_face_store(c, P)) return false;
    if (!c.initialized) {
        void* f0 = c.storage == Storage::F32Fast ? (void*)c.f32[c.f_cur] : aa_buffer(c);
        launch_fast_init(P, C, fk, f0, c.T[c.t_cur], c.T[1 - c.t_cur], c.u_state, rt().bench.initial_temperature, st);
        c.initialized = true;
        c.post_collision = false;
        c.aa_layout = 3;
    }
    if (c.storage == Storage::F32Fast) {
        const FastRare rare{c.rare_list, c.rare_count, c.rare_cap, c.rare_vals};
        launch_fast_step(P, C, fk, c.mask, c.f32[c.f_cur], c.f32[1 - c.f_cur], c.T[c.t_cur], c.T[1 - c.t_cur],
                         c.u_state, nullptr, nullptr, 0, c.post_collision ? 0 : 1, rare, st);
        c.f_cur ^= 1;
        c.post_collision = true;
    } else {
        const bool even = c.aa_layout == 3 || c.aa_layout == 2;
        const int mode = even ? 1 : 2;
        const FastRare rare{c.rare_list, c.rare_count, c.rare_cap, c.rare_vals};
        launch_fast_step(P, C, fk, c.mask, aa_buffer(c), aa_buffer(c), c.T[c.t_cur], c.T[1 - c.t_cur], c.u_state,
                         c.face_store[c.store_cur], c.face_store[1 - c.store_cur], mode, c.aa_layout == 3 ? 1 : 0,
                         rare, st);
        c.store_cur ^= 1;
        c.aa_layout = even ? 1 : 2;
        c.post_collision = true;
    }
    if (P.thermal_on) c.t_cur ^= 1;
    return hip_ok(hipGetLastError(), "fast step");
}

bool storage_readout(Context& c, const StepParams& P, int stride, int out_ch, float scale, float* out_dev) {
    hipStream_t st = rt().stream;
    const CellFields C = cell_fields(c);
    switch (c.storage) {
        case Storage::F32Exact:
            launch_exact_readout(P, C, c.f32[c.f_cur], c.post_collision ? 1 : 0, stride, out_ch, scale, out_dev, st);
            break;
        case Storage::F32Fast:
            launch_fast_readout(P, C, FastStorage::F32, c.mask, c.f32[c.f_cur], nullptr, c.post_collision ? 0 : 3,
                                stride, out_ch, scale, out_dev, st);
            break;
        case Storage::F16:
        case Storage::BF16:
        case Storage::S16:
        case Storage::F32AA:
            launch_fast_readout(P, C, fast_kind(c.storage), c.mask, aa_buffer(c), c.face_store[c.store_cur],
                                c.aa_layout, stride, out_ch, scale, out_dev, st);
            break;
    }
    return hip_ok(hipGetLastError(), "readout");
}

bool storage_sample(Context& c, const StepParams& P, int x, int y, int z, float* out_dev4) {
    hipStream_t st = rt().stream;
    const CellFields C = cell_fields(c);
    if (c.storage == Storage::F32Exact) {
        launch_exact_sample(P, C, c.f32[c.f_cur], c.post_collision ? 1 : 0, x, y, z, out_dev4, st);
        return hip_ok(hipGetLastError(), "sample");
    }
    // the fast paths read the whole field once and copy the requested cell
    if (!storage_readout(c, P, 1, 4, 1.0f, c.out_dev)) return false;
    const size_t cell = ((size_t)x * c.ny + y) * c.nz + z;
    return hip_ok(hipMemcpyAsync(out_dev4, c.out_dev + cell * 4, 16, hipMemcpyDeviceToDevice, st), "sample copy");
}

}  // namespace aero
