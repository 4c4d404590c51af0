"""Writes the oracle fixtures under tests/golden/ (run in the build container).

Long oracle runs are too slow to repeat inside a GPU test, so their results are
stored as fixtures: SHA-256 of the FP32 output bytes (the exact GPU path must
reproduce them bit for bit), 512 seeded probe cells, per-x-plane sums, and for
accuracy checks the FP64 oracle's velocity field at the probes.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py config1    # one
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "aerodynamics4mc-fabric_amd"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import pyoracle  # noqa: E402
import workloads as W  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def probes(cells: int, count: int = 512, seed: int = 99) -> list:
    return sorted(set(int(v) for v in (W.splitmix_array(seed, count) * cells).astype(np.int64)))


def summarize(flow: np.ndarray, temp: np.ndarray, nx: int, idx: list) -> dict:
    f4 = flow.reshape(-1, 4)
    planes = f4.reshape(nx, -1, 4).astype(np.float64).sum(axis=1)
    return {
        "flow_sha256": digest(flow),
        "temperature_sha256": digest(temp),
        "probe_cells": idx,
        "probe_flow": f4[idx].astype(float).tolist(),
        "probe_temperature": temp[idx].astype(float).tolist(),
        "plane_sums": planes.tolist(),
    }


def config1(steps=1000, n=32):
    cfg, pk, solid = W.config1(n)
    out = {"name": "config1", "n": n, "steps": steps,
           "desc": "D3Q27 cumulant 32^3, 15% solid boxes (seed 2), u=(0.05,0,0)+-1e-3 (seed 1), "
                   "nu 1e-3 (Re 50), reference SGS C=0.025, buoyancy off, velocity inlet x-, convective x+, "
                   "periodic y/z"}
    idx = probes(n ** 3)
    for prec in (32, 64):
        t0 = time.time()
        o = pyoracle.Oracle(n, n, n, cfg, precision=prec)
        o.run(pk, steps)
        out[f"fp{prec}"] = summarize(o.flow(), o.temperature(), n, idx)
        if prec == 64:
            u = o.velocity()
            out["fp64"]["probe_velocity_f64"] = u[idx].tolist()
        out[f"fp{prec}"]["seconds"] = time.time() - t0
        print(f"config1 fp{prec}: {time.time() - t0:.1f} s", flush=True)
    return out


def config1_horizons(horizons=(50, 100, 200, 400), n=32):
    """Config 1 (32^3) through the FP32 and FP64 oracles, probes at short horizons: the
    flow (nu 1e-3, tau ~0.503) amplifies every rounding difference, so after 1000
    steps the FP32 oracle is ~2e-2 relative away from the FP64 one and only a
    bit-exact path meets a 1e-5 relative bar there; at these horizons the two
    oracles still agree, and a faster FP32 kernel is held to 1e-5 relative of the
    FP32 oracle."""
    cfg, pk, solid = W.config1(n)
    out = {"name": "config1_horizons", "n": n, "horizons": list(horizons),
           "desc": "config 1 workload; FP32 and FP64 oracle probes (density and velocity) at each horizon"}
    idx = probes(n ** 3)
    out["probe_cells"] = idx
    for prec in (32, 64):
        o = pyoracle.Oracle(n, n, n, cfg, precision=prec)
        done = 0
        for h in horizons:
            if done == 0:
                o.run(pk, h)
            else:
                for _ in range(h - done):
                    o.step(None)
            done = h
            out[f"fp{prec}_{h}"] = o.flow().reshape(-1, 4)[idx].astype(float).tolist()
    for h in horizons:
        a, b = np.array(out[f"fp32_{h}"]), np.array(out[f"fp64_{h}"])
        out[f"fp32_vs_fp64_rel_{h}"] = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)))
        print(f"config1 horizon {h}: FP32 oracle vs FP64 {out[f'fp32_vs_fp64_rel_{h}']:.3e}", flush=True)
    return out


def config3_small(steps=300, n=32):
    cfg, ov, pk, t, solid = W.config3(n, beta=0.1, cs=0.1)
    out = {"name": "config3_small", "n": n, "steps": steps,
           "desc": "D3Q27 cumulant + SGS (C=0.1) + Boussinesq (beta 0.1, gravity -z) at 32^3: ~10% solid "
                   "boxes (seed 5), Gaussian hot blob dT 0.02 sigma 8 on the z=0 floor (seed 4), inflow 0.05"}
    idx = probes(n ** 3)
    for prec in (32, 64):
        t0 = time.time()
        o = pyoracle.Oracle(n, n, n, cfg, overrides=ov, precision=prec)
        o.set_temperature(t)
        o.run(pk, steps)
        out[f"fp{prec}"] = summarize(o.flow(), o.temperature(), n, idx)
        out[f"fp{prec}"]["seconds"] = time.time() - t0
        print(f"config3_small fp{prec}: {time.time() - t0:.1f} s", flush=True)
    return out


def config3_full(steps=1000, n=128):
    """FP64 oracle only (about 15 min on 8 cores): the reference for the FP16 / BF16 /
    FP32 accuracy of the 128^3 workload after 1000 steps."""
    cfg, ov, pk, t, solid = W.config3(n, beta=0.1, cs=0.1)
    out = {"name": "config3_full", "n": n, "steps": steps,
           "desc": "D3Q27 cumulant + SGS (C=0.1) + Boussinesq (beta 0.1, gravity -z) at 128^3: ~10% solid "
                   "boxes (seed 5), Gaussian hot blob dT 0.02 sigma 8 on the z=0 floor (seed 4), inflow 0.05; "
                   "FP64 oracle"}
    idx = probes(n ** 3, count=4096, seed=7)
    t0 = time.time()
    o = pyoracle.Oracle(n, n, n, cfg, overrides=ov, precision=64)
    o.set_temperature(t)
    o.run(pk, steps)
    full = summarize(o.flow(), o.temperature(), n, idx)
    out["fp64"] = {"probe_cells": idx,
                   "probe_flow": full["probe_flow"],  # output channels (u_x, u_y, u_z, rho - 1)
                   "probe_temperature": [float("%.9g" % v) for v in full["probe_temperature"]],
                   "plane_sums": [[float("%.12g" % v) for v in row] for row in full["plane_sums"]],
                   "flow_sha256": full["flow_sha256"], "seconds": time.time() - t0}
    print(f"config3_full fp64: {time.time() - t0:.1f} s", flush=True)
    # the FP32 oracle's own distance from FP64 at the probes (another ~15 min)
    o32 = pyoracle.Oracle(n, n, n, cfg, overrides=ov, precision=32)
    o32.set_temperature(t)
    o32.run(pk, steps)
    ref = np.array(full["probe_flow"])[:, :3]
    got = o32.flow().reshape(-1, 4)[idx, :3].astype(np.float64)
    fluid = solid.reshape(-1)[idx] == 0
    out["fp32_oracle_err"] = float(np.max(np.abs(got[fluid] - ref[fluid]))) / 0.05
    out["fp32_oracle_note"] = "max |u_fp32_oracle - u_fp64_oracle| / U_in over the fluid probes after 1000 steps"
    return out


def config3_horizons(horizons=(100, 200, 300), n=128):
    """FP64 and FP32 oracle of the 128^3 workload at short horizons (about 25 min on 8
    cores): while two FP32 evaluations still agree to ~1e-4 of the inflow speed, the
    16-bit storage formats are held to the 1e-3 bar against the FP64 values here."""
    cfg, ov, pk, t, solid = W.config3(n, beta=0.1, cs=0.1)
    out = {"name": "config3_horizons", "n": n, "horizons": list(horizons),
           "desc": "D3Q27 cumulant + SGS (C=0.1) + Boussinesq (beta 0.1, gravity -z) at 128^3: ~10% solid "
                   "boxes (seed 5), Gaussian hot blob dT 0.02 sigma 8 on the z=0 floor (seed 4), inflow 0.05; "
                   "FP64 oracle probes at each horizon, FP32 oracle distance from them"}
    idx = probes(n ** 3, count=4096, seed=7)
    out["probe_cells"] = idx
    fluid = solid.reshape(-1)[idx] == 0
    runs = {}
    # per-precision results are cached between invocations (a long run that is
    # interrupted resumes from the finished precision)
    cache_dir = os.environ.get("AERO_GOLDEN_CACHE", "/tmp/aero_golden_cache")
    os.makedirs(cache_dir, exist_ok=True)
    for prec in (64, 32):
        cache = os.path.join(cache_dir, f"config3_horizons_fp{prec}_{'_'.join(map(str, horizons))}.npz")
        if os.path.exists(cache):
            z = np.load(cache)
            for h in horizons:
                runs[(prec, h)] = (z[f"f{h}"], z[f"t{h}"])
            continue
        t0 = time.time()
        o = pyoracle.Oracle(n, n, n, cfg, overrides=ov, precision=prec)
        o.set_temperature(t)
        done = 0
        for h in horizons:
            if done == 0:
                o.run(pk, h)
            else:
                for _ in range(h - done):
                    o.step(None)
            done = h
            runs[(prec, h)] = (o.flow().reshape(-1, 4)[idx].astype(np.float64), o.temperature()[idx].astype(np.float64))
            print(f"config3_horizons fp{prec} step {h}: {time.time() - t0:.1f} s", flush=True)
        np.savez(cache, **{f"f{h}": runs[(prec, h)][0] for h in horizons}, **{f"t{h}": runs[(prec, h)][1] for h in horizons})
    for h in horizons:
        f64, t64 = runs[(64, h)]
        f32, _ = runs[(32, h)]
        out[f"fp64_{h}"] = {"probe_flow": [[float("%.9g" % v) for v in row] for row in f64.tolist()],
                            "probe_temperature": [float("%.9g" % v) for v in t64.tolist()]}
        out[f"fp32_oracle_err_{h}"] = float(np.max(np.abs(f32[fluid, :3] - f64[fluid, :3]))) / 0.05
    return out


def config2_meso(checkpoints=(1, 100, 500), dims=(256, 64, 256)):
    """D3Q19 mesoscale 256 x 64 (vertical) x 256 over value-noise terrain (heights up
    to 24 cells), geostrophic wind 0.04 lattice units rotated 30 degrees from +x with a
    log-law profile, lattice viscosity 5e-4: SHA-256 of the FP32 oracle's 6-channel
    state after each checkpoint step, probe values, and the FP64 oracle's probes at
    the last checkpoint for the FP32 error report."""
    cfg = W.meso_config(*dims)
    F = W.meso_forcing(*dims)
    cells = int(np.prod(dims))
    idx = probes(cells)
    out = {"name": "config2_meso", "dims": list(dims), "checkpoints": list(checkpoints),
           "desc": "D3Q19 mesoscale 256x64x256 (y vertical), workloads.meso_forcing defaults (seed 3), "
                   "workloads.meso_config (nu_lattice 5e-4, dx 64 m, dt 5 s), tau_min 0.58"}
    for prec in (32, 64):
        t0 = time.time()
        o = pyoracle.MesoOracle(*dims, dx=cfg.dx_m, dt=cfg.dt_s, nu=cfg.molecular_nu_m2_s, pr=cfg.prandtl_air,
                                prt=cfg.turbulent_prandtl, precision=prec)
        done = 0
        for c in checkpoints:
            while done < c:
                state = o.step(F)
                done += 1
            if prec == 64 and c != checkpoints[-1]:
                continue
            s6 = state.reshape(-1, 6)
            key = f"fp{prec}_step{c}"
            out[key] = {"probe_state": s6[idx].astype(float).tolist()}
            if prec == 32:
                out[key]["state_sha256"] = digest(state)
            print(f"config2 fp{prec} step {c}: {time.time() - t0:.0f} s", flush=True)
    out["probe_cells"] = idx
    return out


def config4_batch(steps=500, probes_per_brick=64):
    """FP64 oracle of each of the 56 batched bricks (workloads.batch_brick_specs) after
    `steps` steps (about 50 min on 8 cores; per-brick results are cached, so an
    interrupted run resumes): velocity and temperature at seeded probe cells, for the
    error report of the batched FP16 run."""
    import aero_lbm as A
    ov = W.reference_overrides(A.KERNEL_FAST)
    cache_dir = os.environ.get("AERO_GOLDEN_CACHE", "/tmp/aero_golden_cache")
    os.makedirs(cache_dir, exist_ok=True)
    out = {"name": "config4_batch", "steps": steps,
           "desc": "40 x 32^3 + 12 x 64^3 + 4 x 128^3 bricks of workloads.batch_brick_specs (seed 2024), each "
                   "stepped alone by the FP64 oracle from its packet and hot blob; probe values after the "
                   "given steps", "bricks": []}
    t0 = time.time()
    for i, spec in enumerate(W.batch_brick_specs()):
        n = spec["edge"]
        idx = probes(n ** 3, count=probes_per_brick, seed=1000 + i)
        cache = os.path.join(cache_dir, f"config4_b{i}_s{steps}_p{probes_per_brick}.json")
        if os.path.exists(cache):
            with open(cache) as fh:
                out["bricks"].append(json.load(fh))
            continue
        cfg, pk, t, solid = W.batch_brick_case(spec)
        o = pyoracle.Oracle(n, n, n, cfg, overrides=ov, precision=64)
        o.set_temperature(t)
        o.run(pk, steps)
        rec = {"index": i, "edge": n, "speed": spec["speed"], "probe_cells": idx,
               "probe_solid": [int(v) for v in solid.reshape(-1)[idx]],
               "probe_flow": [[float("%.9g" % v) for v in row] for row in o.flow().reshape(-1, 4)[idx].tolist()],
               "probe_temperature": [float("%.9g" % v) for v in o.temperature()[idx].tolist()]}
        with open(cache, "w") as fh:
            json.dump(rec, fh)
        out["bricks"].append(rec)
        print(f"config4 brick {i} ({n}^3): {time.time() - t0:.0f} s", flush=True)
    return out


JOBS = {"config1": config1, "config1_horizons": config1_horizons, "config3_small": config3_small, "config3_full": config3_full,
        "config3_horizons": config3_horizons, "config2_meso": config2_meso, "config4_batch": config4_batch}

if __name__ == "__main__":
    names = sys.argv[1:] or ["config1", "config3_small"]
    for name in names:
        res = JOBS[name]()
        with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
            if name in ("config3_full", "config3_horizons", "config4_batch", "config2_meso"):
                json.dump(res, fh, separators=(",", ":"))
            else:
                json.dump(res, fh, indent=0)
        print("wrote", name, flush=True)
