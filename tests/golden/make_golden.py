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


JOBS = {"config1": config1, "config3_small": config3_small, "config3_full": config3_full}

if __name__ == "__main__":
    names = sys.argv[1:] or ["config1", "config3_small"]
    for name in names:
        res = JOBS[name]()
        with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
            if name == "config3_full":
                json.dump(res, fh, separators=(",", ":"))
            else:
                json.dump(res, fh, indent=0)
        print("wrote", name, flush=True)
