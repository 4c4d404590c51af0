"""Throughput benchmark of the D3Q27 cumulant step on one MI355X.

Workload (one "step" = one lattice update of every cell of one brick): D3Q27
cumulant + Smagorinsky-type SGS (C = 0.1) + Boussinesq (beta 0.1, gravity -z)
on one 128^3 brick with ~10% solid boxes and a Gaussian hot blob on the floor,
inflow 0.05 -- the 128^3 fp16_inplace workload. The headline `value` is the
GLUPS of the FP16 equilibrium-shifted A-A path; the FP32 A-B path runs on the
same inputs for the FP16/FP32 ratio. Inputs are resident on the device before
timing; no readback inside the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--modes fp16,fp32]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "aerodynamics4mc-fabric_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "GLUPS + % HBM roofline, D3Q27 128³ FP16-shifted vs FP32 plain, 1 MI355X"

MODES = {
    # name: (solver mode, kernel variant, bytes per population)
    "fp16": ("fp16", 2),
    "bf16": ("bf16", 2),
    "s16": ("s16", 2),
    "fp32": ("fp32_fast", 4),
    "fp32_exact": ("fp32_exact", 4),
}


# Algorithmic bytes per lattice update: every population read once and written once, one flag byte per cell, the temperature scalar read and
# written once. The velocity carried to the next step's temperature advection
# (3 x 2 B read + written for 16-bit storage) is real traffic this figure leaves out.
B_LUP = {"fp16": 27 * 2 * 2 + 1 + 8, "bf16": 27 * 2 * 2 + 1 + 8, "s16": 27 * 2 * 2 + 1 + 8, "fp32": 27 * 4 * 2 + 1 + 8,
         "fp32_exact": 27 * 4 * 2 + 1 + 8}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "d3q27_step_128_traffic.json")


def algorithmic_bytes(mode: str, cells: int) -> dict:
    return {"bytes": cells * B_LUP[mode], "per_lup": B_LUP[mode]}


def profile_entry(mode: str, key: str):
    """A field of the committed rocprofv3 summary for this storage mode (HBM bytes per
    step launch from FETCH_SIZE / WRITE_SIZE scaled by the copy calibration recorded in
    the same file, the profiled kernel time), or None when the summary lacks it."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
        return t["modes"][mode][key]
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic(mode: str):
    return profile_entry(mode, "bytes_per_launch")


def set_mode(lib, A, W, mode: str, ov):
    ov.kernel_variant = A.KERNEL_EXACT if mode == "fp32_exact" else A.KERNEL_FAST
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(ov)), "set_solver_overrides")
    lib.aero_lbm_set_realtime_solver_mode({"fp16": A.SOLVER_D3Q27_FP16_INPLACE, "bf16": A.SOLVER_D3Q27_BF16_INPLACE,
                                           "s16": A.SOLVER_D3Q27_S16_INPLACE}.get(mode, A.SOLVER_CLASSIC_D3Q27))


def run_mode(lib, A, W, mode, n, steps, warmup, inputs):
    import torch
    cfg, ov, packet, temp, solid = inputs
    A.check(lib.aero_lbm_init_rect(n, n, n, 11, 4), "init_rect")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "benchmark_set_config")
    set_mode(lib, A, W, mode, ov)
    ctx = A.Context(n, n, n, key=100)
    ctx.set_temperature(temp)
    ctx.step(packet)  # upload + initialise (untimed)
    for _ in range(max(0, warmup - 1)):
        ctx.step(None)
    A.check(lib.aero_lbm_finish(), "finish")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    device_ms = ctx.run_timed(steps)  # HIP events on the library's stream around the step launches
    A.check(lib.aero_lbm_finish(), "finish")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    flow = ctx.flow().reshape(-1, 4)
    finite = bool(np.isfinite(flow).all())
    ctx.release()
    cells = n ** 3
    return {"mode": mode, "wall_s": wall, "device_ms": device_ms, "glups_wall": cells * steps / wall / 1e9,
            "glups_device": cells * steps / (device_ms / 1e3) / 1e9, "kernel_ms": device_ms / steps,
            "finite": finite, "max_speed": float(np.abs(flow[:, :3]).max())}


def host_cpu_facts() -> dict:
    """lscpu-style facts of this host: model name, logical CPUs, CPUs this process may
    use, and whether SMT is on."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    smt = None
    try:
        smt = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        pass
    allowed = sorted(os.sched_getaffinity(0))
    return {"model": model, "logical_cpus": os.cpu_count(), "allowed_cpus": len(allowed), "smt": smt}


def cpu_worker(n: int, warm: int, steps: int, cpu: int) -> None:
    """One CPU-baseline process: pinned to `cpu`, the FP32 oracle with one thread on
    its own context of the 128^3 workload; prints its rate as JSON. Touches no GPU."""
    os.sched_setaffinity(0, {cpu})
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    import workloads as W
    pyoracle.set_threads(1)  # the reference CPU step is single-threaded
    cfg, ov, packet, temp, solid = W.config3(n, beta=0.1, cs=0.1)
    o = pyoracle.Oracle(n, n, n, cfg, overrides=ov, precision=32)
    o.set_temperature(temp)
    o.run(packet, warm)  # warm-up steps (the first includes initialisation)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(None)
    dt = time.perf_counter() - t0
    print(json.dumps({"cpu": cpu, "steps": steps, "seconds": dt, "lups": n ** 3 * steps / dt}), flush=True)


def run_cpu_workers(n, warm, steps, cpus):
    """Independent worker processes, one per listed CPU, started together."""
    import subprocess
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               f"{n},{warm},{steps},{c}"], stdout=subprocess.PIPE, text=True) for c in cpus]
    out = []
    for p in procs:
        text, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"cpu worker failed with status {p.returncode}")
        out.append(json.loads(text.strip().splitlines()[-1]))
    return out


def cpu_baseline(n, sample_steps=8, agg_steps=3, max_procs=16):
    """The FP32 oracle (plain-C restatement of the reference's single-threaded CPU
    step) on this host's cores, on bounded samples of the same 128^3 workload: one
    process pinned to one core (the reference's own configuration), then the
    aggregate of independent contexts, one process per core, over the cores this
    process may use (at most max_procs)."""
    facts = host_cpu_facts()
    allowed = sorted(os.sched_getaffinity(0))
    one = run_cpu_workers(n, 2, sample_steps, allowed[:1])[0]
    k = min(max_procs, len(allowed))
    t0 = time.perf_counter()
    many = run_cpu_workers(n, 1, agg_steps, allowed[:k])
    wall = time.perf_counter() - t0
    agg = sum(w["lups"] for w in many)
    return {"value": one["lups"] / 1e9, "unit": "GLUPS", "cores": 1, "kind": "port",
            "sample": f"{sample_steps} steps of the same {n}^3 workload after 2 warm-up steps, FP32 oracle "
                      f"(oracle/liboracle.so), one thread pinned to CPU {allowed[0]} of a {facts['model']}; "
                      f"{one['seconds']:.1f} s",
            "all_cores": {"value": agg / 1e9, "unit": "GLUPS", "cores": k,
                          "sample": f"{k} independent processes, one pinned per CPU, each {agg_steps} steps of "
                                    f"its own {n}^3 context after 1 warm-up step; {wall:.1f} s wall"},
            "host": facts}


MESO_DIMS = (256, 64, 256)  # x, y (vertical), z
MESO_B_LUP = 2 * 19 * 4 * 2 + 1  # f and g read and written once, one flag byte


def run_meso(steps, warmup, no_cpu, tau_floor=0.501):
    """D3Q19 mesoscale step on 256 x 64 x 256 terrain (4-octave value noise, heights
    0-24 cells, geostrophic wind 0.04 lattice units at 30 degrees from +x with a log-law
    vertical profile, lattice viscosity 5e-4). The reference clamps both relaxation
    times at 0.58; the line runs the configured viscosity itself (tau 0.5015) through
    the library's tau-floor extension, tau_floor=None keeps the reference's clamp."""
    import aero_lbm as A
    import workloads as W
    lib = A.load()
    cfg = W.meso_config(*MESO_DIMS)
    F = W.meso_forcing(*MESO_DIMS)
    cells = int(np.prod(MESO_DIMS))
    A.check(lib.aero_lbm_mesoscale_set_tau_floor(float("nan") if tau_floor is None else tau_floor), "tau floor")
    tau = max(0.58 if tau_floor is None else tau_floor, 0.5 + 3.0 * 5e-4)
    ctx = A.Mesoscale(cfg)
    ctx.step(F)  # upload, coefficient planes, seeding (untimed)
    if warmup > 0:
        ctx.run_timed(F, warmup)
    t0 = time.perf_counter()
    device_ms = ctx.run_timed(F, steps)
    wall = time.perf_counter() - t0
    state = ctx.state().reshape(-1, 6)
    # the same step with every coefficient plane varying from cell to cell (the
    # workload's forcing leaves 13 of the 17 planes uniform, which the kernels take
    # as parameters instead of reading)
    Fv = W.meso_forcing_varying(*MESO_DIMS)
    ctx.step(Fv)
    ms_varying = ctx.run_timed(Fv, steps) / steps
    ctx.release()
    lib.aero_lbm_mesoscale_set_tau_floor(float("nan"))
    ms = device_ms / steps
    out = {"workload": "d3q19_mesoscale_256x64x256_fp32", "cells": cells, "steps": steps, "tau": round(tau, 5),
           "glups": cells * steps / wall / 1e9, "ms_per_step": wall * 1e3 / steps, "kernel_ms_per_step": ms,
           "finite": bool(np.isfinite(state).all()),
           "varying_planes": {"kernel_ms_per_step": ms_varying, "glups_device": cells / (ms_varying * 1e-3) / 1e9,
                              "forcing": "workloads.meso_forcing_varying: every coefficient plane varies"},
           "roofline": {"bound": "hbm", "achieved": cells * MESO_B_LUP / (ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": cells * MESO_B_LUP / (ms * 1e-3) / HBM_PEAK,
                        "traffic": measured_traffic("meso"), "bytes_per_lup": MESO_B_LUP}}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        pyoracle.set_threads(1)
        o = pyoracle.MesoOracle(*MESO_DIMS, dx=cfg.dx_m, dt=cfg.dt_s, nu=cfg.molecular_nu_m2_s, pr=cfg.prandtl_air,
                                prt=cfg.turbulent_prandtl, precision=32)
        o.step(F)
        n = 3
        t0 = time.perf_counter()
        for _ in range(n):
            o.step(F)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cells * n / dt / 1e9, "unit": "GLUPS", "cores": 1, "kind": "port",
                               "sample": f"{n} steps of the same 256x64x256 workload after 1 seeding step, FP32 "
                                         f"mesoscale oracle, single thread; {dt:.1f} s"}
    return out


def run_batched(steps, warmup, per_brick_steps=40):
    """40 x 32^3 + 12 x 64^3 + 4 x 128^3 FP16 bricks (workloads.batch_brick_specs), each
    with seeded solid boxes (5-20 % fill), its own inflow (speed 0.02-0.08, direction
    uniform in the x-y plane), Boussinesq buoyancy and a seeded hot blob; stepped
    together with the int16 atlas of every brick published every 4 steps (stride 4),
    the cycle replayed from a HIP graph. The same bricks stepped one launch sequence
    per brick (with the per-brick atlas read back and quantised on the host) give the
    per-brick rate."""
    import aero_lbm as A
    import workloads as W
    lib = A.load()
    A.check(lib.aero_lbm_init_rect(32, 32, 32, 11, 4), "init_rect")
    ov = W.reference_overrides(A.KERNEL_FAST)
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(ov)), "set_solver_overrides")
    keys, cells = [], 0
    key = 20000
    for spec in W.batch_brick_specs():
        edge = spec["edge"]
        cfg, pk, t, _ = W.batch_brick_case(spec)
        A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
        lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_D3Q27_FP16_INPLACE)
        ctx = A.Context(edge, edge, edge, key=key)
        ctx.set_temperature(t)
        ctx.step(pk)  # upload, initialise, first step (untimed)
        keys.append(key)
        cells += edge ** 3
        key += 1
    lib.aero_lbm_set_realtime_solver_mode(0)
    K = 4
    batch = A.Batch(keys, steps_per_publish=K, stride=4, scales=[20.0] * len(keys), autotune=True)
    info = batch.info()
    batch.run(max(1, warmup // K), graph=True)
    cycles = max(1, steps // K)
    t0 = time.perf_counter()
    device_ms = batch.run(cycles, graph=True)
    wall = time.perf_counter() - t0
    pb_cycles = max(1, per_brick_steps // K)
    pb_ms = batch.run_per_brick(pb_cycles)
    batch.release()
    for k in keys:
        lib.aero_lbm_release_context(k)
    nsteps = cycles * K
    ms = wall * 1e3 / nsteps
    b_lup = B_LUP["fp16"]  # FP16 populations, flag byte, temperature
    out = {"workload": "batched_bricks_40x32^3_12x64^3_4x128^3_fp16_graph_step_publish4", "bricks": len(keys),
           "cells": cells, "steps": nsteps, "glups": cells * nsteps / wall / 1e9, "ms_per_step": ms,
           "device_ms_per_step": device_ms / nsteps,
           "per_brick_glups": cells * pb_cycles * K / (pb_ms / 1e3) / 1e9, "per_brick_ms_per_step": pb_ms / (pb_cycles * K),
           "launch": info,
           "roofline": {"bound": "hbm", "achieved": cells * b_lup / (device_ms / nsteps * 1e-3) / 1e9,
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": cells * b_lup / (device_ms / nsteps * 1e-3) / HBM_PEAK,
                        "traffic": measured_traffic("batched"), "bytes_per_lup": b_lup}}
    out["speedup_vs_per_brick"] = out["glups"] / out["per_brick_glups"]
    return out


def world_run(edge, grid, steps, batching, publish_every=4, stride=4):
    """One brick world of grid^3 bricks of edge^3 cells in FP16 in-place mode (seeded
    solid boxes and inflow per brick), stepped through the drop-in C ABI one world step
    at a time with every brick's packed atlas copied every publish_every steps.
    Returns (wall seconds of the timed steps, cells)."""
    import aero_lbm as A
    import workloads as W
    rng = np.random.default_rng(7)
    w = A.BrickWorld(world=21, size=edge, mode=A.SOLVER_D3Q27_FP16_INPLACE)
    coords = [(x, y, z) for x in range(grid) for y in range(grid) for z in range(grid)]
    try:
        w.batching(batching)
        w.hints(coords, exact=True)
        for i, c in enumerate(coords):
            ob = W.solid_boxes(edge, edge, edge, 0.08, seed=100 + i, edge=(2, 6), keep_clear_x=2)
            u, a = rng.uniform(0.02, 0.06), rng.uniform(0, 2 * np.pi)
            flow = np.tile(np.array([u * np.cos(a), u * np.sin(a), 0.0, 0.0], np.float32), edge ** 3)
            w.upload_static(c, ob.reshape(-1))
            w.upload_dynamic(c, flow)
        w.step(2)  # packet steps: contexts start and become cached (untimed)
        t_pub = 0.0
        t0 = time.perf_counter()
        for k in range(steps):
            w.step(1)
            if (k + 1) % publish_every == 0:
                t1 = time.perf_counter()
                for c in coords:
                    w.atlas(c, stride)
                t_pub += time.perf_counter() - t1
        wall = time.perf_counter() - t0
        info = A.load().aero_lbm_simulation_runtime_info().decode()
        print(f"bench: world batching={batching}: {info}", file=sys.stderr, flush=True)
        return wall, len(coords) * edge ** 3, t_pub
    finally:
        w.release()


def run_world(steps=200, edge=32, grid=4):
    """The brick-world runtime as the mod drives it (aero_lbm_simulation_step_brick_world_runtime
    plus a packed-atlas copy per brick every 4 steps), 4 x 4 x 4 bricks of 32^3 FP16:
    cached bricks in one batched launch set per world step against the per-brick
    cached-step calls the reference's runtime makes."""
    wb, cells, pb = world_run(edge, grid, steps, True)
    sp = max(8, steps // 4)
    wp, _, pp = world_run(edge, grid, sp, False)
    out = {"workload": f"brick_world_{grid ** 3}x{edge}^3_fp16_step_publish4", "bricks": grid ** 3, "cells": cells,
           "steps": steps, "glups": cells * steps / wb / 1e9, "ms_per_step": wb * 1e3 / steps,
           "publish_ms_per_step": pb * 1e3 / steps,
           "per_brick_glups": cells * sp / wp / 1e9, "per_brick_ms_per_step": wp * 1e3 / sp,
           "per_brick_publish_ms_per_step": pp * 1e3 / sp}
    out["speedup_vs_per_brick"] = out["glups"] / out["per_brick_glups"]
    return out


def run_adjoint(reps=3):
    """Inflow/heat fit of one 64^3 brick (FP32): 200-step forward with checkpoints
    every 20 steps, backward over the 64x64x3 inflow profile and the 16^3 heat
    source, loss over 512 probes against velocities of a hidden parameter set.
    Timed: forward and backward at the starting guess (best of `reps`). Checked:
    the gradient against FP64 central differences on 32 random unit directions."""
    import aero_lbm as A
    import workloads as W
    case = W.adjoint_case(64, 16, 512)
    steps, every = 200, 20
    nu, beta, alpha = 5e-3, 0.1, 0.035
    mk = lambda tgt: A.Adjoint(64, 16, steps, every, nu, beta, alpha, (0.04, 0.0, 0.0), case["solid"],
                               case["probes"], tgt)
    m0 = mk(np.zeros(3 * 512, np.float32))
    _, u = m0.forward(case["inflow_truth"], case["heat_truth"])
    m0.release()
    m = mk(u.reshape(-1))
    inflow, heat = case["inflow_start"], case["heat_start"]
    fwd, bwd = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        loss, _ = m.forward(inflow, heat)
        t1 = time.perf_counter()
        gi, gh = m.backward()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
    g = np.concatenate([gi, gh]).astype(np.float64)
    p = np.concatenate([inflow, heat]).astype(np.float64)
    ni = inflow.size
    rng = np.random.default_rng(11)
    eps, rel = 1e-5, []
    for _ in range(32):
        d = rng.standard_normal(p.size)
        d /= np.linalg.norm(d)
        fd = (m.forward_fp64(p[:ni] + eps * d[:ni], p[ni:] + eps * d[ni:]) -
              m.forward_fp64(p[:ni] - eps * d[:ni], p[ni:] - eps * d[ni:])) / (2 * eps)
        rel.append(abs(float(np.dot(g, d)) - fd) / abs(fd))
    m.release()
    cells = 64 ** 3
    return {"workload": "adjoint_fit_64^3_200steps_ckpt20_inflow64x64x3_heat16^3_probes512_fp32",
            "parameters": int(p.size), "loss_at_start": loss,
            "forward_ms": min(fwd) * 1e3, "backward_ms": min(bwd) * 1e3,
            "gradient_ms": (min(fwd) + min(bwd)) * 1e3,
            "forward_glups": cells * steps / min(fwd) / 1e9,
            "backward_glups": cells * steps / min(bwd) / 1e9,
            "fd_check": {"directions": 32, "eps": eps, "max_rel": max(rel), "median_rel": float(np.median(rel)),
                         "reference": "fp64 forward, central differences"}}


CODEC_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "codec_zfp_traffic.json")


def codec_profile(key):
    """A field of the committed rocprofv3 summary of the codec's kernels, or None."""
    try:
        with open(CODEC_TRAFFIC_FILE) as fh:
            return json.load(fh)["grid_256"][key]
    except (OSError, KeyError, ValueError):
        return None


def run_codec(reps=50):
    """Analysis codec: zfp 1.0.1 fixed-accuracy streams of 3D float grids, byte for byte
    the reference's nativeCompressFloatGrid3d. Two workloads: one 256^3 residual
    velocity grid (a swirl of amplitude 0.05 plus Gaussian noise of 0.003, tolerance
    0.02) for throughput, and the analysis payload the server codes per region (four
    64^3 residual grids: vx, vy, vz at tolerance 0.02, pressure at 5e-4). Inputs are
    resident in device memory; device time per call from HIP events over `reps`
    back-to-back calls of the batched entry point; the host-memory entry point (the
    JNI path, PCIe both ways) timed beside it. CPU baseline: the reference codec built
    from its sources (oracle/_ref) when present, else the oracle, one thread."""
    import torch
    import aero_lbm as A
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import codec_cases as CC
    lib = A.load()

    def timed(grids, tols):
        dev = [torch.from_numpy(g).cuda() for g in grids]
        streams, ms = A.compress_device_grids(dev, tols, lib, repeats=reps)
        t0, k = time.perf_counter(), 0
        while time.perf_counter() - t0 < 0.5 or k < 3:
            for g, t in zip(grids, tols):
                A.compress_float_grid_3d(g, t, lib)
            k += 1
        return streams, ms, (time.perf_counter() - t0) / k * 1e3

    big = [CC.grid("velocity", (256, 256, 256), 3)]
    streams, ms, host_ms = timed(big, [CC.VELOCITY_TOL])
    values = big[0].size
    comp = len(streams[0])
    alg = values * 4 + comp  # read every value once, write the stream once
    traffic = codec_profile("bytes_per_call")
    out = {"workload": "zfp_fixed_accuracy_256^3_velocity_residual_tol0.02", "values": values,
           "compressed_bytes": comp, "bits_per_value": round(8 * comp / values, 3),
           "device_us_per_call": round(ms * 1e3, 2), "input_gbps": values * 4 / (ms * 1e-3) / 1e9,
           "host_call_ms": round(host_ms, 3),
           "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": alg / (ms * 1e-3) / HBM_PEAK, "traffic": traffic,
                        "algorithmic_bytes": alg,
                        "note": "encode + assembly launches per call; the encoder is integer-VALU bound"}}
    payload = [CC.grid(k, (64, 64, 64), 40 + i) for i, k in enumerate(["velocity", "velocity", "velocity", "pressure"])]
    ptols = [CC.VELOCITY_TOL] * 3 + [CC.PRESSURE_TOL]
    pstreams, pms, phost = timed(payload, ptols)
    out["payload_4x64^3"] = {"device_us_per_call": round(pms * 1e3, 2), "host_call_ms": round(phost, 3),
                             "compressed_bytes": sum(len(x) for x in pstreams)}
    ref = CC.reference()
    cpu, kind = (ref, "reference") if ref is not None else (CC.oracle(), "port")
    t0, k = time.perf_counter(), 0
    while time.perf_counter() - t0 < 2.0 or k < 2:
        for g, t in zip(payload, ptols):
            cpu.compress(g, t)
        k += 1
    cpu_ms = (time.perf_counter() - t0) / k * 1e3
    out["payload_4x64^3"]["cpu_ms"] = round(cpu_ms, 3)
    out["cpu_baseline"] = {"value": sum(g.size for g in payload) * 4 / (cpu_ms * 1e-3) / 1e9, "unit": "GB/s of input",
                           "cores": 1, "kind": kind,
                           "sample": f"{k} codings of the 4 x 64^3 payload on the host, "
                                     f"{'reference codec with vendored zfp' if kind == 'reference' else 'oracle'}, "
                                     f"{cpu_ms:.2f} ms each"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--modes", default="fp16,fp32,s16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-meso", action="store_true", help="skip the D3Q19 mesoscale line")
    ap.add_argument("--meso-steps", type=int, default=500)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-bricks line")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-fit line")
    ap.add_argument("--no-world", action="store_true", help="skip the brick-world runtime line")
    ap.add_argument("--no-codec", action="store_true", help="skip the analysis-codec line")
    ap.add_argument("--codec-only", action="store_true", help="only the analysis-codec line")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB (read 1 GiB + write 1 GiB) five times, "
                         "to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.cpu_worker:
        n_, warm_, steps_, cpu_ = (int(v) for v in args.cpu_worker.split(","))
        cpu_worker(n_, warm_, steps_, cpu_)
        return
    if args.calibrate_copy:
        import torch
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_each_way": src.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the solver has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.codec_only:
        print(json.dumps({"codec": run_codec()}), flush=True)
        return
    import aero_lbm as A
    import workloads as W
    lib = A.load()
    n = args.n
    inputs = W.config3(n, beta=0.1, cs=0.1)
    solid = inputs[4]
    modes = [m for m in args.modes.split(",") if m]
    results = {}
    for m in modes:
        print(f"bench: mode {m} n={n} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
        results[m] = run_mode(lib, A, W, m, n, args.steps, args.warmup, inputs)
        print(f"bench: {m}: {results[m]['glups_wall']:.2f} GLUPS (wall), kernel {results[m]['kernel_ms']:.4f} ms",
              file=sys.stderr, flush=True)
    head = modes[0]
    r = results[head]
    ab = algorithmic_bytes(head, n ** 3)
    achieved = ab["bytes"] / (r["kernel_ms"] / 1e3)
    solid_fraction = float(solid.mean())
    line = {
        "metric": METRIC,
        "value": round(r["glups_wall"], 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"d3q27_cumulant_sgs_boussinesq_{n}^3_{head}", "cells": n ** 3,
                   "storage": head, "solid_fraction": round(solid_fraction, 4),
                   "sgs_c": 0.1, "beta": 0.1, "gravity": "-z", "inflow": 0.05},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": measured_traffic(head),
                     "bytes_per_launch": ab["bytes"], "bytes_per_lup": ab["per_lup"],
                     "kernel_ms": r["kernel_ms"]},
        # the committed rocprofv3 kernel trace of the same workload (profiler-serialised
        # launches) beside the live HIP-event time that `frac` uses
        "roofline_profiled": {"kernel_us": profile_entry(head, "avg_kernel_us_profiled"),
                     "frac": (ab["bytes"] / (profile_entry(head, "avg_kernel_us_profiled") * 1e-6) / HBM_PEAK
                              if profile_entry(head, "avg_kernel_us_profiled") else None),
                     "source": "profiles/d3q27_step_128_kernel_stats.md"},
        "modes": {m: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in res.items()}
                  for m, res in results.items()},
    }
    if "fp32" in results and head != "fp32":
        line["fp16_over_fp32"] = round(results[head]["glups_wall"] / results["fp32"]["glups_wall"], 3)
    if not args.no_cpu:
        print("bench: cpu baseline (oracle, 1 core, then one process per core)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(n)
    if not args.no_meso:
        print(f"bench: mesoscale d3q19 {MESO_DIMS} steps={args.meso_steps}", file=sys.stderr, flush=True)
        line["mesoscale"] = run_meso(args.meso_steps, min(args.warmup, 20), args.no_cpu)
        print(f"bench: mesoscale {line['mesoscale']['glups']:.2f} GLUPS "
              f"(every plane varying: {line['mesoscale']['varying_planes']['glups_device']:.2f} GLUPS device)",
              file=sys.stderr, flush=True)
    if not args.no_batch:
        print(f"bench: batched bricks steps={args.batch_steps}", file=sys.stderr, flush=True)
        line["batched"] = run_batched(args.batch_steps, min(args.warmup, 20))
        print(f"bench: batched {line['batched']['glups']:.2f} GLUPS vs per-brick "
              f"{line['batched']['per_brick_glups']:.2f} ({line['batched']['speedup_vs_per_brick']:.2f}x); "
              f"{line['batched']['launch']}", file=sys.stderr, flush=True)
    if not args.no_world:
        print("bench: brick world 64 x 32^3 fp16", file=sys.stderr, flush=True)
        line["world"] = run_world()
        print(f"bench: world {line['world']['glups']:.2f} GLUPS batched vs {line['world']['per_brick_glups']:.2f} "
              f"per brick ({line['world']['speedup_vs_per_brick']:.2f}x)", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        print("bench: adjoint fit 64^3 x 200 steps", file=sys.stderr, flush=True)
        line["adjoint"] = run_adjoint()
        print(f"bench: adjoint forward {line['adjoint']['forward_ms']:.1f} ms, backward "
              f"{line['adjoint']['backward_ms']:.1f} ms, fd max rel {line['adjoint']['fd_check']['max_rel']:.2e}",
              file=sys.stderr, flush=True)
    if not args.no_codec:
        line["codec"] = run_codec()
        print(f"bench: codec {line['codec']['device_us_per_call']:.1f} us per 256^3 grid "
              f"({line['codec']['input_gbps']:.0f} GB/s of input), payload "
              f"{line['codec']['payload_4x64^3']['device_us_per_call']:.1f} us on the device vs "
              f"{line['codec']['payload_4x64^3']['cpu_ms']:.2f} ms on the host CPU", file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
