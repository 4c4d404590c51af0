"""The mesoscale workload at full size on the GPU: D3Q19, 256 x 64 (vertical) x 256
cells over value-noise terrain (heights up to 24 cells), geostrophic wind 0.04 lattice
units rotated 30 degrees from +x with a log-law profile, lattice viscosity 5e-4, 500
steps (workloads.meso_forcing / meso_config).

The GPU step follows the reference CPU step's operation order
(native/src/aero_lbm_mesoscale.cpp:548-738), so its state must equal the FP32
oracle's bit for bit: the fixture (tests/golden/make_golden.py config2_meso) holds the
SHA-256 of the oracle's 6-channel state after steps 1, 100 and 500. The FP64 oracle's
probe values give the error of the FP32 arithmetic itself, reported beside it."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config2_meso.json")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_config2_mesoscale_500_steps_bit_exact():
    with open(FIXTURE) as fh:
        fx = json.load(fh)
    dims = tuple(fx["dims"])
    cfg = W.meso_config(*dims)
    F = W.meso_forcing(*dims)
    ctx = A.Mesoscale(cfg)
    idx = np.array(fx["probe_cells"])
    try:
        done = 0
        for c in fx["checkpoints"]:
            log(f"config2 mesoscale {dims} to step {c}")
            if done == 0:
                state = ctx.step(F)
                done = 1
            if c > done:
                ctx.run_timed(F, c - done)
                done = c
                state = ctx.state()
            want = fx[f"fp32_step{c}"]
            got = hashlib.sha256(np.ascontiguousarray(state, np.float32).tobytes()).hexdigest()
            s6 = state.reshape(-1, 6)
            if got != want["state_sha256"]:
                ref = np.array(want["probe_state"], np.float32)
                diff = float(np.max(np.abs(s6[idx] - ref)))
                pytest.fail(f"state after step {c} differs from the FP32 oracle (max probe difference {diff:g})")
            assert np.array_equal(s6[idx], np.array(want["probe_state"], np.float32)), c
        last = fx["checkpoints"][-1]
        ref64 = np.array(fx[f"fp64_step{last}"]["probe_state"])
        err = np.max(np.abs(s6[idx].astype(np.float64) - ref64), axis=0)
        log(f"config2 FP32 GPU state vs FP64 oracle after {last} steps, max |diff| per channel at "
            f"{len(idx)} probes: {np.array2string(err, precision=3)}")
        assert np.all(np.isfinite(s6))
    finally:
        ctx.release()
