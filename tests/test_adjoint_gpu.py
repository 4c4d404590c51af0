"""Adjoint of the brick step (aero_lbm_adjoint_*) on the GPU.

* The FP64 GPU forward equals the FP64 dual-number model on the CPU (same
  function, 1e-12 relative); the FP32 forward is within 1e-5 of it.
* The FP32 adjoint gradient, contracted with random directions, equals the
  FP64 dual-number directional derivative (1e-4 relative).
* On the 64^3 fit (200 steps, checkpoints every 20, 64x64x3 inflow + 16^3 heat,
  512 probes, target from a hidden parameter set), the gradient matches FP64
  central differences on 32 random unit directions (1e-4 relative).
* The checkpoint interval changes nothing but rounding of the atomic scatter.
* Gradient descent with these gradients fits the inflow and heat.
"""
import sys

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

NU, BETA, ALPHA = 5e-3, 0.1, 0.035
U_IN = 0.04


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def model(case, steps, every, target):
    n = case["n"]
    return A.Adjoint(n, case["heat_grid"], steps, every, NU, BETA, ALPHA, (U_IN, 0.0, 0.0), case["solid"],
                     case["probes"], target)


def truth_target(case, steps):
    m = model(case, steps, steps, np.zeros(3 * len(case["probes"]), np.float32))
    try:
        _, u = m.forward(case["inflow_truth"], case["heat_truth"])
    finally:
        m.release()
    return u.reshape(-1).astype(np.float32)


def small_case():
    case = W.adjoint_case(16, 4, 40, fill=0.06, seed=5)
    rng = np.random.default_rng(3)
    inflow = (case["inflow_start"] + rng.uniform(-0.004, 0.004, case["inflow_start"].size)).astype(np.float32)
    heat = rng.uniform(-0.003, 0.003, case["heat_start"].size).astype(np.float32)
    return case, inflow, heat


def dual(oracle_mod, case, steps, target, inflow, heat, di, dh):
    return oracle_mod.dual_directional(case["n"], case["heat_grid"], steps, np.float64(np.float32(NU)),
                                       np.float32(BETA), np.float32(ALPHA), (U_IN, 0.0, 0.0), case["solid"],
                                       case["probes"], target.astype(np.float64), inflow.astype(np.float64),
                                       heat.astype(np.float64), di, dh)


def test_forward_matches_dual_model(oracle_mod):
    case, inflow, heat = small_case()
    steps = 40
    target = np.random.default_rng(4).uniform(-0.01, 0.05, 3 * len(case["probes"])).astype(np.float32)
    log("adjoint forward 16^3 x 40 steps vs FP64 dual model")
    m = model(case, steps, 7, target)
    try:
        l32, _ = m.forward(inflow, heat)
        l64 = m.forward_fp64(inflow.astype(np.float64), heat.astype(np.float64))
    finally:
        m.release()
    ld, _ = dual(oracle_mod, case, steps, target, inflow, heat, np.zeros(inflow.size), np.zeros(heat.size))
    log(f"  loss fp32 {l32!r} fp64 {l64!r} dual {ld!r}")
    assert abs(l64 - ld) <= 1e-12 * ld
    assert abs(l32 - ld) <= 1e-5 * ld


def test_gradient_matches_dual_directional(oracle_mod):
    case, inflow, heat = small_case()
    steps = 40
    target = np.random.default_rng(6).uniform(-0.01, 0.05, 3 * len(case["probes"])).astype(np.float32)
    m = model(case, steps, 7, target)
    try:
        m.forward(inflow, heat)
        gi, gh = m.backward()
    finally:
        m.release()
    rng = np.random.default_rng(7)
    for k in range(4):
        di = rng.standard_normal(inflow.size)
        dh = rng.standard_normal(heat.size)
        _, dl = dual(oracle_mod, case, steps, target, inflow, heat, di, dh)
        ga = float(np.dot(gi.astype(np.float64), di) + np.dot(gh.astype(np.float64), dh))
        log(f"  direction {k}: adjoint {ga!r} dual {dl!r} rel {abs(ga - dl) / abs(dl):.2e}")
        assert abs(ga - dl) <= 1e-4 * abs(dl)


def test_checkpoint_interval_only_changes_rounding():
    case, inflow, heat = small_case()
    steps = 40
    target = np.zeros(3 * len(case["probes"]), np.float32)
    grads = []
    for every in (1, 7, 40):
        m = model(case, steps, every, target)
        try:
            m.forward(inflow, heat)
            grads.append(m.backward())
        finally:
            m.release()
    for gi, gh in grads[1:]:
        assert np.allclose(gi, grads[0][0], rtol=1e-5, atol=1e-6 * np.abs(grads[0][0]).max())
        assert np.allclose(gh, grads[0][1], rtol=1e-5, atol=1e-6 * np.abs(grads[0][1]).max())


def test_fit_gradient_matches_central_differences():
    """The 64^3 fit: 32 random unit directions over the 12288 + 4096 parameters."""
    case = W.adjoint_case(64, 16, 512)
    steps, every = 200, 20
    log("adjoint 64^3 x 200 steps: target from the hidden parameters")
    target = truth_target(case, steps)
    m = model(case, steps, every, target)
    try:
        inflow, heat = case["inflow_start"], case["heat_start"]
        loss, _ = m.forward(inflow, heat)
        gi, gh = m.backward()
        g = np.concatenate([gi, gh]).astype(np.float64)
        p = np.concatenate([inflow, heat]).astype(np.float64)
        ni = inflow.size
        rng = np.random.default_rng(11)
        eps = 1e-5
        rel = []
        for k in range(32):
            d = rng.standard_normal(p.size)
            d /= np.linalg.norm(d)
            lp = m.forward_fp64(p[:ni] + eps * d[:ni], p[ni:] + eps * d[ni:])
            lm = m.forward_fp64(p[:ni] - eps * d[:ni], p[ni:] - eps * d[ni:])
            fd = (lp - lm) / (2 * eps)
            ga = float(np.dot(g, d))
            rel.append(abs(ga - fd) / abs(fd))
            log(f"  direction {k}: adjoint {ga!r} fd {fd!r} rel {rel[-1]:.2e}")
    finally:
        m.release()
    log(f"  loss {loss!r}, max rel {max(rel):.2e}, median {np.median(rel):.2e}")
    assert max(rel) <= 1e-4


def test_fit_reduces_loss():
    case = W.adjoint_case(32, 8, 128, seed=9)
    steps, every = 100, 10
    target = truth_target(case, steps)
    m = model(case, steps, every, target)
    try:
        p = np.concatenate([case["inflow_start"], case["heat_start"]]).astype(np.float64)
        ni = case["inflow_start"].size
        mom, var = np.zeros_like(p), np.zeros_like(p)
        lr, b1, b2 = 1e-3, 0.9, 0.999
        losses = []
        for it in range(1, 21):
            loss, _ = m.forward(p[:ni].astype(np.float32), p[ni:].astype(np.float32))
            gi, gh = m.backward()
            losses.append(loss)
            g = np.concatenate([gi, gh]).astype(np.float64)
            mom = b1 * mom + (1 - b1) * g
            var = b2 * var + (1 - b2) * g * g
            p -= lr * (mom / (1 - b1 ** it)) / (np.sqrt(var / (1 - b2 ** it)) + 1e-12)
        log("  fit losses " + " ".join(f"{v:.3e}" for v in losses))
    finally:
        m.release()
    assert losses[-1] < 0.2 * losses[0]


def test_error_paths():
    case, inflow, heat = small_case()
    m = model(case, 8, 4, np.zeros(3 * len(case["probes"]), np.float32))
    try:
        with pytest.raises(A.AeroLbmError, match="forward first"):
            m.backward()
    finally:
        m.release()
    bad = dict(case)
    bad["probes"] = np.array([16 ** 3], np.int32)
    with pytest.raises(A.AeroLbmError, match="out of range"):
        model(bad, 8, 4, np.zeros(3, np.float32))
