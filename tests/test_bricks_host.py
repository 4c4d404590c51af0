"""Host-side brick runtime through the C ABI (no GPU work): service and world
lifecycle, active hints with and without the neighbour closure, retention and
pruning, runtime status, world deltas, static constraints on uploaded dynamic
state and the packed atlas of stored state, against bricks_oracle and closed forms.
"""
import ctypes

import numpy as np
import pytest

import aero_lbm as A
import bricks_oracle as B


@pytest.fixture()
def world():
    w = A.BrickWorld(world=5, size=8, dx=0.5, dt=0.05, mode=A.SOLVER_D3Q27_FP16_INPLACE)
    yield w
    w.release()


def test_hint_closure_and_exact_hints(world):
    world.hints([(0, 0, 0), (2, 0, 0)], exact=False)
    st = world.status()
    # two hinted bricks plus their 6-neighbour closure: 2 + 12 - 1 shared = 13 bricks
    assert st[0] == 8 and st[1] == 13 and st[2] == 2 and st[3] == 0
    world.hints([(0, 0, 0), (2, 0, 0)], exact=True)
    st = world.status()
    assert st[1] == 2 and st[2] == 2 and st[3] == 0   # no static data yet: not active
    assert world.bricks(active=False) == [(0, 0, 0), (2, 0, 0)]


def test_static_upload_activates_and_marks_reinit(world):
    world.hints([(1, 2, 3)], exact=True)
    world.upload_static((1, 2, 3), np.zeros(512, np.uint8))
    st = world.status()
    assert st[3] == 1 and st[4] == 1 and st[6] == 1   # active, geometry dirty, pending reinit
    assert world.bricks() == [(1, 2, 3)]


def test_retention_prunes_after_16_idle_epochs(world):
    world.hints([(0, 0, 0)], exact=False)
    world.upload_static((0, 0, 0), np.zeros(512, np.uint8))
    world.hints([(9, 9, 9)], exact=False)
    # the first brick keeps its dirty flags until a step, so it stays; a hinted-only
    # closure brick with nothing else goes once 16 epochs have passed
    assert (1, 0, 0) not in world.bricks(active=False)
    before = world.status()[1]
    for _ in range(17):
        world.hints([(9, 9, 9)], exact=False)
    after = world.status()
    assert after[1] < before
    assert (0, 0, 0) in world.bricks(active=False)


def test_dynamic_upload_applies_static_constraints_and_atlas(world):
    n = 8
    rng = np.random.default_rng(4)
    ob = (rng.random(n ** 3) < 0.2).astype(np.uint8)
    kind = np.zeros(n ** 3, np.uint8)
    kind[5] = 3
    world.hints([(0, 0, 0)], exact=True)
    world.upload_static((0, 0, 0), ob, surface_kind=kind)
    flow = rng.uniform(-0.05, 0.05, n ** 3 * 4).astype(np.float32)
    air = rng.uniform(-0.5, 0.5, n ** 3).astype(np.float32)
    surf = rng.uniform(0, 1, n ** 3).astype(np.float32)
    world.upload_dynamic((0, 0, 0), flow, air, surf)
    got_f, got_a, got_s = world.dynamic((0, 0, 0))
    ef, ea, es = flow.copy(), air.copy(), surf.copy()
    B.apply_constraints(ob, np.zeros(n ** 3, np.uint16), kind, np.zeros(n ** 3, np.float32), ef, ea, es)
    assert np.array_equal(got_f, ef) and np.array_equal(got_a, ea) and np.array_equal(got_s, es)
    assert got_s[5] == surf[5] or ob[5]           # a surface kind keeps the surface temperature
    for stride in (1, 2, 3, 4):
        atlas = world.atlas((0, 0, 0), stride)
        assert np.array_equal(atlas, B.packed_atlas(n, stride, ef, velocity_scale=0.5 / 0.05)), stride


def test_world_deltas(world):
    n = 8
    world.hints([(0, 0, 0), (1, 0, 0)], exact=True)
    for c in ((0, 0, 0), (1, 0, 0)):
        world.upload_static(c, np.zeros(n ** 3, np.uint8))
    lib = world.lib
    # type 2 at a cell of brick (1,0,0): that brick and its existing neighbours need a re-init
    d = A.WorldDelta(type=2, x=9, y=3, z=3, data0=5)
    assert lib.aero_lbm_simulation_submit_world_deltas(world.service, ctypes.byref(d), 1) == 1
    # deltas wait for the next step of their world; a world key that does not exist is ignored
    d2 = A.WorldDelta(type=4, x=0, y=0, z=0, data0=99)
    assert lib.aero_lbm_simulation_submit_world_deltas(world.service, ctypes.byref(d2), 1) == 1
    assert lib.aero_lbm_simulation_submit_world_deltas(world.service, None, -1) == 0
    assert b"negative" in lib.aero_lbm_simulation_last_error()


def test_error_paths(world):
    lib = world.lib
    assert lib.aero_lbm_simulation_step_brick_world_runtime(world.service, 12345, 1) == 0
    assert b"missing brick runtime" in lib.aero_lbm_simulation_last_error()
    assert lib.aero_lbm_simulation_step_brick_world_runtime(world.service, 5, 0) == 0
    assert lib.aero_lbm_simulation_get_brick_world_active_brick_count(77, 5) == -1
    out = np.zeros(8, np.int32)
    assert lib.aero_lbm_simulation_get_brick_world_runtime_status(world.service, 5, out.ctypes.data_as(
        ctypes.POINTER(ctypes.c_int)), 4) == 0
    ob = np.zeros(27, np.uint8)
    with pytest.raises(A.AeroLbmError, match="size mismatch"):
        world.size = 3
        world.upload_static((0, 0, 0), ob)
    world.size = 8
    assert lib.aero_lbm_simulation_ensure_brick_world_runtime(world.service, 0, 8, 1.0, 0.05) == 0
    assert lib.aero_lbm_simulation_ensure_brick_world_runtime(world.service, 3, 8, -1.0, 0.05) == 0
    assert world.bricks() == []
