"""GPU brick runtime through the C ABI.

* A brick world steps each brick through this library's own context API: its
  first step builds the 11-channel packet, configures the compact boundary from
  the packet's mean velocity and steps with the packet; later steps are cached.
  The same sequence driven by hand from bricks_oracle's packet, mean velocity and
  seed rules must give bit-identical flow fields.
* The boundary-reference shell blend against the FP32 oracle with the same shell
  (FP32 in-place storage; the two collision codes differ by rounding only).
* The packed atlas of a live brick equals the oracle's quantisation of the
  context's own atlas extraction.
"""
import ctypes
import sys

import numpy as np
import pytest

import aero_lbm as A
import bricks_oracle as B
import workloads as W

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compact_config(v, n):
    """Velocity inlet x-, convective outflow x+, symmetry elsewhere (what the runtime sets)."""
    cfg = A.default_config()
    cfg.enabled = 1
    cfg.preset = 0
    cfg.reference_length = float(n)
    for k in range(3):
        cfg.initial_velocity[k] = v[k]
    for f in (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max):
        ctypes.memset(ctypes.addressof(f), 0, ctypes.sizeof(f))
        f.hydrodynamic_kind = A.HYDRO_SYMMETRY
        f.thermal_kind = 0
    cfg.x_min.hydrodynamic_kind = A.HYDRO_VELOCITY
    for k in range(3):
        cfg.x_min.velocity[k] = v[k]
    cfg.x_max.hydrodynamic_kind = A.HYDRO_CONVECTIVE
    return cfg


class Manual:
    """A standalone context driven the way the runtime drives a brick."""

    def __init__(self, lib, n, key, mode, packet):
        self.lib, self.n = lib, n
        self.v, ok = B.packet_mean_velocity(packet)
        assert ok
        lib.aero_lbm_set_realtime_solver_mode(mode)
        self.ctx = A.Context(n, n, n, key=key)
        self._cfg()
        self.ctx.step(packet)
        lib.aero_lbm_benchmark_reset_config()
        lib.aero_lbm_set_realtime_solver_mode(0)

    def _cfg(self):
        A.check(self.lib.aero_lbm_benchmark_set_config(ctypes.byref(compact_config(self.v, self.n))), "config")

    def step(self, k=1):
        for _ in range(k):
            self._cfg()
            self.ctx.step(None)
            self.lib.aero_lbm_benchmark_reset_config()


@pytest.fixture(scope="module")
def lib():
    lib = A.load()
    A.check(lib.aero_lbm_set_solver_overrides(ctypes.byref(W.reference_overrides(A.KERNEL_FAST))), "overrides")
    return lib


def brick_static(n, seed):
    ob = W.solid_boxes(n, n, n, 0.08, seed=seed, edge=(2, 4), keep_clear_x=2).reshape(-1).astype(np.uint8)
    return ob


def constrained(ob, flow, air):
    f, a, s = flow.copy(), air.copy(), np.zeros_like(air)
    B.apply_constraints(ob, np.zeros(ob.size, np.uint16), np.zeros(ob.size, np.uint8), np.zeros(ob.size, np.float32),
                        f, a, s)
    return f, a


@pytest.mark.parametrize("mode", [A.SOLVER_D3Q27_FP16_INPLACE, A.SOLVER_D3Q27_BF16_INPLACE])
def test_single_brick_matches_manual_driving(lib, mode):
    n = 16
    ob = brick_static(n, 3)
    rng = np.random.default_rng(1)
    flow = np.tile(np.array([0.04, 0.0, 0.0, 0.0], np.float32), n ** 3)
    flow[0::4] += rng.uniform(-1e-3, 1e-3, n ** 3).astype(np.float32)
    air = np.zeros(n ** 3, np.float32)
    w = A.BrickWorld(world=11, size=n, mode=mode)
    try:
        w.hints([(0, 0, 0)], exact=True)
        w.upload_static((0, 0, 0), ob)
        w.upload_dynamic((0, 0, 0), flow, air)
        log(f"world step 1 mode={mode}")
        w.step(1)
        f, a = constrained(ob, flow, air)
        pk = B.build_packet(n, ob, np.zeros(n ** 3, np.uint8), np.zeros(n ** 3, np.float32), f, a, [True] * 6)
        m = Manual(lib, n, 4001, mode, pk)
        got, _, _ = w.dynamic((0, 0, 0))
        assert np.array_equal(got, m.ctx.flow())
        w.step(6)
        m.step(6)
        got, air_got, _ = w.dynamic((0, 0, 0))
        assert np.array_equal(got, m.ctx.flow())
        assert np.array_equal(air_got, m.ctx.temperature())
        st = w.status()
        assert st[3] == 1 and st[4] == 0 and st[6] == 0 and st[7] >= 7
        # packed atlas of the live brick: quantised context atlas
        for stride in (1, 4):
            sn = (n + stride - 1) // stride
            raw = np.zeros(sn ** 3 * 4, np.float32)
            A.check(lib.aero_lbm_extract_flow_atlas_rect(n, n, n, m.ctx.key, stride, A.fptr(raw), raw.size), "atlas")
            q = np.empty((sn ** 3, 4), np.int16)
            r4 = raw.reshape(-1, 4)
            q[:, :3] = B.quantize((r4[:, :3] * np.float32(1.0 / 0.05)).astype(np.float32), 5.6)
            q[:, 3] = B.quantize(r4[:, 3], 0.03)
            assert np.array_equal(w.atlas((0, 0, 0), stride), q.reshape(-1)), stride
        m.ctx.release()
    finally:
        w.release()


def test_cell_patch_deltas_patch_the_cached_context(lib):
    """World deltas of type 8 (cell patches) on a cached in-place brick go through
    aero_lbm_patch_d3q27_f16_static_cells_rect and the brick steps on cached
    (simulation_bridge.cpp:1617-1704): the world equals a context driven by hand with
    the same patch, bit for bit, and its runtime info counts the patch."""
    n = 16
    mode = A.SOLVER_D3Q27_FP16_INPLACE
    ob = brick_static(n, 7)
    rng = np.random.default_rng(2)
    flow = np.tile(np.array([0.04, 0.0, 0.0, 0.0], np.float32), n ** 3)
    flow[0::4] += rng.uniform(-1e-3, 1e-3, n ** 3).astype(np.float32)
    air = np.zeros(n ** 3, np.float32)
    w = A.BrickWorld(world=12, size=n, mode=mode)
    try:
        w.hints([(0, 0, 0)], exact=True)
        w.upload_static((0, 0, 0), ob)
        w.upload_dynamic((0, 0, 0), flow, air)
        w.step(3)
        f, a = constrained(ob, flow, air)
        pk = B.build_packet(n, ob, np.zeros(n ** 3, np.uint8), np.zeros(n ** 3, np.float32), f, a, [True] * 6)
        m = Manual(lib, n, 4002, mode, pk)
        m.step(2)
        # a fluid cell turns solid, a solid cell turns fluid
        fluid = np.argwhere(ob.reshape(n, n, n) == 0)
        solid = np.argwhere(ob.reshape(n, n, n) != 0)
        to_solid = next(c for c in fluid if 4 < c[0] < n - 4)
        to_fluid = solid[len(solid) // 2]
        deltas = (A.WorldDelta * 2)()
        for d, c, v in zip(deltas, (to_solid, to_fluid), (1, 0)):
            d.type, d.x, d.y, d.z, d.data0, d.data1 = 8, int(c[0]), int(c[1]), int(c[2]), 12, v
        A.check(lib.aero_lbm_simulation_submit_world_deltas(w.service, deltas, 2), "deltas")
        log("world step after two cell patches")
        w.step(3)
        cells = np.array([(c[0] * n + c[1]) * n + c[2] for c in (to_solid, to_fluid)], np.int32)
        A.check(lib.aero_lbm_patch_d3q27_f16_static_cells_rect(n, n, n, m.ctx.key, 2, cells.ctypes.data_as(A._I),
                                                               np.array([1, 0], np.uint8).ctypes.data_as(A._U8),
                                                               np.zeros(2, np.uint8).ctypes.data_as(A._U8)), "patch")
        m.step(3)
        got, air_got, _ = w.dynamic((0, 0, 0))
        assert np.array_equal(got, m.ctx.flow())
        assert np.array_equal(air_got, m.ctx.temperature())
        assert np.all(got.reshape(-1, 4)[cells[0]] == 0.0)
        info = lib.aero_lbm_simulation_runtime_info().decode()
        assert "patch_steps=1" in info and "patch_cells=2" in info and "patch_failures=0" in info, info
        m.ctx.release()
    finally:
        w.release()


def test_boundary_reference_shell_and_seeding(lib):
    """Brick A carries flow and a boundary reference; brick B (x+ neighbour) has no
    state and is seeded from A's mean with A's adjacent layer as its x- ghost."""
    n = 16
    mode = A.SOLVER_D3Q27_FP16_INPLACE
    ob_a, ob_b = brick_static(n, 5), brick_static(n, 6)
    flow = np.tile(np.array([0.03, 0.005, 0.0, 0.001], np.float32), n ** 3)
    air = np.zeros(n ** 3, np.float32)
    bref = np.tile(np.array([0.05, 0.0, -0.01, 0.002], np.float32), n ** 3)
    w = A.BrickWorld(world=12, size=n, mode=mode)
    try:
        w.hints([(0, 0, 0), (1, 0, 0)], exact=True)
        w.upload_static((0, 0, 0), ob_a)
        w.upload_static((1, 0, 0), ob_b)
        w.upload_dynamic((0, 0, 0), flow, air)
        w.upload_dynamic((0, 0, 0), bref, air, reference=True)
        log("world with shell: step 1")
        w.step(1)
        # A: its x+ face touches B, the other five are exposed and read the reference
        fa, aa = constrained(ob_a, flow, air)
        exposed = [True, False, True, True, True, True]
        pk_a = B.build_packet(n, ob_a, np.zeros(n ** 3, np.uint8), np.zeros(n ** 3, np.float32), fa, aa, exposed,
                              bref, air)
        ma = Manual(lib, n, 4101, mode, pk_a)
        A.check(lib.aero_lbm_set_d3q27_f16_boundary_shell_rect(n, n, n, ma.ctx.key, A.fptr(bref), 4, 0b111101, 8,
                                                               0.1), "shell")
        # B: seeded with A's (constrained) mean, then A's x = n-1 layer as its x = 0 ghost
        seed = B.seed_from_neighbours([(fa, aa, np.zeros(n ** 3, np.float32))])
        fb = np.tile(seed[0], n ** 3)
        ab = np.full(n ** 3, seed[1], np.float32)
        sb = np.full(n ** 3, seed[2], np.float32)
        B.copy_ghost(n, fa, aa, np.zeros(n ** 3, np.float32), 0, False, fb, ab, sb)
        fb, ab = constrained(ob_b, fb, ab)
        pk_b = B.build_packet(n, ob_b, np.zeros(n ** 3, np.uint8), np.zeros(n ** 3, np.float32), fb, ab,
                              [False] * 6)
        mb = Manual(lib, n, 4102, mode, pk_b)
        got_a, _, _ = w.dynamic((0, 0, 0))
        got_b, _, _ = w.dynamic((1, 0, 0))
        assert np.array_equal(got_a, ma.ctx.flow())
        assert np.array_equal(got_b, mb.ctx.flow())
        w.step(5)
        ma.step(5)
        mb.step(5)
        got_a, _, _ = w.dynamic((0, 0, 0))
        got_b, _, _ = w.dynamic((1, 0, 0))
        assert np.array_equal(got_a, ma.ctx.flow())
        assert np.array_equal(got_b, mb.ctx.flow())
        ma.ctx.release()
        mb.ctx.release()
    finally:
        w.release()


def shell_targets(bref, n):
    """FP16-rounded (ties away) shell targets after the speed and pressure clamps."""
    f = bref.reshape(-1, 4).astype(np.float32).copy()
    kmax = np.float32(0.57735026919 * 0.6)
    s2 = (f[:, :3] ** 2).sum(1)
    over = s2 > kmax * kmax
    f[over, :3] *= (kmax / np.sqrt(s2[over]))[:, None]
    f[:, 3] = np.clip(f[:, 3], -0.03, 0.03)
    bits = np.array([[B.half_bits(v) for v in row] for row in f], np.uint16)
    return bits.view(np.float16).astype(np.float32)


def test_shell_blend_vs_oracle(lib, oracle_mod):
    n, steps = 12, 30
    solid = W.solid_boxes(n, n, n, 0.08, seed=9, edge=(2, 3), keep_clear_x=2)
    pk = W.make_packet(solid, u0=(0.02, 0.0, 0.0), perturb=2e-3, seed_u=4)
    bref = np.tile(np.array([0.06, -0.02, 0.01, 0.004], np.float32), n ** 3)
    bref[0::4] += np.linspace(0, 0.01, n ** 3, dtype=np.float32)
    v, _ = B.packet_mean_velocity(pk.reshape(-1, 11))
    cfg = compact_config(v, n)
    ov = W.reference_overrides(A.KERNEL_FAST)
    A.check(lib.aero_lbm_init_rect(n, n, n, 11, 4), "init")
    A.check(lib.aero_lbm_benchmark_set_config(ctypes.byref(cfg)), "config")
    lib.aero_lbm_set_realtime_solver_mode(A.SOLVER_D3Q27_FP32_INPLACE)
    ctx = A.Context(n, n, n, key=4201)
    ctx.step(pk)
    mask, layers, relax = 0b110101, 5, 0.3
    A.check(lib.aero_lbm_set_d3q27_f16_boundary_shell_rect(n, n, n, ctx.key, A.fptr(bref), 4, mask, layers, relax),
            "shell")
    o = oracle_mod.Oracle(n, n, n, cfg, overrides=ov, precision=32)
    o.step(pk)
    o.set_shell(B.shell_layers(n, mask, layers, solid.reshape(-1)), shell_targets(bref, n), layers, relax)
    for _ in range(steps - 1):
        ctx.step(None)
        o.step(None)
    got = ctx.flow().reshape(-1, 4)
    ref = o.flow().reshape(-1, 4)
    fluid = solid.reshape(-1) == 0
    err = float(np.max(np.abs(got[fluid, :3] - ref[fluid, :3])))
    log(f"shell blend FP32 in place vs FP32 oracle: max velocity difference {err:.3e}")
    assert err < 1e-5
    # the shell acts: without it the run ends elsewhere
    o2 = oracle_mod.Oracle(n, n, n, cfg, overrides=ov, precision=32)
    o2.run(pk, steps)
    assert float(np.max(np.abs(o2.flow().reshape(-1, 4)[fluid, :3] - ref[fluid, :3]))) > 1e-3
    ctx.release()
    lib.aero_lbm_set_realtime_solver_mode(0)


def test_classic_world_runtime_mode_matches_oracle_chain(oracle_mod, lib):
    """The mod's default classic solver mode: every step rebuilds the brick's packet
    from its dynamic state (fan disks from fan surface kinds, heat from an emitter),
    steps the context in runtime mode with the benchmark configuration reset, reads
    the flow and air temperature back and applies the static constraints
    (native/src/aero_lbm_simulation_bridge.cpp:1706-1832). The same chain on the FP32
    oracle gives bit-identical dynamic state after every step."""
    n = 16
    ob = brick_static(n, 9)
    kind = np.zeros(n ** 3, np.uint8)
    em = np.zeros(n ** 3, np.float32)
    for x in range(4, 7):
        for y in range(6, 9):
            for z in range(6, 9):
                c = B.cell(n, x, y, z)
                if not ob[c]:
                    kind[c] = 33   # fan blowing toward x+
    for x in range(9, 11):
        c = B.cell(n, x, 2, 8)
        if not ob[c]:
            em[c] = 4000.0        # watts
    rng = np.random.default_rng(4)
    flow = np.tile(np.array([0.03, 0.0, 0.01, 0.0], np.float32), n ** 3)
    flow[0::4] += rng.uniform(-2e-3, 2e-3, n ** 3).astype(np.float32)
    air = (0.05 * rng.uniform(-1.0, 1.0, n ** 3)).astype(np.float32)
    zeros16 = np.zeros(n ** 3, np.uint16)
    surf = np.zeros(n ** 3, np.float32)

    def constrain(f, a):
        f, a, s = f.copy(), a.copy(), surf.copy()
        B.apply_constraints(ob, zeros16, kind, em, f, a, s)
        return f, a

    lib.aero_lbm_benchmark_reset_config()  # the mod's classic bricks step with no benchmark configuration
    w = A.BrickWorld(world=13, size=n, mode=A.SOLVER_CLASSIC_D3Q27)
    try:
        w.hints([(0, 0, 0)], exact=True)
        w.upload_static((0, 0, 0), ob, surface_kind=kind, emitter=em)
        w.upload_dynamic((0, 0, 0), flow, air)
        cfg = A.default_config()  # the classic brick steps with the configuration reset: runtime mode
        o = oracle_mod.Oracle(n, n, n, cfg, precision=32)
        f, a = constrain(flow, air)
        for step in range(1, 9):
            log(f"classic world step {step}")
            w.step(1)
            pk = B.build_packet(n, ob, kind, em, f, a, [True] * 6)
            o.step(pk.reshape(-1))
            f, a = constrain(o.flow(), o.temperature())
            got_f, got_a, _ = w.dynamic((0, 0, 0))
            assert np.array_equal(got_f, f), (step, float(np.max(np.abs(got_f - f))))
            assert np.array_equal(got_a, a), step
        st = w.status()
        assert st[3] == 1
        assert np.abs(f[0::4]).max() > 0.04    # the fan disk has pushed the flow (0.032 at the start)
    finally:
        w.release()


def test_classic_world_two_bricks_exchange_halo_like_oracle_chain(oracle_mod, lib):
    """Two face-adjacent bricks along x in the classic solver mode. The first world
    step bootstraps each brick's ghost face from its neighbour's uploaded state
    (simulation_bridge.cpp:1354-1409); before every later step the two exact FP32
    contexts exchange one ghost layer (aero_lbm_exchange_halo_layers_rect, 1056-1089,
    1411-1439). The same chain on two FP32 oracle contexts (the oracle's halo exchange,
    pinned in test_oracle_context_ops.py) gives bit-identical dynamic states."""
    n = 16
    obs = {(0, 0, 0): brick_static(n, 31), (1, 0, 0): brick_static(n, 32)}
    rng = np.random.default_rng(6)
    zeros8, zeros16, zf = np.zeros(n ** 3, np.uint8), np.zeros(n ** 3, np.uint16), np.zeros(n ** 3, np.float32)

    def constrain(ob, f, a):
        f, a, s = f.copy(), a.copy(), zf.copy()
        B.apply_constraints(ob, zeros16, zeros8, zf, f, a, s)
        return f, a

    up = {}
    for c, u in (((0, 0, 0), 0.03), ((1, 0, 0), 0.045)):
        fl = np.tile(np.array([u, 0.0, 0.005, 0.0], np.float32), n ** 3)
        fl[0::4] += rng.uniform(-2e-3, 2e-3, n ** 3).astype(np.float32)
        up[c] = (fl, (0.02 * rng.uniform(-1.0, 1.0, n ** 3)).astype(np.float32))
    lib.aero_lbm_benchmark_reset_config()  # classic bricks step in runtime mode
    w = A.BrickWorld(world=15, size=n, mode=A.SOLVER_CLASSIC_D3Q27)
    try:
        w.hints(list(obs), exact=True)
        for c in obs:
            w.upload_static(c, obs[c])
            w.upload_dynamic(c, *up[c])
        cfg = A.default_config()
        orc = {c: oracle_mod.Oracle(n, n, n, cfg, precision=32) for c in obs}
        st = {c: constrain(obs[c], *up[c]) for c in obs}  # the snapshots the bootstrap copies from
        a_, b_ = (0, 0, 0), (1, 0, 0)
        boot = {}
        for c, nb, positive in ((a_, b_, True), (b_, a_, False)):
            f, a = st[c][0].copy(), st[c][1].copy()
            B.copy_ghost(n, st[nb][0], st[nb][1], zf, 0, positive, f, a, zf.copy())
            boot[c] = constrain(obs[c], f, a)
        st = boot
        for step in range(1, 7):
            log(f"classic two-brick world step {step}")
            w.step(1)
            if step > 1:
                orc[a_].exchange_halo(orc[b_], (1, 0, 0), 1)
            for c in obs:
                pk = B.build_packet(n, obs[c], zeros8, zf, st[c][0], st[c][1], [True] * 6)
                orc[c].step(pk.reshape(-1))
            st = {c: constrain(obs[c], orc[c].flow(), orc[c].temperature()) for c in obs}
            for c in obs:
                got_f, got_a, _ = w.dynamic(c)
                assert np.array_equal(got_f, st[c][0]), (step, c, float(np.max(np.abs(got_f - st[c][0]))))
                assert np.array_equal(got_a, st[c][1]), (step, c)
    finally:
        w.release()


def run_world(bricks, n, mode, steps, batching):
    """Steps a world of the given bricks {coord: (obstacle, flow)}; returns the
    dynamic state and stride-1/stride-4 atlases of every brick and the runtime info."""
    w = A.BrickWorld(world=14, size=n, mode=mode)
    try:
        w.batching(batching)
        w.hints(list(bricks), exact=True)
        for c, (ob, flow) in bricks.items():
            w.upload_static(c, ob)
            w.upload_dynamic(c, flow, np.zeros(n ** 3, np.float32))
        for k in range(steps):
            w.step(1)
            if k == 2:
                # an atlas between steps: the device publish must not disturb the state
                w.atlas(next(iter(bricks)), 4)
        out = {}
        for c in bricks:
            fl, air, _ = w.dynamic(c)
            out[c] = (fl, air, w.atlas(c, 1), w.atlas(c, 4))
        info = A.load().aero_lbm_simulation_runtime_info().decode()
        return out, info
    finally:
        w.release()


@pytest.mark.parametrize("mode", [A.SOLVER_D3Q27_FP16_INPLACE, A.SOLVER_D3Q27_S16_INPLACE])
def test_world_batched_steps_equal_per_brick_steps(lib, mode):
    """A world of 2 x 2 x 2 bricks, each with its own solid boxes and inflow: with
    batching on, the cached bricks of a world step take one launch set per parity
    group and their atlases are quantised on the device; the dynamic state, the
    temperature and the packed atlases equal the per-brick cached steps (batching
    off, each brick through aero_lbm_step_rect_cached) bit for bit."""
    n, steps = 16, 9
    rng = np.random.default_rng(11)
    bricks = {}
    for i, c in enumerate([(x, y, z) for x in range(2) for y in range(2) for z in range(2)]):
        ob = brick_static(n, 20 + i)
        u = rng.uniform(0.02, 0.06)
        ang = rng.uniform(0, 2 * np.pi)
        flow = np.tile(np.array([u * np.cos(ang), 0.0, u * np.sin(ang), 0.0], np.float32), n ** 3)
        flow[0::4] += rng.uniform(-1e-3, 1e-3, n ** 3).astype(np.float32)
        bricks[c] = (ob, flow)
    log(f"world 2x2x2 mode={mode} batched")
    on, info_on = run_world(bricks, n, mode, steps, True)
    log(f"world 2x2x2 mode={mode} per brick")
    off, info_off = run_world(bricks, n, mode, steps, False)
    assert "batched_world_steps=0" not in info_on, info_on
    assert "batched_world_steps=0" in info_off, info_off
    for c in bricks:
        for a, b, what in zip(on[c], off[c], ("flow", "air", "atlas1", "atlas4")):
            assert np.array_equal(a, b), (c, what)
    # the bricks did evolve, and differently
    assert not np.array_equal(on[(0, 0, 0)][0], on[(1, 1, 1)][0])
