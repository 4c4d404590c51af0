"""Pins of the brick-runtime oracle (oracle/bricks_oracle.py), CPU only.

The reference holds no fixtures for its brick runtime and its bridge source
(native/src/aero_lbm_simulation_bridge.cpp) needs <jni.h>, absent here, so it is
not compiled. Each rule is pinned by values fixed by hand from the rule's
statement (closed forms on tiny bricks), chosen so that a wrong axis, a wrong
face order or an off-by-one layer fails.
"""
import numpy as np
import pytest

import bricks_oracle as B


def test_quantize_closed_form():
    r = 5.6
    v = np.array([0.0, r, -r, 2 * r, -3 * r, 0.5 * r, -0.5 * r, np.nan, np.inf, 1e-9], np.float32)
    # lround rounds halves away from zero: 16383.5 -> 16384
    assert B.quantize(v, r).tolist() == [0, 32767, -32767, 32767, -32767, 16384, -16384, 0, 0, 0]
    assert B.quantize(np.float32(0.03), 0.03) == 32767


def test_half_bits_ties_away_and_agrees_with_ieee_elsewhere():
    # 1 + 2^-11 lies halfway between 1 and 1 + 2^-10: ties away from zero -> 0x3C01
    assert B.half_bits(1.0 + 2.0 ** -11) == 0x3C01
    assert B.half_bits(-(1.0 + 2.0 ** -11)) == 0xBC01
    assert B.half_bits(1.0) == 0x3C00 and B.half_bits(-2.0) == 0xC000 and B.half_bits(0.0) == 0
    assert B.half_bits(1e6) == 0x7C00            # overflow -> inf
    assert B.half_bits(2.0 ** -26) == 0          # below half the smallest subnormal
    assert B.half_bits(2.0 ** -24) == 0x0001     # smallest subnormal
    rng = np.random.default_rng(3)
    vals = rng.uniform(-0.4, 0.4, 4000).astype(np.float32)
    ours = np.array([B.half_bits(v) for v in vals], np.uint16)
    ieee = vals.astype(np.float16).view(np.uint16)
    # round-to-nearest-even and ties-away differ only on exact ties, which random
    # FP32 values hit with probability ~2^-13 per value
    assert np.count_nonzero(ours != ieee) <= 2


def test_power_to_source_closed_form():
    # 1000 W * 0.05 s / (1.225 * 1005 * 1 * 20) = 0.00203066...
    assert abs(float(B.power_to_source(1000.0)) - 50.0 / 24622.5) < 1e-9
    assert float(B.power_to_source(1e6)) == pytest.approx(0.006)
    assert float(B.power_to_source(-1e6)) == pytest.approx(-0.006)


def test_packet_channels_and_reference_layers():
    n = 10
    cells = n ** 3
    ob = np.zeros(cells, np.uint8)
    ob[B.cell(n, 5, 5, 5)] = 1
    kind = np.zeros(cells, np.uint8)
    kind[B.cell(n, 2, 3, 4)] = 34           # fan toward y-
    em = np.zeros(cells, np.float32)
    em[B.cell(n, 6, 6, 6)] = 1000.0
    flow = np.tile(np.array([0.01, 0.02, 0.03, 0.001], np.float32), cells)
    air = np.full(cells, 0.2, np.float32)
    bref = np.tile(np.array([-0.05, 0.0, 0.0, -0.002], np.float32), cells)
    bair = np.full(cells, -0.1, np.float32)
    # only the y+ face is exposed: the 8 layers y >= 2 read the reference
    pk = B.build_packet(n, ob, kind, em, flow, air, [False, False, False, True, False, False], bref, bair)
    assert pk[B.cell(n, 5, 5, 5)].tolist() == [1.0] + [0.0] * 10
    fan = pk[B.cell(n, 2, 3, 4)]
    assert fan[1] == 1.0 and fan[2:5].tolist() == [0.0, -4.0, 0.0]
    assert pk[B.cell(n, 0, 1, 0), 5:9].tolist() == pytest.approx([0.01, 0.02, 0.03, 0.001])   # y = 1: own state
    assert pk[B.cell(n, 0, 2, 0), 5:9].tolist() == pytest.approx([-0.05, 0.0, 0.0, -0.002])   # y = 2: reference
    assert pk[B.cell(n, 0, 2, 0), 10] == pytest.approx(-0.1)
    assert pk[B.cell(n, 6, 6, 6), 9] == pytest.approx(50.0 / 24622.5)
    # no reference: every fluid cell carries its own state
    pk2 = B.build_packet(n, ob, kind, em, flow, air, [True] * 6)
    assert np.allclose(pk2[ob == 0][:, 5:9], [0.01, 0.02, 0.03, 0.001])
    v, ok = B.packet_mean_velocity(pk)
    assert not ok  # a fan cell disables the compact boundary


def test_packet_mean_velocity_skips_solid_and_nonfinite():
    pk = np.zeros((4, 11), np.float32)
    pk[0, 5:8] = (0.1, 0.0, 0.0)
    pk[1, 5:8] = (0.3, 0.2, 0.0)
    pk[2, 0] = 1.0
    pk[2, 5:8] = (9.0, 9.0, 9.0)   # solid: ignored
    pk[3, 5:8] = (np.nan, 0.0, 0.0)  # non-finite: ignored
    v, ok = B.packet_mean_velocity(pk)
    assert ok and v == pytest.approx((0.2, 0.1, 0.0))


def test_seed_from_neighbours_means():
    n = 4
    cells = n ** 3
    a = (np.tile(np.array([0.04, 0.0, 0.0, 0.0], np.float32), cells), np.full(cells, 0.1, np.float32),
         np.zeros(cells, np.float32))
    b = (np.tile(np.array([0.0, 0.02, 0.0, 0.01], np.float32), cells), np.full(cells, 0.3, np.float32),
         np.full(cells, 2.0, np.float32))
    still = (np.zeros(cells * 4, np.float32), np.full(cells, 5.0, np.float32), np.zeros(cells, np.float32))
    f, air, surf = B.seed_from_neighbours([a, None, still, b, None, None])  # the flow-free one does not count
    assert f.tolist() == pytest.approx([0.02, 0.01, 0.0, 0.005])
    assert air == pytest.approx(0.2) and surf == pytest.approx(1.0)
    assert B.seed_from_neighbours([still, None]) is None


@pytest.mark.parametrize("axis,positive", [(0, False), (0, True), (1, True), (2, False)])
def test_ghost_copy_takes_the_adjacent_layer(axis, positive):
    n = 5
    cells = n ** 3
    nb = np.arange(cells * 4, dtype=np.float32) + 1.0
    nb_air = np.arange(cells, dtype=np.float32)
    flow = np.zeros(cells * 4, np.float32)
    air = np.zeros(cells, np.float32)
    surf = np.zeros(cells, np.float32)
    B.copy_ghost(n, nb, nb_air, nb_air * 2, axis, positive, flow, air, surf)
    f4 = flow.reshape(n, n, n, 4)
    src = nb.reshape(n, n, n, 4)
    dst_i, src_i = (n - 1, 0) if positive else (0, n - 1)
    take = lambda a, i: np.take(a, i, axis=axis)
    assert np.array_equal(take(f4, dst_i), take(src, src_i))
    # everything else untouched
    mask = np.ones((n, n, n), bool)
    idx = [slice(None)] * 3
    idx[axis] = dst_i
    mask[tuple(idx)] = False
    assert not f4[mask].any()


def test_derive_from_sources_fan_disk_and_heat_plume():
    n = 8
    cells = n ** 3
    ob = np.zeros(cells, np.uint8)
    ob[B.cell(n, 3, 2, 2)] = 1          # blocks the centre lane of the fan beyond x = 2
    ob[B.cell(n, 6, 2, 6)] = 1          # caps the heat column two cells above the source
    fan = np.zeros(cells, np.uint8)
    fan[B.cell(n, 0, 2, 2)] = 2         # fan at x = 0 pushing +x
    heat = np.zeros(cells, np.float32)
    heat[B.cell(n, 6, 0, 6)] = 100.0
    kind, open_face, emitter = B.derive_from_sources(n, ob, fan, heat)
    k = kind.reshape(n, n, n)
    # disk y, z in {1, 2, 3} for x = 1..5 (45 cells) minus the solid cell and the
    # two cells behind it on the centre lane
    assert np.count_nonzero(k == 33) == 45 - 3
    assert k[2, 2, 2] == 33 and k[3, 2, 2] == 0 and k[4, 2, 2] == 0 and k[5, 2, 2] == 0
    assert k[5, 1, 3] == 33 and k[6, 2, 2] == 0
    e = emitter.reshape(n, n, n)
    assert e[6, 0, 6] == pytest.approx(100.0) and e[6, 1, 6] == pytest.approx(85.0)
    assert e[6, 2, 6] == 0.0 and e[6, 3, 6] == 0.0      # solid, then blocked
    assert np.count_nonzero(e) == 2
    # open-face mask bit order y-, y+, z-, z+, x-, x+
    assert open_face[B.cell(n, 0, 0, 0)] == 0b101010      # y+, z+, x+ open at the corner
    assert open_face[B.cell(n, 4, 4, 4)] == 0b111111
    assert open_face[B.cell(n, 3, 2, 2)] == 0              # solid
    assert open_face[B.cell(n, 4, 2, 2)] == 0b101111       # x- neighbour solid


def test_shell_layers_closed_form():
    n = 12
    solid = np.zeros(n ** 3, np.uint8)
    solid[B.cell(n, 1, 5, 5)] = 1
    lay = B.shell_layers(n, 0b000011, 8, solid).reshape(n, n, n)   # x- and x+ exposed
    assert lay[0, 5, 5] == 1 and lay[11, 5, 5] == 1 and lay[3, 0, 0] == 4 and lay[8, 7, 7] == 4
    assert lay[1, 5, 5] == 0                                          # solid
    assert lay[7, 3, 3] == 5 and lay[4, 3, 3] == 5
    lay2 = B.shell_layers(n, 0b010000, 3, np.zeros(n ** 3)).reshape(n, n, n)  # z- only, depth 3
    assert lay2[5, 5, 0] == 1 and lay2[5, 5, 2] == 3 and lay2[5, 5, 3] == 0 and lay2[0, 0, 11] == 0


def test_packed_atlas_sampling():
    n, stride = 6, 4
    flow = np.zeros((n, n, n, 4), np.float32)
    flow[4, 4, 0] = (0.25, -0.1, 0.0, 0.015)   # samples at 0 and 4 per axis
    flow[5, 5, 5] = (1.0, 1.0, 1.0, 0.01)     # never sampled
    out = B.packed_atlas(n, stride, flow.reshape(-1), velocity_scale=10.0).reshape(2, 2, 2, 4)
    assert out[1, 1, 0].tolist() == [14628, -5851, 0, 16384]   # 2.5 / 5.6, -1 / 5.6 of 32767; 0.5 -> 16384
    assert not out[0].any()
