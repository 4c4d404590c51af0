"""Pins of the oracle's restatement of the reference's context operations
(native/src/aero_lbm_jni.cpp): the context shift (8432-8476), the halo exchange
between two contexts (8201-8426) and the static-cell patch (6156-6415, 4725-4778).

The checks come from what the operations mean, not from their code: a shift moves
every value by an exact index offset and fills what enters from outside with the
rest equilibrium (the product of the 1-D weights); the halo exchange of two bricks
cut with an overlap from one larger field restores each brick's ghost layers on the
core cross-section and touches nothing else; the patch restarts exactly the cells
an entry changed."""
import itertools

import numpy as np
import pytest

import aero_lbm as A
import workloads as W

W1 = np.array([1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0])
WQ = np.array([W1[q // 9] * W1[(q // 3) % 3] * W1[q % 3] for q in range(27)])


def stepped_oracle(oracle_mod, dims, seed=3, steps=3, precision=64):
    cfg = W.channel_config(0.05, 50.0, sgs=True)
    solid = W.solid_boxes(*dims, 0.1, seed=seed, edge=(2, 3), keep_clear_x=2)
    t = W.hot_blob(*dims, 0.02, 2.0, seed=seed + 1, axis=2)
    t[solid.astype(bool)] = 0.0
    pk = W.make_packet(solid, u0=(0.05, 0.01, 0.0), perturb=1e-3, seed_u=seed + 2, temperature=t)
    o = oracle_mod.Oracle(*dims, cfg, precision=precision)
    o.set_temperature(t)
    o.step(pk)
    for _ in range(steps - 1):
        o.step(None)
    return o, solid


def test_shift_moves_by_the_offset_and_fills_rest_equilibrium(oracle_mod):
    dims = (10, 9, 8)
    o, _ = stepped_oracle(oracle_mod, dims)
    f0 = o.populations().reshape(27, *dims)
    t0 = o.temperature().reshape(dims)
    d = (2, -1, 3)
    o.shift(*d)
    f1 = o.populations().reshape(27, *dims)
    t1 = o.temperature().reshape(dims)
    for x, y, z in itertools.product(*(range(n) for n in dims)):
        sx, sy, sz = x + d[0], y + d[1], z + d[2]
        inside = 0 <= sx < dims[0] and 0 <= sy < dims[1] and 0 <= sz < dims[2]
        if inside:
            assert np.array_equal(f1[:, x, y, z], f0[:, sx, sy, sz])
            assert t1[x, y, z] == t0[sx, sy, sz]
        else:
            assert np.array_equal(f1[:, x, y, z], WQ), (x, y, z)
            assert t1[x, y, z] == 0.0
    # the rest equilibrium carries density 1 and no momentum
    assert np.isclose(WQ.sum(), 1.0, rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        o.shift(10, 0, 0)


def test_shift_and_back_restores_the_cells_that_stayed_inside(oracle_mod):
    dims = (9, 8, 7)
    o, _ = stepped_oracle(oracle_mod, dims, seed=5)
    f0 = o.populations().reshape(27, *dims)
    o.shift(1, 2, -1)
    o.shift(-1, -2, 1)
    f2 = o.populations().reshape(27, *dims)
    # new(x) = old(x + d1) then new2(x) = new(x + d2) = old(x): x keeps its value
    # when x + d2 lies inside
    keep = np.zeros(dims, bool)
    keep[1:, 2:, :-1] = True
    assert np.array_equal(f2[:, keep], f0[:, keep])
    assert np.array_equal(f2[:, ~keep][:, 0], WQ)


@pytest.mark.parametrize("offset,halo", [((1, 0, 0), 1), ((0, -1, 0), 2), ((0, 0, 1), 2), ((1, -1, 0), 1)])
def test_halo_exchange_restores_ghost_layers_of_overlapping_bricks(oracle_mod, offset, halo):
    n = 10
    core = n - 2 * halo
    a, _ = stepped_oracle(oracle_mod, (n, n, n), seed=7, steps=1)
    b, _ = stepped_oracle(oracle_mod, (n, n, n), seed=8, steps=1)
    # one field over the union; brick b lies `core` cells from brick a along the offset
    rng = np.random.default_rng(11)
    span = [n + core * abs(o) for o in offset]
    g = rng.random((27, *span))
    ta = [0 if o >= 0 else core for o in offset]      # brick a's origin in the union
    tb = [ta[k] + offset[k] * core for k in range(3)]  # brick b's origin
    sl = lambda o: tuple(slice(o[k], o[k] + n) for k in range(3))
    fa, fb = g[(slice(None),) + sl(ta)].copy(), g[(slice(None),) + sl(tb)].copy()
    ga, gb = fa.copy(), fb.copy()
    # spoil each brick's ghost layers facing the other
    for k, o in enumerate(offset):
        if o == 0:
            continue
        idx_a = [slice(None)] * 4
        idx_b = [slice(None)] * 4
        idx_a[k + 1] = slice(n - halo, n) if o > 0 else slice(0, halo)
        idx_b[k + 1] = slice(0, halo) if o > 0 else slice(n - halo, n)
        fa[tuple(idx_a)] = -1.0
        fb[tuple(idx_b)] = -2.0
    a.set_populations(fa)
    b.set_populations(fb)
    a.exchange_halo(b, offset, halo)
    ra = a.populations().reshape(27, n, n, n)
    rb = b.populations().reshape(27, n, n, n)
    # the exchanged boxes: across the offset axes the ghost layers, along the others
    # the core cross-section [halo, n - halo)
    box_a = [slice(n - halo, n) if o > 0 else (slice(0, halo) if o < 0 else slice(halo, n - halo)) for o in offset]
    box_b = [slice(0, halo) if o > 0 else (slice(n - halo, n) if o < 0 else slice(halo, n - halo)) for o in offset]
    ba = (slice(None),) + tuple(box_a)
    bb = (slice(None),) + tuple(box_b)
    assert np.array_equal(ra[ba], ga[ba]) and np.array_equal(rb[bb], gb[bb])
    # everything else as it was (spoiled ghost cells outside the box included)
    ma = np.ones((27, n, n, n), bool)
    ma[ba] = False
    mb = np.ones((27, n, n, n), bool)
    mb[bb] = False
    assert np.array_equal(ra[ma], fa[ma]) and np.array_equal(rb[mb], fb[mb])
    with pytest.raises(ValueError):
        a.exchange_halo(b, (2, 0, 0), halo)


def test_patch_restarts_exactly_the_changed_cells(oracle_mod):
    dims = (10, 8, 9)
    o, solid = stepped_oracle(oracle_mod, dims, seed=13)
    f0 = o.populations()
    flat = solid.reshape(-1)
    fluid = np.flatnonzero(flat == 0)
    solids = np.flatnonzero(flat != 0)
    c_to_solid, c_keep, c_toggle = int(fluid[5]), int(fluid[40]), int(fluid[77])
    c_to_fluid = int(solids[3])
    cells = [c_to_solid, c_to_fluid, c_keep, c_toggle, c_toggle]
    values = [1, 0, 0, 1, 0]
    o.patch_cells(cells, values)
    f1 = o.populations()
    assert np.array_equal(f1[:, c_to_solid], np.zeros(27))
    assert np.array_equal(f1[:, c_to_fluid], WQ)
    assert np.array_equal(f1[:, c_toggle], WQ)      # changed twice: restarted fluid
    assert np.array_equal(f1[:, c_keep], f0[:, c_keep])
    others = np.setdiff1d(np.arange(f0.shape[1]), [c_to_solid, c_to_fluid, c_toggle])
    assert np.array_equal(f1[:, others], f0[:, others])
    # the new flags take effect: a fluid neighbour of the new solid cell bounces back
    o.step(None)
    u = o.flow().reshape(-1, 4)
    assert np.all(u[c_to_solid] == 0.0)
    assert np.all(np.isfinite(u))
