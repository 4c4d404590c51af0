"""Oracle pins for the cylinder crossflow preset (ref: native/src/aero_lbm_jni.cpp:190-228
geometry and inlet profile, 366-380 preset, 1540-1544 inlet rule, 2452-2492 drag).

From the mathematics: the populations entering the x = 0 plane along +x are the
textbook D3Q27 equilibrium at the parabolic inlet velocity 4 u_max s (1 - s); and
for a fluid at rest with a linear density rho = 1 + g x the momentum-exchange force
on the solid body obeys a discrete Archimedes law. Pairing every link that enters a
solid run along c_q with the link that leaves it gives
    F = -2 g sum_{q in a half set, c_qz = 0} w_q c_q c_qx (S + R_q)
per z plane, S = solid cells in the plane, R_q = number of maximal solid runs along
lines of direction c_q. The divergence argument holds for any solid shape away from
the domain edges; a dropped factor, a wrong link sign or reading the solid side's
population breaks it.
"""
import numpy as np
import pytest

import aero_lbm as A
import workloads as W


def cylinder_cfg():
    return W.cylinder_config()


def d3q27():
    c = np.array([(ix - 1, iy - 1, iz - 1) for ix in range(3) for iy in range(3) for iz in range(3)])
    w1 = {-1: 1.0 / 6.0, 0: 2.0 / 3.0, 1: 1.0 / 6.0}
    w = np.array([w1[a] * w1[b] * w1[d] for a, b, d in c])
    return c, w


def test_inlet_populations_are_equilibrium_of_the_parabola(oracle_mod):
    nx, ny, nz = 12, 11, 4
    cfg = cylinder_cfg()
    solid = np.zeros((nx, ny, nz), np.uint8)
    packet = W.make_packet(solid, u0=(0.0, 0.0, 0.0), perturb=0.0)
    o = oracle_mod.Oracle(nx, ny, nz, cfg, overrides=W.reference_overrides(A.KERNEL_EXACT), precision=64)
    o.step(packet)
    f = o.populations().reshape(27, nx, ny, nz)
    c, w = d3q27()
    u_max = float(np.float32(0.12))
    kmax = 0.57735026919 * 0.60
    for y in range(1, ny - 1):  # y = 0 / ny-1: diagonal links come from the y walls
        s = y / (ny - 1)
        ux = min(4.0 * u_max * s * (1.0 - s), kmax)
        for q in range(27):
            if c[q, 0] != 1:
                continue
            cu = 3.0 * c[q, 0] * ux
            feq = w[q] * 1.0 * (1.0 + cu + 0.5 * cu * cu - 1.5 * ux * ux)
            assert f[q, 0, y, :] == pytest.approx(np.full(nz, feq), abs=1e-12), (q, y)


def runs_along(mask2d, d):
    """Number of maximal runs of True along lines of direction d = (dx, dy) in a 2-D mask."""
    nx, ny = mask2d.shape
    count = 0
    for x in range(nx):
        for y in range(ny):
            if not mask2d[x, y]:
                continue
            px, py = x - d[0], y - d[1]
            if 0 <= px < nx and 0 <= py < ny and mask2d[px, py]:
                continue
            count += 1
    return count


@pytest.mark.parametrize("dims", [(24, 17, 3), (90, 33, 2)])
def test_rest_drag_is_discrete_archimedes(oracle_mod, dims):
    nx, ny, nz = dims
    g = 1.0e-3
    cfg = cylinder_cfg()
    # the preset's own cylinder as the solid mask, fluid at rest with rho = 1 + g x
    disk = W.cylinder_mask(nx, ny, 1)[:, :, 0].astype(bool)
    assert disk.sum() > 4 and not disk[0].any() and not disk[:, 0].any() and not disk[:, -1].any()
    solid = np.repeat(disk[:, :, None], nz, axis=2).astype(np.uint8)
    packet = W.make_packet(solid, u0=(0.0, 0.0, 0.0), perturb=0.0).reshape(-1, 11)
    xs = np.repeat(np.arange(nx), ny * nz).astype(np.float64)
    packet[:, 8] = (g * xs).astype(np.float32)
    o = oracle_mod.Oracle(nx, ny, nz, cfg, overrides=W.reference_overrides(A.KERNEL_EXACT), precision=64)
    o.step(packet.reshape(-1))
    got = o.force().astype(np.float64)
    c, w = d3q27()
    S = int(disk.sum())
    want = np.zeros(3)
    gx = np.float64(np.float32(g))
    for q in range(27):
        cq = c[q]
        if cq[2] != 0 or not (cq[0] > 0 or (cq[0] == 0 and cq[1] > 0)):
            continue  # half set of the in-plane links
        R = runs_along(disk, (cq[0], cq[1]))
        want += -2.0 * gx * w[q] * cq * cq[0] * (S + R)
    # the ramp is stored as float32 per cell: the pairing argument holds to that rounding
    assert np.max(np.abs(got - want)) < 1e-6 * max(1.0, np.abs(want).max()) + 1e-9
    assert abs(want[0]) > 1e-3
