"""Brick-runtime oracle (numpy) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/ as the checker. A plain restatement of the reference's
host-side brick runtime rules, function by function, each citing the reference
file:line it follows (native/src/aero_lbm_simulation_bridge.cpp unless noted).
Arrays are cell-major with cell = (x * n + y) * n + z; flow state has 4 floats per
cell (u_x, u_y, u_z, p in lattice units).
"""
from __future__ import annotations

import numpy as np

PACKET_CH = 11
REF_LAYERS = 8          # 261
GHOST_LAYERS = 1        # 487
FAN_SPEED = 4.0         # 522
FAN_LENGTH, FAN_RADIUS = 5, 1          # 505-506
HEAT_HEIGHT, HEAT_COUPLING = 4, 0.85   # 507-508
NBR = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]           # 490-493
JAVA_FACE = [(0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1), (-1, 0, 0), (1, 0, 0)]     # 478-481
FAN_DIR = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]  # 482-486


def cell(n, x, y, z):
    return (x * n + y) * n + z


def fan_velocity(kind: int):
    """decode_client_fan_surface_kind (524-537): kinds 32..37 = fans toward x-, x+, y-, y+, z-, z+."""
    table = {32: (-FAN_SPEED, 0, 0), 33: (FAN_SPEED, 0, 0), 34: (0, -FAN_SPEED, 0), 35: (0, FAN_SPEED, 0),
             36: (0, 0, -FAN_SPEED), 37: (0, 0, FAN_SPEED)}
    return table.get(int(kind))


def power_to_source(watts):
    """temperature_source_from_power_watts (3152-3158), FP32."""
    w = np.float32(watts)
    s = w * np.float32(0.05) / (np.float32(1.225) * np.float32(1005.0) * np.float32(1.0) * np.float32(20.0))
    return np.float32(min(max(s, np.float32(-0.006)), np.float32(0.006)))


def build_packet(n, obstacle, surface_kind, emitter, flow, air, exposed, bref_flow=None, bref_air=None):
    """build_brick_step_packet (1497-1596). exposed: 6 booleans x-, x+, y-, y+, z-, z+
    (a face counts only when a boundary reference is present)."""
    pk = np.zeros((n ** 3, PACKET_CH), np.float32)
    layers = max(1, min(REF_LAYERS, n))
    ref_ok = bref_flow is not None
    for x in range(n):
        for y in range(n):
            for z in range(n):
                c = cell(n, x, y, z)
                if obstacle[c]:
                    pk[c, 0] = 1.0
                    continue
                fv = fan_velocity(surface_kind[c])
                if fv is not None:
                    pk[c, 1] = 1.0
                    pk[c, 2:5] = fv
                use_ref = ref_ok and ((exposed[0] and x < layers) or (exposed[1] and x >= n - layers) or
                                      (exposed[2] and y < layers) or (exposed[3] and y >= n - layers) or
                                      (exposed[4] and z < layers) or (exposed[5] and z >= n - layers))
                src_f, src_a = (bref_flow, bref_air) if use_ref else (flow, air)
                pk[c, 5:9] = src_f[4 * c:4 * c + 4]
                pk[c, 10] = src_a[c]
                if emitter[c] > 0.0:
                    pk[c, 9] = power_to_source(emitter[c])
    return pk


def packet_mean_velocity(pk):
    """summarize_compact_brick_packet (317-358): FP64 mean of the finite reference
    velocities of fluid cells, rounded to FP32; supported when no fan cell."""
    fluid = pk[:, 0] <= 0.5
    fans = int(np.sum(fluid & (pk[:, 1] > 0.5)))
    v = pk[fluid][:, 5:8].astype(np.float64)
    ok = np.all(np.isfinite(v), axis=1)
    if not ok.any():
        return (0.0, 0.0, 0.0), False
    m = v[ok].sum(axis=0) / ok.sum()
    return tuple(float(np.float32(c)) for c in m), fans == 0


def brick_mean(flow, air, surface):
    """compute_brick_mean_state (1264-1285): FP32 running sums in cell order."""
    f = flow.reshape(-1, 4)
    acc = np.zeros(4, np.float32)
    a = np.float32(0)
    s = np.float32(0)
    for c in range(f.shape[0]):
        acc = (acc + f[c]).astype(np.float32)
        a = np.float32(a + air[c])
        s = np.float32(s + surface[c])
    inv = np.float32(1.0) / np.float32(f.shape[0])
    return (acc * inv).astype(np.float32), np.float32(a * inv), np.float32(s * inv)


def nonzero_flow(flow):
    """dynamic_has_nonzero_flow (1249-1262)."""
    f = flow.reshape(-1, 4)[:, :3]
    return bool(np.any(np.abs(f) > np.float32(1e-8)))


def seed_from_neighbours(neighbours):
    """seed_brick_from_neighbor_means (1287-1334): neighbours is the list of
    (flow, air, surface) of face neighbours in the order x-, x+, y-, y+, z-, z+;
    returns the per-cell seed (flow4, air, surface) or None."""
    acc = np.zeros(4, np.float32)
    a = np.float32(0)
    s = np.float32(0)
    count = 0
    for nb in neighbours:
        if nb is None or not nonzero_flow(nb[0]):
            continue
        mf, ma, ms = brick_mean(*nb)
        acc = (acc + mf).astype(np.float32)
        a = np.float32(a + ma)
        s = np.float32(s + ms)
        count += 1
    if count == 0:
        return None
    inv = np.float32(1.0) / np.float32(count)
    return (acc * inv).astype(np.float32), np.float32(a * inv), np.float32(s * inv)


def copy_ghost(n, nb_flow, nb_air, nb_surface, axis, positive, flow, air, surface):
    """copy_brick_face_ghost_from_neighbor (1354-1409), in place."""
    if not nonzero_flow(nb_flow):
        return
    f = flow.reshape(n, n, n, 4)
    nf = nb_flow.reshape(n, n, n, 4)
    a3, na = air.reshape(n, n, n), nb_air.reshape(n, n, n)
    s3, ns = surface.reshape(n, n, n), nb_surface.reshape(n, n, n)
    for layer in range(max(1, min(GHOST_LAYERS, n))):
        dst = n - 1 - layer if positive else layer
        src = layer if positive else n - 1 - layer
        sl_d = [slice(None)] * 3
        sl_s = [slice(None)] * 3
        sl_d[axis] = dst
        sl_s[axis] = src
        f[tuple(sl_d)] = nf[tuple(sl_s)]
        a3[tuple(sl_d)] = na[tuple(sl_s)]
        s3[tuple(sl_d)] = ns[tuple(sl_s)]


def apply_constraints(obstacle, open_face, surface_kind, emitter, flow, air, surface):
    """apply_brick_static_constraints (1204-1240), in place."""
    f = flow.reshape(-1, 4)
    solid = obstacle != 0
    f[solid] = 0.0
    air[solid] = 0.0
    surface[solid] = 0.0
    bare = (~solid) & (open_face == 0) & (surface_kind == 0) & ~(emitter > 0.0)
    surface[bare] = 0.0


def derive_from_sources(n, obstacle, fan_dir_src, emitter_src):
    """rebuild_static_region_derived_fields_from_sources (1987-2097): returns
    (surface_kind, open_face_mask, emitter)."""
    def solid(x, y, z):
        if min(x, y, z) < 0 or max(x, y, z) >= n:
            return True
        return obstacle[cell(n, x, y, z)] != 0

    kind = np.zeros(n ** 3, np.uint8)
    open_face = np.zeros(n ** 3, np.uint16)
    emitter = np.zeros(n ** 3, np.float32)
    for x in range(n):
        for y in range(n):
            for z in range(n):
                if solid(x, y, z):
                    continue
                m = 0
                for f, (dx, dy, dz) in enumerate(JAVA_FACE):
                    if not solid(x + dx, y + dy, z + dz):
                        m |= 1 << f
                open_face[cell(n, x, y, z)] = m
    for fx in range(n):
        for fy in range(n):
            for fz in range(n):
                d = min(int(fan_dir_src[cell(n, fx, fy, fz)]), 6)
                if d == 0:
                    continue
                dx, dy, dz = FAN_DIR[d]
                axis = 0 if d <= 2 else (1 if d <= 4 else 2)
                for dist in range(1, FAN_LENGTH + 1):
                    for a in range(-FAN_RADIUS, FAN_RADIUS + 1):
                        for b in range(-FAN_RADIUS, FAN_RADIUS + 1):
                            lane = [fx, fy, fz]
                            others = [k for k in range(3) if k != axis]
                            lane[others[0]] += a
                            lane[others[1]] += b
                            t = (lane[0] + dx * dist, lane[1] + dy * dist, lane[2] + dz * dist)
                            if solid(*t):
                                continue
                            if any(solid(lane[0] + dx * k, lane[1] + dy * k, lane[2] + dz * k) for k in range(1, dist)):
                                continue
                            kind[cell(n, *t)] = 32 + d - 1
    for x in range(n):
        for y in range(n):
            for z in range(n):
                p = np.float32(emitter_src[cell(n, x, y, z)])
                if not (p > 0) or not np.isfinite(p):
                    continue
                for dist in range(HEAT_HEIGHT + 1):
                    ty = y + dist
                    if ty >= n or solid(x, ty, z):
                        continue
                    if any(solid(x, k, z) for k in range(y + 1, ty)):
                        continue
                    c = cell(n, x, ty, z)
                    add = p if dist == 0 else np.float32(p * np.float32(HEAT_COUPLING) / np.float32(dist * dist))
                    emitter[c] = np.float32(emitter[c] + add)
    return kind, open_face, emitter


def quantize(v, rng):
    """quantize_signed (3122-3128): lround(clamp(v / range, -1, 1) * 32767), 0 for NaN."""
    v = np.asarray(v, np.float32)
    t = np.clip(v / np.float32(rng), -1.0, 1.0).astype(np.float32) * np.float32(32767.0)
    out = np.where(t >= 0, np.floor(t.astype(np.float64) + 0.5), np.ceil(t.astype(np.float64) - 0.5))
    return np.where(np.isfinite(v), out, 0).astype(np.int16)


def packed_atlas(n, stride, flow, velocity_scale):
    """copy_brick_world_packed_flow_atlas, stored-state branch (5383-5408)."""
    sn = (n + stride - 1) // stride
    f = flow.reshape(n, n, n, 4)
    g = np.minimum(n - 1, np.arange(sn) * stride)
    s = f[np.ix_(g, g, g)].reshape(-1, 4)
    out = np.empty((s.shape[0], 4), np.int16)
    out[:, :3] = quantize((s[:, :3] * np.float32(velocity_scale)).astype(np.float32), 5.6)
    out[:, 3] = quantize(s[:, 3], 0.03)
    return out.reshape(-1)


def half_bits(v):
    """float_to_half_bits (native/src/aero_lbm_jni.cpp:1205-1235): nearest, ties away
    from zero, flush below the subnormal range; finite inputs."""
    b = int(np.float32(v).view(np.uint32))
    sign = (b >> 16) & 0x8000
    man = b & 0x7FFFFF
    e = ((b >> 23) & 0xFF) - 127 + 15
    if e <= 0:
        if e < -10:
            return sign
        man |= 0x800000
        shift = 14 - e
        h = man >> shift
        if (man >> (shift - 1)) & 1:
            h += 1
        return sign | h
    if e >= 31:
        return sign | 0x7C00
    h = sign | (e << 10) | (man >> 13)
    if man & 0x1000:
        h += 1
    return h


def shell_layers(n, face_mask, layers, solid):
    """Shell layer per cell (native/src/aero_lbm_jni.cpp:6475-6514): 1 + distance to the
    nearest flagged face within `layers`, 0 outside or on solid cells."""
    x, y, z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    best = np.full((n, n, n), layers + 1)
    cand = [(1, x < layers, x + 1), (2, x >= n - layers, n - x), (4, y < layers, y + 1),
            (8, y >= n - layers, n - y), (16, z < layers, z + 1), (32, z >= n - layers, n - z)]
    for bit, inside, val in cand:
        if face_mask & bit:
            best = np.where(inside, np.minimum(best, val), best)
    out = np.where(best <= layers, best, 0).reshape(-1)
    out[np.asarray(solid).reshape(-1) != 0] = 0
    return out.astype(np.uint8)
