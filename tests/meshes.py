"""Small hand-built meshes whose vertices have many incident elements (for the
limits of the element-task vertex blocks)."""
from __future__ import annotations

import numpy as np

# Hexa27 node k at tensor indices (a, b, c) = kNodeABC[k] // 9, // 3 % 3, % 3
# (corners, edge mid-points, face centres, centre), the engine's node order.
NODE_ABC = [0, 18, 24, 6, 2, 20, 26, 8, 9, 21, 15, 3, 1, 19, 25, 7, 11, 23, 17, 5, 12, 10, 22, 16, 4, 14, 13]
P2_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]


def _dedupe(points):
    """Unique points (rounded to 1e-9) and the index of every input point."""
    keys = np.round(np.asarray(points, np.float64) / 1e-9).astype(np.int64)
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    return (uniq * 1e-9).astype(np.float32), inv.reshape(-1).astype(np.int32)


def tet_fan(m: int):
    """2 m linear tetrahedra around the z axis sharing the origin: (O, r_i,
    r_i+1, top) and (O, r_i+1, r_i, bottom), r_i on the unit circle."""
    ang = 2.0 * np.pi * np.arange(m) / m
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(m)], 1)
    xyz = np.concatenate([[[0.0, 0.0, 0.0]], ring, [[0.0, 0.0, 1.0]], [[0.0, 0.0, -1.0]]]).astype(np.float32)
    top, bot = m + 1, m + 2
    topo = []
    for i in range(m):
        a, b = 1 + i, 1 + (i + 1) % m
        topo.append([0, a, b, top])
        topo.append([0, b, a, bot])
    return xyz, np.asarray(topo, np.int32)


def p1_to_p2(xyz, topo):
    """Quadratic tetrahedra with straight edges: edge mid-points in the
    engine's Tetra10 edge order."""
    pts = [p for p in xyz.astype(np.float64)]
    for t in topo:
        for i, j in P2_EDGES:
            pts.append(0.5 * (xyz[t[i]].astype(np.float64) + xyz[t[j]].astype(np.float64)))
    uniq, inv = _dedupe(pts)
    n = len(xyz)
    out = np.zeros((len(topo), 10), np.int32)
    for e, t in enumerate(topo):
        out[e, :4] = inv[t]
        out[e, 4:] = inv[n + 6 * e: n + 6 * e + 6]
    return uniq, out


def hex_book(m: int):
    """m straight-sided 27-node hexahedra around the z axis sharing the axis
    edge: bottom quad (O, P_i, Q_i, P_i+1), |P| = 1, Q_i at radius 1.2 halfway
    between, top = bottom + z; the nodes are images of the reference nodes
    under each element's trilinear map."""
    ang = 2.0 * np.pi * np.arange(m) / m
    pts, rows = [], []
    for i in range(m):
        a0, a1 = ang[i], ang[(i + 1) % m] + (2.0 * np.pi if i + 1 == m else 0.0)
        am = 0.5 * (a0 + a1)
        quad = np.array([[0.0, 0.0], [np.cos(a0), np.sin(a0)], [1.2 * np.cos(am), 1.2 * np.sin(am)],
                         [np.cos(a1), np.sin(a1)]])
        corners = np.array([[*quad[k], 0.0] for k in range(4)] + [[*quad[k], 1.0] for k in range(4)])
        sign = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                         [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float64)
        row = []
        for k in range(27):
            t = NODE_ABC[k]
            xi = np.array([t // 9, t // 3 % 3, t % 3], np.float64) - 1.0
            w = np.prod(1.0 + sign * xi, axis=1) / 8.0
            row.append(len(pts))
            pts.append(w @ corners)
        rows.append(row)
    uniq, inv = _dedupe(pts)
    return uniq, inv[np.asarray(rows)].astype(np.int32)
