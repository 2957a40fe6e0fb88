"""ORACLE - test infrastructure only.

Pure-Python restatement of the reference's mesh decimation, EdgeCollapse and
TriangleCollapse (Geometry.inl:50-254), for small meshes: the same visiting
order (last triangle first, removal by swapping the last triangle in), the
same float arithmetic (numpy float32 scalars) for positions, squared
distances and the cross product, and the same double comparison
(sum of squared edges * edge_ratio > area * 12 / sqrt(3)).
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32


def _sqdist(a, b):  # SquareDistance (Geometry.inl:35-38): float sum, returned as double
    dx, dy, dz = a[0] - b[0], a[1] - b[1], a[2] - b[2]
    return float(dx * dx + dy * dy + dz * dz)


def _follow(remap, i):
    while remap[i] < i:
        i = remap[i]
    return i


def decimate(vertices, triangles, mode: int, edge_ratio: float):
    """Returns (vertices float32 (n', 3), triangles int32 (m', 3))."""
    pos = [[F(v[0]), F(v[1]), F(v[2])] for v in np.asarray(vertices, np.float32)]
    tri = [list(map(int, t)) for t in np.asarray(triangles)]
    n = len(pos)
    remap = list(range(n))
    count = [1] * n
    ratio = 12.0 / math.sqrt(3.0)
    er = float(F(edge_ratio))

    def pop_into(i):
        tri[i] = tri[-1]
        tri.pop()

    for i in range(len(tri) - 1, -1, -1):
        idx = [_follow(remap, v) for v in tri[i]]
        if idx[0] == idx[1] or idx[0] == idx[2] or idx[1] == idx[2]:
            pop_into(i)
            continue
        p = [[pos[k][a] / F(count[k]) for a in range(3)] for k in idx]
        q0 = [p[1][a] - p[0][a] for a in range(3)]
        q1 = [p[2][a] - p[0][a] for a in range(3)]
        d = [_sqdist(p[j], p[(j + 1) % 3]) for j in range(3)]
        c = [q0[1] * q1[2] - q0[2] * q1[1], -q0[0] * q1[2] + q0[2] * q1[0], q0[0] * q1[1] - q0[1] * q1[0]]
        area = math.sqrt(float(c[0] * c[0] + c[1] * c[1] + c[2] * c[2])) / 2
        if not ((d[0] + d[1] + d[2]) * er > area * ratio):
            continue
        if mode == 0:
            j = 0
            if d[1] < d[j]:
                j = 1
            if d[2] < d[j]:
                j = 2
            a0, b0 = idx[j], idx[(j + 1) % 3]
            i1, i2 = (a0, b0) if a0 < b0 else (b0, a0)
            for a in range(3):
                pos[i1][a] = pos[i1][a] + pos[i2][a]
            count[i1] += count[i2]
            remap[i2] = i1
        else:
            if idx[0] < idx[1]:
                i1, i2, i3 = (idx[0], idx[2], idx[1]) if idx[0] < idx[2] else (idx[2], idx[0], idx[1])
            else:
                i1, i2, i3 = (idx[1], idx[2], idx[0]) if idx[1] < idx[2] else (idx[2], idx[1], idx[0])
            for a in range(3):
                pos[i1][a] = pos[i1][a] + (pos[i2][a] + pos[i3][a])
            count[i1] += count[i2] + count[i3]
            remap[i2] = i1
            remap[i3] = i1
        pop_into(i)
    out = []
    newid = [0] * n
    for i in range(n):
        avg = [pos[i][a] / F(count[i]) for a in range(3)]
        if remap[i] == i:
            newid[i] = len(out)
            out.append(avg)
    for i in range(len(tri) - 1, -1, -1):
        idx = [_follow(remap, v) for v in tri[i]]
        tri[i] = [newid[k] for k in idx]
        if idx[0] == idx[1] or idx[0] == idx[2] or idx[1] == idx[2]:
            pop_into(i)
    return np.array(out, np.float32).reshape(-1, 3), np.array(tri, np.int32).reshape(-1, 3)
