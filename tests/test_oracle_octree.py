"""Pins the oracle's octree (A2-A8, main.cu:126-841) by brute force.

A set-based model builds the same sibling-complete levels from integer cell
coordinates: a depth-d level holds the 8 children of every non-empty depth
d-1 cell in Morton order; neighbours are the cells at offset (dx,dy,dz) whose
parent is non-empty.  Every field of the oracle's arrays must agree.
"""
import numpy as np
import pytest

import oracle as orc


def _cells(p, D):
    """Leaf cell of each normalised point with encodePoint's strict '>' (main.cu:132-164)."""
    q = p.astype(np.float64) * (1 << D)
    c = np.ceil(q).astype(np.int64) - 1
    return np.clip(c, 0, (1 << D) - 1)


def _key(cx, cy, cz, d, D):
    k = 0
    for l in range(1, d + 1):
        sh = d - l
        k |= ((cx >> sh) & 1) << (3 * (D - l) + 2)
        k |= ((cy >> sh) & 1) << (3 * (D - l) + 1)
        k |= ((cz >> sh) & 1) << (3 * (D - l))
    return k


def brute(p, D):
    n = len(p)
    cell = _cells(p, D)
    keys = np.array([_key(*cell[i], D, D) for i in range(n)], np.int64)
    order = np.lexsort((np.arange(n), keys))
    keys = keys[order]
    cell = cell[order]
    occupied = [set() for _ in range(D + 1)]
    for c in cell:
        for d in range(D + 1):
            occupied[d].add(tuple(int(v) >> (D - d) for v in c))
    levels = [[(0, 0, 0)]]
    for d in range(1, D + 1):
        parents = sorted(occupied[d - 1], key=lambda c: _key(*c, d - 1, D))
        levels.append([(2 * px + (s >> 2 & 1), 2 * py + (s >> 1 & 1), 2 * pz + (s & 1))
                       for (px, py, pz) in parents for s in range(8)])
    counts = [len(L) for L in levels]
    bases = np.concatenate([[0], np.cumsum(counts)[:-1]])
    index = [{c: i for i, c in enumerate(L)} for L in levels]
    M = sum(counts)
    out = dict(key=np.zeros(M, np.int64), pnum=np.zeros(M, np.int64), parent=np.full(M, -1, np.int64),
               children=np.full((M, 8), -1, np.int64), neighs=np.full((M, 27), -1, np.int64),
               pidx=np.zeros(M, np.int64), dnum=np.zeros(M, np.int64), didx=np.zeros(M, np.int64))
    point_counts = [dict() for _ in range(D + 1)]
    for c in cell:
        for d in range(D + 1):
            t = tuple(int(v) >> (D - d) for v in c)
            point_counts[d][t] = point_counts[d].get(t, 0) + 1
    leaf_slots = {}  # cell at depth D -> local leaf index
    for i, c in enumerate(levels[D]):
        leaf_slots[c] = i
    for d in range(D + 1):
        for i, c in enumerate(levels[d]):
            g = bases[d] + i
            out["key"][g] = _key(*c, d, D)
            out["pnum"][g] = point_counts[d].get(c, 0)
            if d > 0:
                out["parent"][g] = bases[d - 1] + index[d - 1][tuple(v >> 1 for v in c)]
            if d < D and c in occupied[d]:
                out["children"][g] = bases[d + 1] + index[d + 1][(2 * c[0], 2 * c[1], 2 * c[2])] + np.arange(8)
            for j in range(27):
                nb = (c[0] + j // 9 - 1, c[1] + (j // 3) % 3 - 1, c[2] + j % 3 - 1)
                if nb in index[d]:
                    out["neighs"][g, j] = bases[d] + index[d][nb]
            # leaf-level slots of the subtree (empty leaf slots count: dnum = 1 at depth D)
            if d == D:
                out["dnum"][g] = 1
                out["didx"][g] = i
            else:
                lo = [leaf_slots[s] for s in leaf_slots if tuple(v >> (D - d) for v in s) == c]
                out["dnum"][g] = len(lo)
    # pidx / didx: number of points / leaf slots before the node in key order at its depth
    for d in range(D + 1):
        sl = slice(bases[d], bases[d] + counts[d])
        out["pidx"][sl] = np.concatenate([[0], np.cumsum(out["pnum"][sl])[:-1]])
        if d < D:
            out["didx"][sl] = np.concatenate([[0], np.cumsum(out["dnum"][sl])[:-1]])
    return counts, bases, out


@pytest.mark.parametrize("seed,n,D", [(0, 50, 3), (1, 300, 4), (2, 800, 5), (3, 7, 4), (4, 200, 2)])
def test_oracle_octree_matches_brute_force(seed, n, D):
    rng = np.random.default_rng(seed)
    if seed == 3:  # clustered points: deep single-branch paths
        p = (0.3 + 0.01 * rng.random((n, 3))).astype(np.float32)
    else:
        p = rng.random((n, 3)).astype(np.float32)
    nrm = rng.standard_normal((n, 3)).astype(np.float32)
    o = orc.Oracle(D)
    o.set_points(p, nrm)
    o.build_octree()
    counts, bases = o.level_info()
    pts = o.read_buffer(0)
    bc, bb, ref = brute(pts, D)
    assert counts.tolist() == bc and bases.tolist() == bb.tolist()
    for name, which in (("key", 3), ("pidx", 4), ("pnum", 5), ("parent", 6), ("children", 7), ("neighs", 8),
                        ("didx", 9), ("dnum", 10)):
        got = o.read_buffer(which).astype(np.int64)
        assert np.array_equal(got, ref[name]), name
    # every sorted point lies in its leaf slot
    leaf = o.read_buffer(2)
    base_d = bases[D]
    pidx, pnum = o.read_buffer(4), o.read_buffer(5)
    for i in range(len(pts)):
        g = base_d + leaf[i]
        assert pidx[g] <= i < pidx[g] + pnum[g]


def test_normalisation_restates_reference():
    # bbox -> scale = 1.25 * max extent, centre = mid - scale/2 (main.cu:530-545)
    p = np.array([[1.0, 2.0, 3.0], [3.0, 2.5, 4.0], [2.0, 6.0, 3.5]], np.float32)
    n = np.array([[3.0, 0, 0], [0, 0, 0], [1, 1, 1]], np.float32)
    o = orc.Oracle(4)
    o.set_points(p, n)
    c, s = o.transform()
    assert s == np.float32(4.0) * np.float32(1.25)
    assert np.array_equal(c, (np.float32([2.0, 4.0, 3.5]) - np.float32(s) / 2).astype(np.float32))
    q = o.read_buffer(0)
    assert np.array_equal(q, ((p - c) / np.float32(s)).astype(np.float32))
    nn = o.read_buffer(1)
    assert np.allclose(np.linalg.norm(nn[[0, 2]], axis=1), 2 ** 5, rtol=1e-6)  # 2^(D+1)
    assert np.array_equal(nn[1], [0, 0, 0])  # zero normal stays zero
