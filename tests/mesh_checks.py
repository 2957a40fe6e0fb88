"""Independent FP64 constructions used to pin the oracle (and the GPU) where
the reference ships no numbers: the exact B-spline basis, brute-force vector
field, divergence and implicit function, and a geometric check of marching-
cubes meshes.  Nothing here calls the oracle's arithmetic; the oracle (or the
GPU) supplies only the octree it was given and the values under test.

Reference semantics restated here (read from the reference, not run):
  - F: degree-2 B-spline on [-3/2, 3/2] with F(0) = 1 (PPolynomial.inl:385-411,
    main.cu:103,3310); node basis B_o(t) = F((t - c_o) / w_o) (normalize = 0).
  - vector field: V_i = sum_p F(2^D (c_i - p)) n'_p over the points of the 27
    neighbours of leaf-level node i (computeVectorField, main.cu:932-964).
  - divergence: b_i = sum over the leaf-level nodes k below the 27 neighbours of
    i of V_k . (dDot_x, dDot_y, dDot_z), dDot[node + res*leaf] = <B_node', B_leaf>
    (computeEncodedFinerNodesDivergence main.cu:1007-1060; FunctionData.inl:158-215).
  - implicit function: f(p) = sum over the ancestors a of p's leaf and their 27
    neighbours j of x_j B_j(p) (main.cu:1334-1381), i.e. the sum over every node
    whose basis support holds p.
"""
from __future__ import annotations

import re
from fractions import Fraction as Fr
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]

# exact F = (4/3) * quadratic B-spline on [-3/2, 3/2] (F(0) = 1): pieces (lo, hi, coefficients in x)
F_EXACT = [(Fr(-3, 2), Fr(-1, 2), [Fr(3, 2), Fr(2), Fr(2, 3)]),
           (Fr(-1, 2), Fr(1, 2), [Fr(1), Fr(0), Fr(-4, 3)]),
           (Fr(1, 2), Fr(3, 2), [Fr(3, 2), Fr(-2), Fr(2, 3)])]

EDGE_VERTEX = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 3), (1, 2), (4, 7), (5, 6), (0, 4), (1, 5), (3, 7), (2, 6)]


def F64(t):
    """F in FP64, vectorised."""
    a = np.abs(np.asarray(t, np.float64))
    return np.where(a <= 0.5, 1.0 - (4.0 / 3.0) * a * a, np.where(a < 1.5, (2.0 / 3.0) * (1.5 - a) ** 2, 0.0))


def _pmul(a, b):
    r = [Fr(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] += x * y
    return r


def _affine(p, a, b):  # p(a x + b)
    out, powk = [Fr(0)], [Fr(1)]
    for c in p:
        n = max(len(out), len(powk))
        out = [(out[i] if i < len(out) else 0) + (c * powk[i] if i < len(powk) else 0) for i in range(n)]
        powk = _pmul(powk, [b, a])
    return out


def basis_pieces(depth, off, deriv=False):
    """B_(depth, off) (or its derivative) as exact pieces (lo, hi, coefficients)."""
    w = Fr(1, 2 ** depth)
    c = (Fr(1, 2) + off) * w
    out = []
    for lo, hi, p in F_EXACT:
        q = [p[i] * i for i in range(1, len(p))] if deriv else p
        q = _affine(q, 1 / w, -c / w)
        if deriv:
            q = [v / w for v in q]
        out.append((c + lo * w, c + hi * w, q))
    return out


def inner(A, B):
    s = Fr(0)
    for a0, a1, p in A:
        for b0, b1, q in B:
            lo, hi = max(a0, b0), min(a1, b1)
            if lo < hi:
                pq = _pmul(p, q)
                s += sum(cf * (hi ** (i + 1) - lo ** (i + 1)) / (i + 1) for i, cf in enumerate(pq))
    return s


def exact_entry(which, da, oa, db, ob):
    """<B_a, B_b>, <B_a', B_b>, <B_a', B_b'> - the quantities the reference tabulates."""
    if which == 0:
        return float(inner(basis_pieces(da, oa), basis_pieces(db, ob)))
    if which == 1:
        return float(inner(basis_pieces(da, oa, True), basis_pieces(db, ob)))
    return float(inner(basis_pieces(da, oa, True), basis_pieces(db, ob, True)))


def mc_table():
    txt = (ROOT / "poissonrecon_gpu_amd" / "csrc" / "mc_triangle_table.inc").read_text()
    body = txt.split("#define PRG_MC_TRIANGLE_TABLE_DATA", 1)[1]
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    assert len(vals) == 256 * 16
    return np.array(vals).reshape(256, 16)


def square_xyz(q):
    """Corner q of the reference's square numbering as (x, y, z) bits (MarchingCubes.cuh:693-706)."""
    y = (q >> 1) & 1
    return ((q & 1) ^ y, y, q >> 2)


# ---- octree access -------------------------------------------------------------------------
def spread3(v):
    v = np.asarray(v, np.uint64) & np.uint64(0x3FF)
    v = (v | (v << np.uint64(16))) & np.uint64(0xFF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def compact3(v):
    v = np.asarray(v, np.uint64) & np.uint64(0x09249249)
    v = (v ^ (v >> np.uint64(2))) & np.uint64(0x030C30C3)
    v = (v ^ (v >> np.uint64(4))) & np.uint64(0x0300F00F)
    v = (v ^ (v >> np.uint64(8))) & np.uint64(0xFF0000FF)
    v = (v ^ (v >> np.uint64(16))) & np.uint64(0x3FF)
    return v.astype(np.int64)


def morton(cx, cy, cz):
    return (spread3(cx) << np.uint64(2)) | (spread3(cy) << np.uint64(1)) | spread3(cz)


class Levels:
    """Per-depth node cells (from the D-level keys) and a coefficient per node."""

    def __init__(self, keys, counts, bases, D, coef=None):
        self.D = D
        self.depth = []
        for d in range(D + 1):
            k = np.asarray(keys[bases[d]:bases[d] + counts[d]], np.uint64)
            code = k >> np.uint64(3 * (D - d))  # depth-d Morton code
            cx = compact3(k >> np.uint64(2)) >> (D - d)
            cy = compact3(k >> np.uint64(1)) >> (D - d)
            cz = compact3(k) >> (D - d)
            c = None if coef is None else np.asarray(coef[bases[d]:bases[d] + counts[d]], np.float64)
            self.depth.append((code, cx, cy, cz, c, int(bases[d])))

    def lookup(self, d, cx, cy, cz):
        """Global node index of the depth-d cell (cx, cy, cz), or -1."""
        code, *_, base = self.depth[d]
        W = 1 << d
        ok = (cx >= 0) & (cy >= 0) & (cz >= 0) & (cx < W) & (cy < W) & (cz < W)
        q = morton(np.clip(cx, 0, W - 1), np.clip(cy, 0, W - 1), np.clip(cz, 0, W - 1))
        pos = np.searchsorted(code, q)
        pos = np.minimum(pos, len(code) - 1)
        hit = ok & (len(code) > 0) & (code[pos] == q)
        return np.where(hit, base + pos, -1)


def basis_float_error(d):
    """Bound on the error of the reference's float evaluation of one depth-d basis
    function (ConfirmedPPolynomial::value of float coefficients shifted to the
    node, ConfirmedPPolynomial.cuh:37-91): the coefficients grow like 4^d and a
    few roundings of 2^-24 act on them.  The oracle's evaluation stays below
    4.6 * 4^d * 2^-24 at every depth 0..10 (measured over random offsets and
    arguments); the bound takes twice that."""
    return (8.0 * 4.0 ** d + 8.0) * 2.0 ** -24


def implicit_f64(lv: Levels, x, pts):
    """f(p) = sum over every node o of x_o B_o(p), FP64 (exact spline).  Also
    returns, per point, a bound on the float error of the reference's
    evaluation: sum over the nodes whose closed support holds p of
    |x_o| * 3 * basis_float_error(depth) (one erring factor per axis), plus
    2^-20 times the sum of |terms| for the products and the running sum."""
    pts = np.asarray(pts, np.float64)
    x = np.asarray(x, np.float64)
    f = np.zeros(len(pts))
    absf = np.zeros(len(pts))
    err = np.zeros(len(pts))
    for d in range(lv.D + 1):
        W = 1 << d
        u = pts * W
        cell = np.clip(np.floor(u).astype(np.int64), 0, W - 1)
        eb = basis_float_error(d)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    cx, cy, cz = cell[:, 0] + dx, cell[:, 1] + dy, cell[:, 2] + dz
                    idx = lv.lookup(d, cx, cy, cz)
                    rx, ry, rz = u[:, 0] - cx - 0.5, u[:, 1] - cy - 0.5, u[:, 2] - cz - 0.5
                    b = F64(rx) * F64(ry) * F64(rz)
                    xv = np.where(idx >= 0, x[np.maximum(idx, 0)], 0.0)
                    t = xv * b
                    f += t
                    absf += np.abs(t)
                    inside = (np.abs(rx) <= 1.5) & (np.abs(ry) <= 1.5) & (np.abs(rz) <= 1.5)
                    err += np.where(inside, np.abs(xv) * 3.0 * eb, 0.0)
    return f, absf, err + 2.0 ** -20 * absf


# ---- geometric check of a marching-cubes mesh ---------------------------------------------
def check_mc_mesh(V, T, D, fval, iso, sample=None, rng=None, value_tol=1.0, pos_tol=0.02):
    """Every triangle lies in one depth-D cell, on the edges and in the order the
    triangle table gives for that cell's cube index, and every vertex sits where
    linear interpolation of the corner values puts it.  Corner values come from
    `fval` (an independent FP64 evaluation of f returning value, sum of |terms|
    and a bound on the float error of the evaluation under test); cells with a
    corner within value_tol error bounds of the isovalue are skipped (their sign
    is decided by float rounding).  Returns (checked triangles, skipped)."""
    tab = mc_table()
    V = np.asarray(V, np.float64)
    T = np.asarray(T)
    if sample is not None and len(T) > sample:
        T = T[(rng or np.random.default_rng(0)).choice(len(T), sample, replace=False)]
    W = float(1 << D)
    g = V * W
    isint = g == np.round(g)
    # vertex edges: the axis that is not on the grid (a vertex exactly at a corner has none)
    nint = isint.sum(axis=1)
    assert np.all(nint >= 2), "a vertex is off the depth-D grid lines"
    # candidate cells of each triangle: per axis, intersect the vertices' cell ranges
    lo = np.full((len(T), 3), -1 << 40, np.int64)
    hi = np.full((len(T), 3), 1 << 40, np.int64)
    for k in range(3):
        gv = g[T[:, k]]
        iv = isint[T[:, k]]
        fl = np.floor(gv).astype(np.int64)
        vlo = np.where(iv, fl - 1, fl)
        vhi = fl
        lo = np.maximum(lo, vlo)
        hi = np.minimum(hi, vhi)
    assert np.all(lo <= hi), "a triangle's vertices share no cell"
    checked = skipped = 0
    corner_off = np.array([square_xyz(q) for q in range(8)])
    # candidate cells: lo, and hi where a triangle's vertices share a face plane
    cand = [lo, hi]
    corners = (cand[0][:, None, :] + corner_off[None, :, :]).reshape(-1, 3) / W
    fv, _, fa = fval(corners)
    fv = fv.reshape(-1, 8) - iso
    fa = fa.reshape(-1, 8)
    corners2 = (cand[1][:, None, :] + corner_off[None, :, :]).reshape(-1, 3) / W
    gv2, _, ga2 = fval(corners2)
    gv2 = gv2.reshape(-1, 8) - iso
    ga2 = ga2.reshape(-1, 8)
    for t in range(len(T)):
        ok = False
        amb = False
        for ci, (vals, mags) in enumerate(((fv[t], fa[t]), (gv2[t], ga2[t]))):
            cell = cand[ci][t]
            if np.any(np.abs(vals) <= value_tol * np.maximum(mags, 1e-30)):
                amb = True
                continue
            cube = sum(1 << q for q in range(8) if vals[q] < 0)
            # edges of this triangle's vertices in the cell
            es = []
            for k in range(3):
                v = g[T[t, k]]
                rel = v - cell
                axes = [a for a in range(3) if not isint[T[t, k], a]]
                if len(axes) != 1:
                    es = None
                    break
                ax = axes[0]
                if not (0 < rel[ax] < 1):
                    es = None
                    break
                side = tuple(int(round(rel[a])) for a in range(3))
                if any(side[a] not in (0, 1) for a in range(3) if a != ax):
                    es = None
                    break
                e = None
                for j, (qa, qb) in enumerate(EDGE_VERTEX):
                    pa, pb = square_xyz(qa), square_xyz(qb)
                    if (j >> 2) == ax and all(pa[a] == side[a] for a in range(3) if a != ax):
                        e = j
                es.append(e)
            if es is None:
                continue
            row = tab[cube]
            n = int(np.argmax(row < 0)) if (row < 0).any() else 16
            tris = [tuple(row[i:i + 3]) for i in range(0, n, 3)]
            if tuple(es) not in tris:
                continue
            # interpolation: t = f1 / (f1 - f2) along the edge, in cell units
            good = True
            for k, e in enumerate(es):
                qa, qb = EDGE_VERTEX[e]
                f1, f2 = vals[qa], vals[qb]
                ax = e >> 2
                pa, pb = square_xyz(qa), square_xyz(qb)
                want = pa[ax] + (pb[ax] - pa[ax]) * f1 / (f1 - f2)
                have = g[T[t, k], ax] - cell[ax]
                # the values under test carry float error up to value_tol * |terms|
                slack = pos_tol + 2 * value_tol * (mags[qa] + mags[qb]) / abs(f1 - f2)
                if abs(want - have) > slack:
                    good = False
            if good:
                ok = True
                break
        if ok:
            checked += 1
        elif amb:
            skipped += 1
        else:
            raise AssertionError(f"triangle {t} ({T[t].tolist()}) does not match the cube of its cell")
    return checked, skipped


def check_mc_structure(V, T, D, sample=None, rng=None):
    """Value-free check of a marching-cubes mesh, for depths where the float
    evaluation under test errs by more than its distance to the isovalue
    (the reference's depth-d basis values carry errors of about 4^d 2^-22):
    every vertex lies on a depth-D grid edge, and every triangle's three
    vertices lie on edges of one depth-D cell, forming a triangle the
    triangle table holds for some cube index.  Returns the triangles checked."""
    tab = mc_table()
    tris = set()
    for cube in range(256):
        row = tab[cube]
        n = int(np.argmax(row < 0)) if (row < 0).any() else 16
        tris.update(tuple(row[i:i + 3]) for i in range(0, n, 3))
    V = np.asarray(V, np.float64)
    T = np.asarray(T)
    if sample is not None and len(T) > sample:
        T = T[(rng or np.random.default_rng(0)).choice(len(T), sample, replace=False)]
    g = V * float(1 << D)
    isint = g == np.round(g)
    assert np.all(isint.sum(axis=1) >= 2), "a vertex is off the depth-D grid lines"
    for t in T:
        lo = np.full(3, -(1 << 40))
        hi = np.full(3, 1 << 40)
        for k in range(3):
            fl = np.floor(g[t[k]]).astype(np.int64)
            lo = np.maximum(lo, np.where(isint[t[k]], fl - 1, fl))
            hi = np.minimum(hi, fl)
        assert np.all(lo <= hi), f"triangle {t.tolist()}: vertices share no cell"
        found = False
        cells = {tuple(lo), tuple(hi)} | {(a, b, c) for a in (lo[0], hi[0]) for b in (lo[1], hi[1])
                                          for c in (lo[2], hi[2])}
        for cell in cells:
            cand = []  # per vertex: the cell edges it can lie on
            for k in range(3):
                rel = g[t[k]] - np.array(cell)
                if np.any(rel < 0) or np.any(rel > 1):
                    cand = None
                    break
                axes = [a for a in range(3) if not isint[t[k], a]]
                # a vertex interpolated to within rounding of a corner sits on the
                # corner: any of the corner's three cell edges
                axes = axes if axes else [0, 1, 2]
                e = [j for j, (qa, qb) in enumerate(EDGE_VERTEX) for ax in axes
                     if (j >> 2) == ax and all(square_xyz(qa)[a] == int(round(rel[a])) for a in range(3) if a != ax)]
                if not e:
                    cand = None
                    break
                cand.append(e)
            if cand is None:
                continue
            if any((e0, e1, e2) in tris for e0 in cand[0] for e1 in cand[1] for e2 in cand[2]):
                found = True
                break
        assert found, (f"triangle {t.tolist()} is no triangle-table triangle of a depth-{D} cell: "
                       f"{[g[t[k]].tolist() for k in range(3)]}")
    return len(T)
