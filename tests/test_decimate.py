"""Mesh decimation (EdgeCollapse / TriangleCollapse, Geometry.inl:50-254)
through the C ABI against the oracle's pure-Python restatement, bit for bit,
on marching-cubes meshes of the oracle pipeline and on perturbed grids; and
the invariants any result must meet (no degenerate triangle, indices in
range, every output vertex the mean of the input vertices merged into it, so
inside their bounding box).  The pass is host code in the library: each
collapse depends on every earlier one."""
import numpy as np
import pytest

import oracle as orc
from oracle.decimate_ref import decimate as decimate_ref


def _mc_mesh(D, n, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    p, nn = (0.5 + 0.3 * g).astype(np.float32), g.astype(np.float32)
    o = orc.Oracle(D)
    o.set_points(p, nn)
    o.build_octree()
    o.compute_divergence()
    o.solve(mode=0)
    o.isovalue()
    v, t = o.extract_leaf_mesh()
    o.close()
    return v, t


def _grid_mesh(k, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="ij")
    v = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], axis=1) / k
    v = (v + rng.normal(0, 0.25 / k, v.shape)).astype(np.float32)
    t = []
    for i in range(k):
        for j in range(k):
            a, b, c, d = i * (k + 1) + j, (i + 1) * (k + 1) + j, (i + 1) * (k + 1) + j + 1, i * (k + 1) + j + 1
            t += [[a, b, c], [a, c, d]]
    return v, np.array(t, np.int32)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("ratio", [0.6, 1.0, 2.5])
@pytest.mark.parametrize("src", ["mc_d5", "grid"])
def test_decimation_matches_oracle(prg, mode, ratio, src):
    v, t = _mc_mesh(5, 4000, 1) if src == "mc_d5" else _grid_mesh(40, 2)
    mesh = prg.Mesh(v, t, np.zeros(3, np.float32), 1.0, len(v), len(t))
    got = prg.decimate_mesh(mesh, mode, ratio)
    vr, tr = decimate_ref(v, t, mode, ratio)
    assert got.vertices.shape == vr.shape and got.triangles.shape == tr.shape
    assert np.array_equal(got.vertices.view(np.uint32), vr.view(np.uint32))
    assert np.array_equal(got.triangles, tr)
    # invariants
    tt = got.triangles
    assert np.all((tt >= 0) & (tt < len(got.vertices)))
    assert not np.any((tt[:, 0] == tt[:, 1]) | (tt[:, 0] == tt[:, 2]) | (tt[:, 1] == tt[:, 2]))
    lo, hi = v.min(0), v.max(0)
    assert np.all(got.vertices >= lo - 1e-6) and np.all(got.vertices <= hi + 1e-6)
    assert len(tt) <= len(t) and len(got.vertices) <= len(v)
    if ratio >= 1.0:
        assert len(tt) < len(t)  # thin triangles exist in both inputs


def test_decimation_of_equilateral_triangle_is_identity(prg):
    # sum of squared edges = 12/sqrt(3) * area exactly at ratio 1: not collapsed (strict >)
    v = np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.int32)
    mesh = prg.Mesh(v, t, np.zeros(3, np.float32), 1.0, 3, 1)
    for mode in (0, 1):
        got = prg.decimate_mesh(mesh, mode, 0.5)
        assert len(got.triangles) == 1 and np.array_equal(got.vertices, v)
        got = prg.decimate_mesh(mesh, mode, 4.0)
        assert len(got.triangles) == 0
        assert len(got.vertices) == (2 if mode == 0 else 1)


def test_decimation_rejects_bad_arguments(prg):
    v = np.zeros((3, 3), np.float32)
    mesh = prg.Mesh(v, np.array([[0, 1, 5]], np.int32), np.zeros(3, np.float32), 1.0, 3, 1)
    with pytest.raises(prg.PrgError):
        prg.decimate_mesh(mesh, 0, 1.0)
    mesh = prg.Mesh(v, np.array([[0, 1, 2]], np.int32), np.zeros(3, np.float32), 1.0, 3, 1)
    for mode, r in ((2, 1.0), (0, 0.0), (1, -1.0)):
        with pytest.raises(prg.PrgError):
            prg.decimate_mesh(mesh, mode, r)


def test_hand_worked_case(prg):
    """Two triangles worked by hand: a thin one [0, 2, 1] (visited first: the last
    triangle) and a fat one [0, 1, 3].  Ratio 0.8: the thin one collapses (sum of
    squared edges 1.52 * 0.8 > 12/sqrt(3) * area 0.05 = 0.346), the fat one does
    not (2.98 * 0.8 < 12/sqrt(3) * 0.3875 = 2.685 after the first merge).
    EdgeCollapse merges the first shortest edge (0, 2) into vertex 0 (mean
    (0.25, 0.05, 0)); TriangleCollapse merges 1 and 2 into 0 (mean (0.5, 0.1/3, 0)),
    which makes the fat triangle degenerate."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0.1, 0], [0.5, -1, 0]], np.float32)
    t = np.array([[0, 1, 3], [0, 2, 1]], np.int32)
    mesh = prg.Mesh(v, t, np.zeros(3, np.float32), 1.0, 4, 2)
    want_edge = (np.array([[0.25, 0.05, 0], [1, 0, 0], [0.5, -1, 0]], np.float32), np.array([[0, 1, 2]], np.int32))
    want_tri = (np.array([[0.5, np.float32(0.1) / np.float32(3), 0], [0.5, -1, 0]], np.float32), np.zeros((0, 3), np.int32))
    for mode, (wv, wt) in ((0, want_edge), (1, want_tri)):
        for got_v, got_t in (decimate_ref(v, t, mode, 0.8), (lambda m: (m.vertices, m.triangles))(prg.decimate_mesh(mesh, mode, 0.8))):
            assert np.allclose(got_v, wv, rtol=0, atol=1e-7), (mode, got_v)
            assert np.array_equal(got_t.reshape(-1, 3), wt)
