"""Marching-cubes table consistency and the oracle's solvers against dense algebra.

- MC: for every cube configuration, the edges used by the triangle table are
  exactly the edges whose two square-numbered corners (edgeVertex,
  MarchingCubes.cuh:693-706) differ in sign; every edge of the triangle fan
  appears in an even number of triangles (a closed surface patch per cube);
  the table counts (MarchingCubes.cuh:653-691) follow from the -1 terminator.
- CG: the oracle's FP32 CG (the build's variant and the reference's) and its
  FP64 CG agree with a dense numpy solve of the same stencil system.
"""
import re
from pathlib import Path

import numpy as np
import pytest

import oracle as orc

ROOT = Path(__file__).resolve().parents[1]
EDGE_VERTEX = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 3), (1, 2), (4, 7), (5, 6), (0, 4), (1, 5), (3, 7), (2, 6)]


def _table():
    txt = (ROOT / "poissonrecon_gpu_amd" / "csrc" / "mc_triangle_table.inc").read_text()
    body = txt.split("#define PRG_MC_TRIANGLE_TABLE_DATA", 1)[1]
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    assert len(vals) == 256 * 16
    return np.array(vals).reshape(256, 16)


def _square_xyz(q):
    y = (q >> 1) & 1
    return ((q & 1) ^ y, y, q >> 2)


def test_edge_vertex_pairs_are_cube_edges_along_their_orientation():
    for e, (a, b) in enumerate(EDGE_VERTEX):
        pa, pb = _square_xyz(a), _square_xyz(b)
        diff = [i for i in range(3) if pa[i] != pb[i]]
        assert diff == [e >> 2], e


def test_triangle_table_matches_sign_changes():
    T = _table()
    for cube in range(256):
        row = T[cube]
        n = int(np.argmax(row < 0)) if (row < 0).any() else 16
        assert n % 3 == 0 and np.all(row[n:] == -1)
        used = set(row[:n].tolist())
        crossing = {e for e, (a, b) in enumerate(EDGE_VERTEX) if ((cube >> a) & 1) != ((cube >> b) & 1)}
        assert used == crossing, cube
        # triangle edges pair up inside the cube's faces: each segment shared by 0 or 2 triangles
        # unless it lies on a cube face (then it is shared with the neighbour cube)
        tris = row[:n].reshape(-1, 3)
        assert len({tuple(sorted(t)) for t in tris}) == len(tris)
    # complementary configurations use the same edge sets
    for cube in range(256):
        a = set(T[cube][T[cube] >= 0].tolist())
        b = set(T[255 - cube][T[255 - cube] >= 0].tolist())
        assert a == b


def _dense_system(o, d):
    counts, bases = o.level_info()
    n = int(counts[d])
    neighs = o.read_buffer(8)[bases[d]: bases[d] + n]
    st = orc.fs_stencil(o.max_depth, d)
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(27):
            nb = neighs[i, j]
            if nb != -1:
                A[i, nb - bases[d]] += float(st[j])
    b = o.read_buffer(12)[bases[d]: bases[d] + n].astype(np.float64)
    return A, b, slice(int(bases[d]), int(bases[d] + n))


@pytest.fixture(scope="module")
def sphere_state():
    import poissonrecon_gpu_amd as prg
    p, n = prg.synth_sphere(6000, 11)
    o = orc.Oracle(5)
    o.set_points(p, n)
    o.build_octree()
    o.compute_divergence()
    return o


def test_stencil_matrix_is_symmetric_positive_definite(sphere_state):
    for d in (3, 4):
        A, _, _ = _dense_system(sphere_state, d)
        assert np.array_equal(A, A.T)
        assert np.linalg.eigvalsh(A).min() > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_cg_matches_dense_solve(sphere_state, mode):
    o = sphere_state
    if mode == 0:
        it, rr = o.solve(mode=0, tol=1e-6)
    else:
        it, rr = o.solve(mode=1, tol=1e-7)  # the reference's absolute criterion
    x = o.read_buffer(13).astype(np.float64)
    for d in range(2, 6):
        A, b, sl = _dense_system(o, d)
        xs = np.linalg.solve(A, b)
        err = np.linalg.norm(x[sl] - xs) / np.linalg.norm(xs)
        cond = np.linalg.cond(A) if d <= 4 else 30.0
        assert cond < 40  # the truncated 27-point operator is well conditioned
        assert err <= cond * max(rr[d], 1e-7) * 4 + 1e-5, (d, err, rr[d])
    x64 = o.solve_fp64(tol=1e-13)
    for d in range(2, 6):
        A, b, sl = _dense_system(o, d)
        xs = np.linalg.solve(A, b)
        assert np.linalg.norm(x64[sl] - xs) <= 1e-9 * np.linalg.norm(xs)


def test_fixed_iterations_stop_on_exact_solution(sphere_state):
    o = sphere_state
    it, rr = o.solve(mode=0, fixed_iters=30)
    assert it[0] <= 1 and np.isfinite(o.read_buffer(13)).all()
