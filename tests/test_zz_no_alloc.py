"""prg_weld_mesh_noalloc (the weld with every device array in the library's
fixed per-GPU array) against prg_weld_mesh, through the C ABI: two calls in a
row, the second on a smaller mesh (the array keeps the first call's contents),
bit for bit; a mesh beyond the array's limits is refused and left unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mesh_with_duplicates(prg, nv, nt, dup, seed):
    rng = np.random.default_rng(seed)
    v = rng.random((nv, 3)).astype(np.float32)
    k = int(dup * nv)
    src = rng.integers(0, nv, k)
    dst = rng.integers(0, nv, k)
    v[dst] = v[src]
    t = rng.integers(0, nv, (nt, 3)).astype(np.int32)
    return prg.Mesh(v, t, np.zeros(3, np.float32), 1.0, nv // 2, nt // 2)


def test_noalloc_weld_matches_weld(prg):
    vmax, tmax = 1 << 19, 1 << 20  # PRG_WELD_NOALLOC_MAX_VERTICES / _TRIANGLES
    with prg.Reconstructor(max_depth=6) as g:
        for nv, nt, dup, seed in ((vmax, tmax, 0.3, 1), (70000, 90000, 0.5, 2)):
            print(f"\n[no-alloc weld: {nv} vertices, {nt} triangles]", flush=True)
            m = _mesh_with_duplicates(prg, nv, nt, dup, seed)
            a, ka = g.weld_mesh(m)
            b, kb = g.weld_mesh(m, noalloc=True)
            assert ka == kb and ka > 0
            assert (a.n_leaf_vertices, a.n_leaf_triangles) == (b.n_leaf_vertices, b.n_leaf_triangles)
            assert np.array_equal(a.vertices.view(np.uint32), b.vertices.view(np.uint32))
            assert np.array_equal(a.triangles, b.triangles)
        # beyond the fixed array: refused, mesh untouched
        m = _mesh_with_duplicates(prg, vmax + 1, 10, 0.1, 3)
        v0, t0 = m.vertices.copy(), m.triangles.copy()
        with pytest.raises(prg.PrgError) as e:
            g.weld_mesh(m, noalloc=True)
        assert e.value.status == prg.E_ARG
        assert np.array_equal(m.vertices, v0) and np.array_equal(m.triangles, t0)
