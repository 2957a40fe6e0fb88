"""prg_extract_mesh_into: the mesh written into caller-owned host buffers
(the leaf-level batch's copy overlaps the subdivision's kernels) equals the
mesh of prg_extract_mesh bit for bit, with and without the subdivision; a
buffer too small for the whole mesh is an error that leaves the context
usable."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _into(prg, g, subdivide, vcap, tcap):
    v = np.full(max(vcap, 1), np.nan, np.float32)
    t = np.full(max(tcap, 1), -7, np.int32)
    m = prg.PrgMesh()
    st = prg.lib().prg_extract_mesh_into(g._h, int(subdivide), v.ctypes.data_as(C.c_void_p), vcap,
                                        t.ctypes.data_as(C.c_void_p), tcap, C.byref(m))
    return st, m, v, t


@pytest.mark.parametrize("subdivide", [0, 1])
def test_mesh_into_caller_buffers(prg, subdivide):
    p, n = prg.synth_superquadrics(80000, 3, 0.001, 3)
    print(f"\n[mesh into caller buffers: N={len(p)} D=9 subdivide={subdivide}]", flush=True)
    with prg.Reconstructor(max_depth=9) as g:
        g.set_points(p, n)
        g.build_octree()
        g.compute_divergence()
        g.solve()
        g.isovalue()
        ref = g.extract_mesh(subdivide=bool(subdivide))
        nv, nt = len(ref.vertices), len(ref.triangles)
        st, m, v, t = _into(prg, g, subdivide, 3 * nv + 30, 3 * nt + 30)
        assert st == prg.OK
        assert (m.n_vertices, m.n_triangles) == (nv, nt)
        assert (m.n_leaf_vertices, m.n_leaf_triangles) == (ref.n_leaf_vertices, ref.n_leaf_triangles)
        assert np.array_equal(v[: 3 * nv].view(np.uint32), ref.vertices.reshape(-1).view(np.uint32))
        assert np.array_equal(t[: 3 * nt], ref.triangles.reshape(-1))
        if subdivide:
            assert nv > ref.n_leaf_vertices
            # room for the leaf-level batch only: an error, and the context still works
            st2, _, _, _ = _into(prg, g, 1, 3 * ref.n_leaf_vertices, 3 * ref.n_leaf_triangles)
            assert st2 != prg.OK
            st3, m3, v3, t3 = _into(prg, g, 1, 3 * nv, 3 * nt)
            assert st3 == prg.OK and m3.n_vertices == nv
            assert np.array_equal(t3[: 3 * nt], ref.triangles.reshape(-1))
