"""Many clouds in one pass (prg_reconstruct_batch): each cloud's mesh equals
the mesh of the same cloud reconstructed alone (the pipeline is
deterministic), whatever the interleaving of the workers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_matches_single_reconstructions(prg):
    print("\n[batch: 6 clouds, 3 workers]", flush=True)
    clouds = [prg.synth_sphere(35947, 1), prg.synth_torus(60000, 2), prg.synth_superquadrics(90000, 2, 0.001, 9),
              prg.synth_sphere(12000, 4), prg.synth_superquadrics(30000, 4, 0.001, 12), prg.synth_torus(20000, 6)]
    depths = [8, 9, 9, 7, 8, 7]
    st, res = prg.reconstruct_batch(clouds, depths, concurrency=3, return_meshes=True)
    assert st == prg.OK
    for (p, n), d, r in zip(clouds, depths, res):
        assert r["status"] == prg.OK
        assert r["max_true_rel_residual"] <= 2e-6
        with prg.Reconstructor(max_depth=d) as g:
            g.set_points(p, n)
            g.build_octree()
            g.compute_divergence()
            g.solve()
            g.isovalue()
            m = g.extract_mesh()
        assert np.array_equal(m.triangles, r["mesh_triangles"])
        assert np.array_equal(m.vertices.view(np.uint32), r["mesh_vertices"].view(np.uint32))
    print("meshes:", [(r["vertices"], r["triangles"]) for r in res])


def test_batch_reports_bad_items(prg):
    p, n = prg.synth_sphere(1000, 1)
    st, res = prg.reconstruct_batch([(p, n), (p, n)], [6, 12], concurrency=2)
    assert st != prg.OK
    assert res[0]["status"] == prg.OK and res[1]["status"] == prg.E_ARG


def _mesh(g, p, n):
    g.set_points(p, n)
    g.build_octree()
    g.compute_divergence()
    g.solve()
    g.isovalue()
    return g.extract_mesh()


def test_depth_switch_in_one_context(prg):
    """prg_set_max_depth: a context that changes depth (deeper, then shallower,
    then back) gives each cloud the mesh a fresh context gives it."""
    print("\n[one context: depths 9, 7, 10, 8]", flush=True)
    cases = [(prg.synth_torus(60000, 2), 9), (prg.synth_sphere(12000, 4), 7),
             (prg.synth_superquadrics(150000, 3, 0.001, 21), 10), (prg.synth_superquadrics(40000, 2, 0.001, 22), 8)]
    with prg.Reconstructor(max_depth=6) as shared:
        for (p, n), d in cases:
            shared.set_max_depth(d)
            ms = _mesh(shared, p, n)
            with prg.Reconstructor(max_depth=d) as fresh:
                mf = _mesh(fresh, p, n)
            assert np.array_equal(ms.triangles, mf.triangles)
            assert np.array_equal(ms.vertices.view(np.uint32), mf.vertices.view(np.uint32))
        with pytest.raises(prg.PrgError):
            shared.set_max_depth(11)
