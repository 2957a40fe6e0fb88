"""Many clouds (the 64-cloud workload: sizes log-uniform in [10k, 4M], seed
5, unions of 1-4 superquadrics, depth clamp(7 + round(log2(N/10k)/2), 7, 11)
capped at 10) through prg_reconstruct_batch, on a subset at full per-cloud
size: the four smallest and the two largest clouds together in one batch.

Every cloud must converge (largest per-depth true relative residual <= 2e-6)
and give a mesh.  The small clouds are checked against the CPU oracle's own
pipeline (own solve, own isovalue): symmetric Hausdorff distance of the
leaf-level meshes at most 1e-4 of the bounding-box diagonal, the acceptance
bound used for the headline configurations; and each batch mesh must equal
the mesh of the same cloud reconstructed alone, bit for bit.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu


def _workload(prg):
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(4e6), 64)).astype(np.int64)
    order = np.argsort(sizes)
    pick = list(order[:4]) + list(order[-2:])
    clouds, depths = [], []
    for k in pick:
        n = int(sizes[k])
        clouds.append(prg.synth_superquadrics(n, 1 + int(k) % 4, 0.001, 5000 + int(k)))
        depths.append(min(10, int(np.clip(7 + round(np.log2(n / 1e4) / 2), 7, 11))))
    return clouds, depths


def _hausdorff_ok(v1, t1, v2, t2, pts):
    from scipy.spatial import cKDTree

    def samples(v, t):
        v = v.astype(np.float64)
        return np.concatenate([v, v[t].mean(axis=1)])

    a, b = samples(v1, t1), samples(v2, t2)
    diag = float(np.linalg.norm(pts.max(0) - pts.min(0)))
    h = max(cKDTree(b).query(a)[0].max(), cKDTree(a).query(b)[0].max())
    return h, diag


def test_many_clouds_subset(prg):
    clouds, depths = _workload(prg)
    print(f"\n[many clouds: sizes {[len(p) for p, _ in clouds]}, depths {depths}]", flush=True)
    st, res = prg.reconstruct_batch(clouds, depths, concurrency=4, return_meshes=True)
    assert st == prg.OK
    for r in res:
        assert r["status"] == prg.OK
        assert r["max_true_rel_residual"] <= 2e-6
        assert r["vertices"] > 0 and r["triangles"] > 0
    print("meshes:", [(r["vertices"], r["triangles"], round(r["ms"], 1)) for r in res])
    for (p, n), d, r in list(zip(clouds, depths, res))[:4]:
        # the same cloud alone: the same mesh
        with prg.Reconstructor(max_depth=d) as g:
            g.set_points(p, n)
            g.build_octree()
            g.compute_divergence()
            g.solve()
            g.isovalue()
            m = g.extract_mesh()
            leaf = g.extract_mesh(subdivide=False)
            pts = g.read_buffer(prg.BUF_POINTS).astype(np.float64)
        assert np.array_equal(m.triangles, r["mesh_triangles"])
        assert np.array_equal(m.vertices.view(np.uint32), r["mesh_vertices"].view(np.uint32))
        # the oracle's own pipeline
        o = orc.Oracle(d)
        try:
            o.set_points(p, n)
            o.build_octree()
            o.compute_divergence()
            o.solve(mode=0)
            o.isovalue()
            vo, to = o.extract_leaf_mesh()
        finally:
            o.close()
        h, diag = _hausdorff_ok(leaf.vertices, leaf.triangles, vo, to, pts)
        print(f"N={len(p)} D={d}: symmetric Hausdorff {h / diag:.2e} of the diagonal")
        assert h <= 1e-4 * diag
