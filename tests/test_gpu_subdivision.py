"""Full mesh with coarse-leaf subdivision: GPU (HIP, through the C ABI)
versus the CPU oracle, bit-exact.

The oracle restates the reference's subdivision with its own data
structures (explicit virtual subtrees per group, neighbours through the
parents with redirected child pointers, owners found by distance tests);
the GPU decides the same facts from keys alone (see csrc/subdivide.hip), so
agreement checks one formulation against the other.  Both sides get the
same solution and isovalue.  The cases are sized so that the oracle's
virtual trees stay within seconds.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu


def _workload(name, prg):
    if name == "sphere_c1_d6":
        return prg.synth_sphere(35947, 1), 6
    if name == "sphere_c1_d8":
        return prg.synth_sphere(35947, 1), 8
    if name == "torus_d7":
        return prg.synth_torus(60000, 2), 7
    if name == "superquadrics_d7":
        return prg.synth_superquadrics(80000, 3, 0.001, 3), 7
    if name == "sparse_d5":
        rng = np.random.default_rng(11)
        p = (rng.random((300, 3)) * 0.6 + 0.2).astype(np.float32)
        n = rng.standard_normal((300, 3)).astype(np.float32)
        return (p, n), 5
    if name == "two_points_d3":
        return (np.array([[0.2, 0.7, 0.4], [0.6, 0.1, 0.9]], np.float32),
                np.array([[0, 0, 1], [1, 0, 0]], np.float32)), 3
    raise KeyError(name)


CASES = ["sphere_c1_d6", "torus_d7", "superquadrics_d7", "sparse_d5", "two_points_d3", "sphere_c1_d8"]


@pytest.mark.parametrize("name", CASES)
def test_full_mesh_bit_exact(name, prg):
    (p, n), D = _workload(name, prg)
    print(f"\n[case {name}: N={len(p)} D={D}]", flush=True)
    o = orc.Oracle(D)
    g = prg.Reconstructor(max_depth=D)
    try:
        o.set_points(p, n)
        o.build_octree()
        o.compute_divergence()
        o.solve(mode=0)
        iso = o.isovalue()
        x = o.read_buffer(prg.BUF_SOLUTION)
        vl, tl = o.extract_leaf_mesh()
        vo, to = o.extract_subdivision()
        roots_o, leaves_o = o.subdivision_info()

        g.set_points(p, n)
        g.build_octree()
        g.compute_divergence()
        g.write_buffer(prg.BUF_SOLUTION, x)
        g.set_isovalue(iso)
        m = g.extract_mesh(subdivide=True)
        roots_g, leaves_g = g.subdivision_info()
        print(f"leaf V={len(vl)} T={len(tl)}; full V={len(vo)} T={len(to)}; roots {roots_o}, virtual leaves {leaves_o}")
        assert (roots_g, leaves_g) == (roots_o, leaves_o), "subdivided coarse leaves differ"
        assert (m.n_leaf_vertices, m.n_leaf_triangles) == (len(vl), len(tl))
        assert m.vertices.shape == vo.shape and m.triangles.shape == to.shape
        assert np.array_equal(m.triangles, to), "triangle connectivity differs"
        assert np.array_equal(np.nan_to_num(m.vertices).view(np.uint32), np.nan_to_num(vo).view(np.uint32))
    finally:
        g.close()
        o.close()
