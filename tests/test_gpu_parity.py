"""GPU (HIP, through the C ABI) versus the CPU oracle, stage by stage.

Each stage gets identical inputs on both sides (earlier stages' oracle
outputs are injected into the GPU context where a floating-point reduction
order could differ), so every comparison isolates one stage:
  octree, neighbours, sorted points: bit-exact
  vector field (float, same op order):  bit-exact
  divergence (double sums, other order): rel L2 <= 1e-5
  CG solution (same b, fixed iterations): rel L2 <= 1e-5
  implicit values at samples / corners (same x): bit-exact
  leaf-level mesh (same x and iso): bit-exact vertices and triangles
  mesh from the device-resident isovalue: bit-exact with the host-set one
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

BUFS_OCTREE = ["BUF_POINTS", "BUF_NORMALS", "BUF_POINT_LEAF", "BUF_KEY", "BUF_PIDX", "BUF_PNUM", "BUF_PARENT",
               "BUF_CHILDREN", "BUF_NEIGHS", "BUF_DIDX", "BUF_DNUM"]


def _edge_points():
    # coordinates on cell boundaries exercise the strict '>' of encodePoint
    rng = np.random.default_rng(7)
    p = rng.random((400, 3), dtype=np.float32) * 0.5 + 0.25
    p[::5, 0] = 0.5
    p[1::7, 1] = 0.375
    p[2::11, 2] = 0.625
    n = rng.standard_normal((400, 3)).astype(np.float32)
    n[3] = 0.0  # zero-length normal: not normalised (main.cu:561-568)
    return p, n


def _workload(name, prg):
    if name == "sphere_c1":
        return prg.synth_sphere(35947, 1), 8
    if name == "torus_d9":
        return prg.synth_torus(60000, 2), 9
    if name == "superquadrics_d9":
        return prg.synth_superquadrics(80000, 3, 0.001, 3), 9
    if name == "edge_d5":
        return _edge_points(), 5
    if name == "two_points_d3":
        return (np.array([[0.2, 0.7, 0.4], [0.6, 0.1, 0.9]], np.float32),
                np.array([[0, 0, 1], [1, 0, 0]], np.float32)), 3
    raise KeyError(name)


CASES = ["sphere_c1", "torus_d9", "superquadrics_d9", "edge_d5", "two_points_d3"]


@pytest.fixture(scope="module", params=CASES)
def pair(request, prg):
    (p, n), D = _workload(request.param, prg)
    print(f"\n[case {request.param}: N={len(p)} D={D}]", flush=True)
    g = prg.Reconstructor(max_depth=D, cg_fixed_iters=0)
    o = orc.Oracle(D)
    g.set_points(p, n)
    o.set_points(p, n)
    g.build_octree()
    o.build_octree()
    yield request.param, D, g, o
    g.close()
    o.close()


def test_transform(pair):
    _, _, g, o = pair
    cg, sg = g.transform()
    co, so = o.transform()
    assert np.array_equal(cg, co) and sg == so


def test_octree_bit_exact(pair, prg):
    name, D, g, o = pair
    cg, bg = g.level_info()
    co, bo = o.level_info()
    assert cg.tolist() == co.tolist() and bg.tolist() == bo.tolist()
    for b in BUFS_OCTREE:
        which = getattr(prg, b)
        a, r = g.read_buffer(which), o.read_buffer(which)
        assert a.shape == r.shape, b
        if a.dtype.kind == "f":
            assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), b
        else:
            bad = np.flatnonzero((a != r).reshape(len(a), -1).any(axis=1)) if a.ndim > 1 else np.flatnonzero(a != r)
            assert bad.size == 0, f"{b}: {bad.size} mismatching rows, first {bad[:5]}"


def test_vector_field_and_divergence(pair, prg):
    name, D, g, o = pair
    g.compute_divergence()
    o.compute_divergence()
    vg, vo = g.read_buffer(prg.BUF_VECTOR_FIELD), o.read_buffer(prg.BUF_VECTOR_FIELD)
    assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32)), "vector field differs"
    dg, do = g.read_buffer(prg.BUF_DIVERGENCE), o.read_buffer(prg.BUF_DIVERGENCE)
    den = np.linalg.norm(do.astype(np.float64))
    err = np.linalg.norm(dg.astype(np.float64) - do) / max(den, 1e-300)
    print(f"divergence rel L2 = {err:.3e}")
    assert err <= 1e-5


def test_solve_fixed_iterations(pair, prg):
    """Same b, 12 iterations of the build's CG on both sides."""
    name, D, g, o = pair
    b = o.read_buffer(prg.BUF_DIVERGENCE)
    g2 = prg.Reconstructor(max_depth=D, cg_fixed_iters=12)
    (p, n), _ = _workload(name, prg)
    g2.set_points(p, n)
    g2.build_octree()
    g2.write_buffer(prg.BUF_DIVERGENCE, b)
    g2.solve()
    o.solve(mode=0, fixed_iters=12)
    xg, xo = g2.read_buffer(prg.BUF_SOLUTION), o.read_buffer(prg.BUF_SOLUTION)
    g2.close()
    den = np.linalg.norm(xo.astype(np.float64))
    err = np.linalg.norm(xg.astype(np.float64) - xo) / max(den, 1e-300)
    print(f"x rel L2 after 12 iterations = {err:.3e}")
    assert err <= 1e-5


def test_solve_converges_to_fp64(pair, prg):
    """Full solve to 1e-6 relative residual; error against the FP64 oracle solve."""
    name, D, g, o = pair
    b = o.read_buffer(prg.BUF_DIVERGENCE)
    g.write_buffer(prg.BUF_DIVERGENCE, b)
    stats = g.solve()
    x64 = o.solve_fp64(tol=1e-12)
    xg = g.read_buffer(prg.BUF_SOLUTION).astype(np.float64)
    counts, bases = g.level_info()
    for d, s in enumerate(stats):
        sl = slice(bases[d], bases[d] + counts[d])
        bd = b[sl].astype(np.float64)
        if not np.any(bd):
            continue
        assert s["true_rel_residual"] <= 2e-6, (d, s)
        # kappa ~ 25: x error <= kappa * residual
        ex = np.linalg.norm(xg[sl] - x64[sl]) / np.linalg.norm(x64[sl])
        print(f"depth {d}: iters {s['iters']} rel res {s['true_rel_residual']:.2e} FP64 err {ex:.2e}")
        assert ex <= 1e-4


def test_point_values_and_isovalue(pair, prg):
    name, D, g, o = pair
    o.solve(mode=0)
    x = o.read_buffer(prg.BUF_SOLUTION)
    g.write_buffer(prg.BUF_SOLUTION, x)
    iso_g = g.isovalue()
    iso_o = o.isovalue()
    pg, po = g.read_buffer(prg.BUF_POINT_VALUES), o.read_buffer(prg.BUF_POINT_VALUES)
    assert np.array_equal(pg.view(np.uint32), po.view(np.uint32)), "implicit values at samples differ"
    # float(sum in double) / N: the sums agree to ~1e-16 relative
    assert abs(iso_g - iso_o) <= 2 * np.spacing(np.float32(abs(iso_o)))


def test_leaf_mesh_bit_exact(pair, prg):
    name, D, g, o = pair
    x = o.read_buffer(prg.BUF_SOLUTION)
    iso = o.isovalue()
    g.write_buffer(prg.BUF_SOLUTION, x)
    g.set_isovalue(iso)
    o.set_isovalue(iso)
    mg = g.extract_mesh(subdivide=False)
    vo, to = o.extract_leaf_mesh()
    cg, co = g.read_buffer(prg.BUF_LEAF_CORNER_VALUES), o.read_buffer(prg.BUF_LEAF_CORNER_VALUES)
    assert np.array_equal(cg.view(np.uint32), co.view(np.uint32)), "leaf corner values differ"
    assert mg.vertices.shape == vo.shape and mg.triangles.shape == to.shape
    assert np.array_equal(mg.triangles, to), "triangle connectivity differs"
    assert np.array_equal(np.nan_to_num(mg.vertices).view(np.uint32), np.nan_to_num(vo).view(np.uint32))


def test_isovalue_kept_on_device(pair, prg):
    """prg_isovalue with out == NULL leaves the value in device memory for the
    mesh: the mesh equals the one made after reading the value back and
    setting it again from the host."""
    name, D, g, o = pair
    g.write_buffer(prg.BUF_SOLUTION, o.read_buffer(prg.BUF_SOLUTION))
    g._call("prg_isovalue", None)
    m_dev = g.extract_mesh(subdivide=False)
    g.set_isovalue(g.isovalue())
    m_host = g.extract_mesh(subdivide=False)
    assert m_dev.vertices.shape == m_host.vertices.shape
    assert np.array_equal(m_dev.triangles, m_host.triangles)
    assert np.array_equal(np.nan_to_num(m_dev.vertices).view(np.uint32),
                          np.nan_to_num(m_host.vertices).view(np.uint32))
