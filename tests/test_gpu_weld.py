"""prg_weld_mesh (GPU, through the C ABI) versus the welding oracle
(oracle/weld_ref.py), bit-exact: synthetic meshes with planted duplicates
and a subdivided mesh, whose leaf-level and subdivision batches share
vertices at their seams."""
import numpy as np
import pytest

from oracle import weld_ref

pytestmark = pytest.mark.gpu


def _planted(seed, n, dup, tris):
    rng = np.random.default_rng(seed)
    v = rng.random((n, 3), dtype=np.float32)
    k = int(n * dup)
    v[rng.choice(n, k, replace=False)] = v[rng.integers(0, n, k)]
    if n == 2:
        v[1] = v[0]
    else:
        v[1] = [-0.0, 0.25, 0.75]  # -0 and +0 differ in their bits: not merged
        v[2] = [0.0, 0.25, 0.75]
    return v, rng.integers(0, n, (tris, 3)).astype(np.int32)


def _check(g, prg, v, t, n_leaf):
    print(f"\n[weld: V={len(v)} T={len(t)} n_leaf={n_leaf}]", flush=True)
    mesh = prg.Mesh(v, t, np.zeros(3, np.float32), 1.0, n_leaf, len(t))
    wm, removed = g.weld_mesh(mesh)
    ov, ot, onleaf, oremoved = weld_ref.weld(v, t, n_leaf)
    assert removed == oremoved and wm.n_leaf_vertices == onleaf
    assert np.array_equal(wm.vertices.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(wm.triangles, ot)
    return removed


@pytest.mark.parametrize("seed,n,dup,tris", [(0, 2, 0.5, 1), (1, 1000, 0.3, 2000), (2, 300000, 0.4, 500000),
                                              (3, 70000, 0.0, 1000)])
def test_weld_planted(prg, seed, n, dup, tris):
    v, t = _planted(seed, n, dup, tris)
    with prg.Reconstructor(max_depth=5) as g:
        removed = _check(g, prg, v, t, n // 3)
    assert (removed > 0) == (dup > 0)


def test_weld_subdivided_mesh(prg):
    p, n = prg.synth_sphere(35947, 1)
    with prg.Reconstructor(max_depth=8) as g:
        g.set_points(p, n)
        g.build_octree()
        g.compute_divergence()
        g.solve()
        g.isovalue()
        m = g.extract_mesh(subdivide=True)
        removed = _check(g, prg, m.vertices, m.triangles, m.n_leaf_vertices)
    print(f"seam vertices merged: {removed} of {len(m.vertices)}", flush=True)


def test_weld_rejects_bad_index(prg):
    v = np.zeros((4, 3), np.float32)
    t = np.array([[0, 1, 4]], np.int32)
    with prg.Reconstructor(max_depth=5) as g:
        with pytest.raises(prg.PrgError):
            g.weld_mesh(prg.Mesh(v, t, np.zeros(3, np.float32), 1.0, 4, 1))
        wm, removed = g.weld_mesh(prg.Mesh(v[:0], t[:0], np.zeros(3, np.float32), 1.0, 0, 0))
        assert removed == 0 and len(wm.vertices) == 0


def test_cli_default_subdivides_and_welds(prg, tmp_path):
    """The CLI writes the reference main()'s mesh by default, leaf-level batch
    plus subdivision batches (main.cu:3799-4561); --weld writes the welded mesh
    of the API path, --no-subdivide the leaf-level batch alone."""
    import json
    import subprocess
    from pathlib import Path

    p, n = prg.synth_sphere(35947, 1)
    src = tmp_path / "sphere.bnpts"
    np.hstack([p, n]).astype(np.float32).tofile(src)
    exe = Path(prg.__file__).resolve().parent / "poisson_recon"
    with prg.Reconstructor(max_depth=7) as g:
        g.set_points(p, n)
        g.build_octree()
        g.compute_divergence()
        g.solve()
        g.isovalue()
        full = g.extract_mesh(subdivide=True)
        leaf = g.extract_mesh(subdivide=False)
        wm, removed = g.weld_mesh(full)
    assert removed > 0 and len(full.triangles) > len(leaf.triangles)
    for flags, want in (([], full), (["--weld"], wm), (["--no-subdivide"], leaf)):
        out = tmp_path / "mesh.ply"
        print(f"\n[cli: {exe.name} --depth 7 {' '.join(flags)}]", flush=True)
        r = subprocess.run([str(exe), "--in", str(src), "--out", str(out), "--depth", "7", *flags, "--json"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        info = json.loads(r.stdout.strip().splitlines()[-1])
        assert info["vertices"] == len(want.vertices) and info["triangles"] == len(want.triangles), flags
        body = out.read_bytes()
        head = body[:400].decode("ascii", "replace")
        assert f"element vertex {len(want.vertices)}" in head and f"element face {len(want.triangles)}" in head
        faces = body.split(b"end_header\n", 1)[1].splitlines()[len(want.vertices):]
        expect = [("3 " + "".join("%d " % i for i in tri)).encode() for tri in want.triangles]
        assert faces == expect, flags
