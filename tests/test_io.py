"""Point readers and the PLY writer (PointStream.inl:29-239, plyfile.cu:2769-2837); host code only."""
import numpy as np
import pytest


def _cloud(n=50, seed=3):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)).astype(np.float32), rng.standard_normal((n, 3)).astype(np.float32)


def test_binary_ply_roundtrip(prg, tmp_path):
    p, n = _cloud()
    path = tmp_path / "pts.ply"
    rec = np.zeros(len(p), dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("q", "<f8"), ("nx", "<f4"),
                                  ("ny", "<f4"), ("nz", "<f4")])
    rec["x"], rec["y"], rec["z"] = p.T
    rec["nx"], rec["ny"], rec["nz"] = n.T
    hdr = ("ply\nformat binary_little_endian 1.0\ncomment test\nelement vertex %d\nproperty float x\nproperty float y\n"
           "property float z\nproperty double q\nproperty float nx\nproperty float ny\nproperty float nz\nend_header\n" % len(p))
    path.write_bytes(hdr.encode() + rec.tobytes())
    a, b = prg.read_points(str(path))
    assert np.array_equal(a, p) and np.array_equal(b, n)


def test_ascii_ply_bnpts_and_text(prg, tmp_path):
    p, n = _cloud(20)
    txt = "\n".join(" ".join(f"{v:.9g}" for v in row) for row in np.hstack([p, n]))
    (tmp_path / "a.ply").write_text("ply\nformat ascii 1.0\nelement vertex 20\nproperty float x\nproperty float y\n"
                                    "property float z\nproperty float nx\nproperty float ny\nproperty float nz\n"
                                    "end_header\n" + txt + "\n")
    (tmp_path / "a.bnpts").write_bytes(np.hstack([p, n]).astype("<f4").tobytes())
    (tmp_path / "a.xyz").write_text(txt + "\n")
    for name in ("a.ply", "a.bnpts", "a.xyz"):
        a, b = prg.read_points(str(tmp_path / name))
        assert np.array_equal(a, p) and np.array_equal(b, n), name


@pytest.mark.parametrize("content", [
    b"not a ply\n",
    b"ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\nend_header\n0 0\n1 1\n",
    b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
    b"property float nx\nproperty float ny\nproperty float nz\nend_header\n0 0 0 1 0 0\n",
])
def test_malformed_ply_is_an_io_error(prg, tmp_path, content):
    path = tmp_path / "bad.ply"
    path.write_bytes(content)
    with pytest.raises(prg.PrgError) as e:
        prg.read_points(str(path))
    assert e.value.status == prg.E_IO


def test_truncated_bnpts_and_missing_file(prg, tmp_path):
    (tmp_path / "t.bnpts").write_bytes(np.zeros(13, "<f4").tobytes())
    for path in (tmp_path / "t.bnpts", tmp_path / "missing.ply"):
        with pytest.raises(prg.PrgError) as e:
            prg.read_points(str(path))
        assert e.value.status == prg.E_IO


def test_ply_writer(prg, tmp_path):
    """Byte-for-byte the layout of the reference's ASCII writer: the header of
    ply_header_complete (plyfile.cu:470-531) for PlyWriteTriangles' two
    elements (:2769-2793), then per element each item as "%g " (float) or
    "%u "/"%d " (uchar count, int index) and a newline (ply_put_element,
    write_ascii_item, :567-626,2102-2147).  Vertices are v * scale + center
    in float (:2800-2806)."""
    v = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [0.7, 0.8, 0.9], [-1.5e-7, 12345.678, 0.0]], np.float32)
    t = np.array([[0, 1, 2], [3, 2, 1]], np.int32)
    center, scale = np.array([1.0, 2.0, -3.0], np.float32), np.float32(2.0)
    mesh = prg.Mesh(v, t, center, float(scale), 4, 2)
    path = tmp_path / "m.ply"
    prg.write_mesh_ply(str(path), mesh)
    w = (v * scale) + center  # float32 arithmetic, as the reference's host loop
    expect = ("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n"
              "element face 2\nproperty list uchar int vertex_indices\nend_header\n")
    for row in w:
        expect += "".join("%g " % float(x) for x in row) + "\n"
    for tri in t:
        expect += "%u " % 3 + "".join("%d " % int(i) for i in tri) + "\n"
    assert path.read_bytes() == expect.encode()
    assert path.read_bytes().splitlines()[-1] == b"3 3 2 1 "
