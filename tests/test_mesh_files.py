"""Mesh files through the C ABI (host only): PLY in ASCII and both binary
byte orders with extra properties and elements, OBJ with normals and
polygons, the reference loaders' error cases, and a file-loaded mesh
building the same octree as the same mesh given in memory."""
import struct

import numpy as np
import pytest

from snogray_amd import native, scenes

V = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0.5)], np.float32)
N = np.array([(0, 0, 1), (0.1, 0, 1), (0, 0.2, 1), (-0.3, 0, 1)], np.float32)
T = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def ply_header(fmt, with_normals=True):
    h = ["ply", f"format {fmt} 1.0", "comment made by hand", "element vertex 4",
         "property float x", "property float y", "property float z"]
    if with_normals:
        h += ["property float nx", "property float ny", "property float nz"]
    h += ["property uchar red", "element face 2", "property list uchar int vertex_indices",
          "property int flags", "element edge 1", "property int vertex1", "property int vertex2",
          "end_header"]
    return ("\n".join(h) + "\n").encode()


def test_ply_ascii(tmp_path):
    rows = [" ".join(map(str, list(V[i]) + list(N[i]) + [7])) for i in range(4)]
    rows += ["3 0 1 2 9", "3 0 2 3 9", "0 1"]
    p = tmp_path / "a.ply"
    p.write_bytes(ply_header("ascii") + ("\n".join(rows) + "\n").encode())
    v, t, n = native.load_mesh(p)
    np.testing.assert_array_equal(v, V)
    np.testing.assert_array_equal(t, T)
    np.testing.assert_array_equal(n, N)


@pytest.mark.parametrize("order,fmt", [("<", "binary_little_endian"), (">", "binary_big_endian")])
def test_ply_binary(tmp_path, order, fmt):
    body = b""
    for i in range(4):
        body += struct.pack(order + "6fB", *V[i], *N[i], 7)
    for tri in T:
        body += struct.pack(order + "B3ii", 3, *tri, 9)
    body += struct.pack(order + "2i", 0, 1)
    p = tmp_path / "b.ply"
    p.write_bytes(ply_header(fmt) + body)
    v, t, n = native.load_mesh(p)
    np.testing.assert_array_equal(v, V)
    np.testing.assert_array_equal(t, T)
    np.testing.assert_array_equal(n, N)


def test_ply_errors(tmp_path):
    p = tmp_path / "q.ply"
    p.write_bytes(ply_header("ascii", False) +
                  b"0 0 0 1\n1 0 0 1\n1 1 0 1\n0 1 0 1\n4 0 1 2 3 0\n3 0 1 2 0\n0 1\n")
    with pytest.raises(native.SgrError, match="Invalid number of indices in face"):
        native.load_mesh(p)
    p.write_bytes(ply_header("ascii", False) + b"0 0 0 1\n1 0 0 1\n")
    with pytest.raises(native.SgrError, match="truncated"):
        native.load_mesh(p)
    p.write_bytes(ply_header("ascii", False) +
                  b"0 0 0 1\n1 0 0 1\n1 1 0 1\n0 1 0 1\n3 0 1 9 0\n3 0 1 2 0\n0 1\n")
    with pytest.raises(native.SgrError, match="out of range"):
        native.load_mesh(p)
    p.write_bytes(b"not a ply\n")
    with pytest.raises(native.SgrError, match="not a PLY file"):
        native.load_mesh(p)


def test_obj_normals_and_polygon_fans(tmp_path):
    text = "\n".join([
        "# a quad and a triangle", "mtllib x.mtl", "o thing", "g part",
        "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0.5", "v 2 2 2",
        "vn 0 0 1", "vn 0.1 0 1", "vn 0 0.2 1", "vn -0.3 0 1", "vn 1 0 0",
        "usemtl red", "f 1//1 2//2 3//3 4//4", "f 5 1 2", ""])
    p = tmp_path / "m.obj"
    p.write_text(text)
    v, t, n = native.load_mesh(p)
    assert len(v) == 5 and n is not None
    # polygon (1 2 3 4) fans into (1 2 3) (1 3 4); indices 1-based -> 0-based
    np.testing.assert_array_equal(t, [[0, 1, 2], [0, 2, 3], [4, 0, 1]])
    np.testing.assert_array_equal(v[:4], V)
    np.testing.assert_array_equal(n[:4], N)


def test_obj_errors(tmp_path):
    p = tmp_path / "e.obj"
    for body, msg in (("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//2 2//2 3//3\n", "identical"),
                      ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1/1 2/2/2 3/3/3\n", "only i and i//n"),
                      ("v 0 0 0\nf 1 2\n", "three vertices"),
                      ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 7\n", "out of range"),
                      ("v 0 0 0\nvn 0 0 1\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "normals for")):
        p.write_text(body)
        with pytest.raises(native.SgrError, match=msg):
            native.load_mesh(p)


def test_loaded_mesh_builds_the_same_octree(tmp_path):
    """An icosphere written as PLY (binary) and as OBJ, loaded back, gives
    the octree of the same mesh given in memory (right-handed, as the
    loaders make it)."""
    v, t = scenes.icosphere(3)
    v = np.asarray(v, np.float32)
    ply = tmp_path / "s.ply"
    hdr = ("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\n"
           "property float y\nproperty float z\nproperty float nx\nproperty float ny\n"
           "property float nz\nelement face %d\nproperty list uchar uint vertex_indices\n"
           "end_header\n" % (len(v), len(t))).encode()
    body = np.concatenate([v, v], 1).astype("<f4").tobytes()
    faces = np.zeros(len(t), dtype=[("n", "u1"), ("i", "<u4", 3)])
    faces["n"], faces["i"] = 3, t
    ply.write_bytes(hdr + body + faces.tobytes())
    obj = tmp_path / "s.obj"
    obj.write_text("".join(f"v {a!r} {b!r} {c!r}\n" for a, b, c in v.tolist()) +
                   "".join(f"f {i + 1} {j + 1} {k + 1}\n" for i, j, k in t.tolist()))
    ref = scenes.SceneDesc()
    ref.mesh(v, t, ref.lambert((0.5, 0.5, 0.5)), left_handed=False, normals=v)
    n0, s0, o0, _ = native.octree_build(ref)
    for path in (ply, obj):
        sd = scenes.SceneDesc()
        sd.mesh_file(path, sd.lambert((0.5, 0.5, 0.5)))
        m = sd.meshes[0]
        np.testing.assert_array_equal(m["v"], v)
        np.testing.assert_array_equal(m["t"], t)
        assert m["left_handed"] is False
        assert (m["vn"] is not None) == (path == ply)
        n1, s1, o1, _ = native.octree_build(sd)
        assert np.array_equal(n0, n1) and np.array_equal(s0, s1) and np.array_equal(o0, o1)
