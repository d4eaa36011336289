"""Mesh files through the C ABI (host only): PLY in ASCII and both binary
byte orders with extra properties and elements, OBJ with normals and
polygons, the reference loaders' error cases, and a file-loaded mesh
building the same octree as the same mesh given in memory."""
import struct

import numpy as np
import pytest

from snogray_amd import abi, native, scenes

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


# ---- STL (load/mesh/stl.lua) --------------------------------------------

def _first_use_dedup(pts):
    """Unique rows of PTS in order of first occurrence, and each row's index
    among them (numpy's unique, reordered): what a vertex group gives."""
    u, first, inv = np.unique(pts, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return u[order], rank[inv.ravel()]


def test_stl_merges_vertices_and_fans_facets(tmp_path):
    facets = [[(0, 0, 0), (1, 0, 0), (1, 1, 0)],
              [(0, 0, 0), (1, 1, 0), (0, 1, 0.5), (-1, 0.5, 0.25)],   # 4 vertices: a fan
              [(1, 0, 0), (2, 0, 0), (1, 1, 0)]]
    text = "solid ascii\n"
    for k, f in enumerate(facets):
        text += "  facet normal 0 0 1\n    outer loop\n"
        text += "".join(f"      vertex {x!r} {y!r} {z!r}\n" for x, y, z in f)
        text += "    endloop\n  endfacet\n" if k % 2 else "    end loop\n  end facet\n"
    text += "endsolid\n"
    # a second solid shares the file's vertex group; -0.0 equals 0.0
    text += "solid  ascii\nfacet\nouter loop\nvertex -0.0 0 0\nvertex 2e0 0. +.0\nvertex 1 1 0\n"
    text += "endloop\nendfacet\nend solid\n"
    p = tmp_path / "m.stl"
    p.write_text(text)
    m = native.load_mesh_file(p)
    pts = np.array([v for f in facets for v in f] + [(0, 0, 0), (2, 0, 0), (1, 1, 0)], np.float32)
    uv, idx = _first_use_dedup(pts)
    np.testing.assert_array_equal(m["v"], uv)
    fi = np.cumsum([0] + [len(f) for f in facets] + [3])
    tris = []
    for a, b in zip(fi[:-1], fi[1:]):
        ids = idx[a:b]
        tris += [(ids[0], ids[k - 1], ids[k]) for k in range(2, len(ids))]
    np.testing.assert_array_equal(m["t"], tris)
    assert len(m["v"]) == 6 and m["n"] is None
    assert m["left_handed"] is True      # the loader leaves the Mesh default
    assert len(m["part_colors"]) == 1 and m["part_colors"][0][0] == -1


def test_stl_errors(tmp_path):
    p = tmp_path / "e.stl"
    good_facet = "facet normal 0 0 1\nouter loop\nvertex 0 0 0\nvertex 1 0 0\nvertex 0 1 0\nendloop\nendfacet\n"
    for body, msg in (("solid cube\n" + good_facet + "endsolid\n", "line 1: parse error"),
                      ("solid ascii\n" + good_facet + "endsolid ascii\n", "line 9: parse error"),
                      ("solid ascii\nfacet\nouter loop\nvertex 0 0 0\nvertex 1 0 0\nendloop\n"
                       "endfacet\nendsolid\n", "not enough vertices in facet"),
                      ("solid ascii\nfacet\nouter loop\nvertex 0 0 1E5\n", "line 4: parse error"),
                      ("solid ascii\nfacet\nouter loop\nvertex 0 0\n1 endloop\n", "parse error")):
        p.write_text(body)
        with pytest.raises(native.SgrError, match=msg):
            native.load_mesh(p)
    p.write_text("")
    v, t, n = native.load_mesh(p)
    assert len(v) == 0 and len(t) == 0


# ---- MSH (load/mesh/load-msh.cc) ----------------------------------------

def test_msh_sections_texcoords_and_normals(tmp_path):
    n1 = np.array([(0, 0, 2), (0.3, 0, 1), (0, 0.2, 3)], np.float32)
    n2 = np.array([(1, 1, 1), (0, -4, 0), (0.1, 0.2, 0.3), (3, 4, 0)], np.float32)
    text = ("3 1\nvertices\n0 0 0\n1 0 0\n0 1 0\ntriangles\n0 1 2\n"
            "texcoords\n0 0\n1 0\n0 1\nnormals\n" + "".join(" ".join(map(repr, r.tolist())) + "\n" for r in n1) +
            "4 2 vertices 0 0 1 1 0 1 1 1 1 0 1 1 triangles 0 1 2 0 2 3 normals\n" +
            " ".join(map(repr, n2.ravel().tolist())))        # no final newline
    p = tmp_path / "m.msh"
    p.write_text(text)
    m = native.load_mesh_file(p)
    np.testing.assert_array_equal(m["v"], [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1),
                                           (1, 1, 1), (0, 1, 1)])
    # the second section's indices count from its first vertex
    np.testing.assert_array_equal(m["t"], [(0, 1, 2), (3, 4, 5), (3, 5, 6)])
    nn = np.concatenate([n1, n2])
    length = np.sqrt(nn[:, 0] * nn[:, 0] + nn[:, 1] * nn[:, 1] + nn[:, 2] * nn[:, 2])
    want = nn * (np.float32(1) / length)[:, None]
    np.testing.assert_array_equal(m["n"], want)
    np.testing.assert_allclose(np.linalg.norm(m["n"], axis=1), 1, atol=1e-6)
    assert m["left_handed"] is False


def test_msh_errors(tmp_path):
    p = tmp_path / "e.msh"
    for body, msg in (("3 1 verts 0 0 0 1 0 0 0 1 0 triangles 0 1 2\n", "expected vertices"),
                      ("3 1 vertices 0 0 0 1 0 0 0 1 triangles 0 1 2\n", "bad format"),
                      ("3 1 vertices 0 0 0 1 0 0 0 1 0 triangles 0 1 5\n", "out of range"),
                      ("3 1 vertices 0 0 0 1 0 0 0 1 0 triangles 0 1 2 normals 0 0 1 0 0 1 0 0 1\n"
                       "3 1 vertices 0 0 0 1 0 0 0 1 0 triangles 0 1 2\n", "some sections only"),
                      ("", "empty")):
        p.write_text(body)
        with pytest.raises(native.SgrError, match=msg):
            native.load_mesh(p)


# ---- UG (load/mesh/ug.lua) ----------------------------------------------

UG = """{ a ug file }
c red 1 0 0;
v a 0 0 0;
v b 1 0 0;  { trailing comment }
v c 1 1 0;
v d 0 1 0;
v e 2 1-1;
f (a b c);
f (a c d e) red;
{ a comment
  over two lines }
c{inline}blue 0 0 1;
f ( b e{x}c ) blue;
c red 0.5 0.25 0.125;
f (a b d) red;
f (c d e) blue;

"""


def test_ug_parts_materials_and_fans(tmp_path):
    p = tmp_path / "m.ug"
    p.write_text(UG)
    m = native.load_mesh_file(p)
    np.testing.assert_array_equal(m["v"], [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (2, 1, -1)])
    # (a c d e) red fans into (a c d) (a d e)
    np.testing.assert_array_equal(m["t"], [(0, 1, 2), (0, 2, 3), (0, 3, 4), (1, 4, 2), (0, 1, 3),
                                           (2, 3, 4)])
    # parts in order of first use: default, red, blue, the redefined red
    np.testing.assert_array_equal(m["tri_parts"], [0, 1, 1, 2, 3, 2])
    np.testing.assert_array_equal(m["part_colors"], [(-1, -1, -1), (1, 0, 0), (0, 0, 1),
                                                     (0.5, 0.25, 0.125)])
    assert m["left_handed"] is True
    sd = scenes.SceneDesc()
    grey = sd.lambert((0.5, 0.5, 0.5))
    sd.mesh_file(p, grey)
    tm = sd.meshes[0]["tm"]
    assert tm[0] == grey and len(set(tm.tolist())) == 4
    for k, c in ((1, (1, 0, 0)), (3, (0, 0, 1)), (4, (0.5, 0.25, 0.125))):
        mat = sd.materials[tm[k]]
        assert mat["type"] == abi.SGR_MAT_LAMBERT and tuple(mat["color"]) == c


def test_ug_errors(tmp_path):
    p = tmp_path / "e.ug"
    for body, msg in (("v a 0 0 0;\nv b 1 0 0;\nv c 0 1 0;\nv d 1 1 0;\nf (a b c d);\n",
                       'line 5: Unknown material name "d"'),
                      ("v a 0 0 0;\nf (a b c);\n", 'line 2: Unknown vertex name "b"'),
                      ("v a 0 0 0;", "line 1: parse error"),            # no final newline
                      ("v a 0 0 0; v b 1 0 0;\n", "line 1: parse error"),
                      ("v a 0 0 0;\n  v b 1 0 0;\n", "line 2: parse error"),  # commands start a line
                      ("{c} v a 0 0 0;\n", "line 1: parse error"),
                      ("w 1 2 3;\n", "line 1: parse error"),
                      ("v a 0 0;\n", "line 1: parse error"),
                      ("f (a b);\n", "line 1: parse error")):
        p.write_text(body)
        with pytest.raises(native.SgrError, match=msg):
            native.load_mesh(p)


def test_stl_loaded_mesh_builds_the_same_octree(tmp_path):
    """An icosphere written as ASCII STL: the merged vertices are the
    sphere's own (each appears in a facet before any repeat is read), and
    the octree is the in-memory mesh's, left-handed as the loader leaves
    it."""
    v, t = scenes.icosphere(2)
    v = np.asarray(v, np.float32)
    order = np.unique(t.ravel(), return_index=True)[1]
    first = t.ravel()[np.sort(order)]          # vertices in order of first use
    remap = np.empty(len(v), np.int64)
    remap[first] = np.arange(len(first))
    stl = tmp_path / "s.stl"
    stl.write_text("solid ascii\n" + "".join(
        "facet normal 0 0 0\nouter loop\n" +
        "".join(f"vertex {a!r} {b!r} {c!r}\n" for a, b, c in v[tri].tolist()) +
        "endloop\nendfacet\n" for tri in t) + "endsolid\n")
    ref = scenes.SceneDesc()
    ref.mesh(v[first], remap[t], ref.lambert((0.5, 0.5, 0.5)), left_handed=True)
    sd = scenes.SceneDesc()
    sd.mesh_file(stl, sd.lambert((0.5, 0.5, 0.5)))
    np.testing.assert_array_equal(sd.meshes[0]["v"], v[first])
    np.testing.assert_array_equal(sd.meshes[0]["t"], remap[t])
    assert sd.meshes[0]["left_handed"] is True and sd.meshes[0]["tm"] is None
    n0, s0, o0, _ = native.octree_build(ref)
    n1, s1, o1, _ = native.octree_build(sd)
    assert np.array_equal(n0, n1) and np.array_equal(s0, s1) and np.array_equal(o0, o1)
