"""Scene-file ingest on the CPU: the oracle's reader pinned by independently
written files, the reference's own scene files and closed forms; then the
product reader (libluxrays_ingest.so) equal to the oracle, bit for bit."""
import os
import struct

import numpy as np
import pytest

from luxcore_amd import scenefile as P
from oracle import scenefile as O
from tests.scene_fixtures import SCENES, unpack


# ---------------------------------------------------------------------------
# an independent PLY writer
# ---------------------------------------------------------------------------
_CODES = {"char": "b", "uchar": "B", "short": "h", "ushort": "H", "int": "i", "uint": "I",
          "float": "f", "double": "d", "int8": "b", "uint8": "B", "int32": "i", "float32": "f",
          "float64": "d", "uint32": "I", "int16": "h", "uint16": "H"}


def write_ply(path, fmt, elements):
    """elements: list of (name, props, rows); props: (name, type) or
    (name, ('list', ctype, vtype)); rows: tuples of values (lists for list props)."""
    head = ["ply", f"format {fmt} 1.0", "comment written by the ingest tests"]
    for name, props, rows in elements:
        head.append(f"element {name} {len(rows)}")
        for pname, ptype in props:
            if isinstance(ptype, tuple):
                head.append(f"property list {ptype[1]} {ptype[2]} {pname}")
            else:
                head.append(f"property {ptype} {pname}")
    head.append("end_header")
    body = bytearray()
    end = "<" if fmt == "binary_little_endian" else ">"
    for name, props, rows in elements:
        for row in rows:
            words = []
            for (pname, ptype), v in zip(props, row):
                if isinstance(ptype, tuple):
                    if fmt == "ascii":
                        words += [str(len(v))] + [_ascii(x, ptype[2]) for x in v]
                    else:
                        body += struct.pack(end + _CODES[ptype[1]], len(v))
                        for x in v:
                            body += struct.pack(end + _CODES[ptype[2]], x)
                else:
                    if fmt == "ascii":
                        words.append(_ascii(v, ptype))
                    else:
                        body += struct.pack(end + _CODES[ptype], v)
            if fmt == "ascii":
                body += (" ".join(words) + "\n").encode()
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(bytes(body))


def _ascii(v, t):
    return repr(float(v)) if _CODES[t] in "fd" else str(int(v))


def _both(path):
    a = P.load_ply(path)
    b = O.read_ply(path)
    return a, b


# ---------------------------------------------------------------------------
# PLY
# ---------------------------------------------------------------------------
def test_ply_face_rule_known_answer(tmp_path):
    """Triangles as given, quads split (v0 v1 v2), (v0 v2 v3), other lengths
    dropped (exttrianglemeshfile.cpp:207-239)."""
    verts = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (2, 2, 2), (3, 3, 3)]
    faces = [[0, 1, 2], [0, 1, 2, 3], [0, 1, 2, 3, 4], [5, 4, 3], [1, 2], [3, 2, 1, 0]]
    want = np.array([[0, 1, 2], [0, 1, 2], [0, 2, 3], [5, 4, 3], [3, 2, 1], [3, 1, 0]], np.uint32)
    for fmt in ("ascii", "binary_little_endian", "binary_big_endian"):
        p = str(tmp_path / f"f_{fmt}.ply")
        write_ply(p, fmt, [("vertex", [("x", "float"), ("y", "float"), ("z", "float")], verts),
                           ("face", [("vertex_indices", ("list", "uchar", "int"))], [(f,) for f in faces])])
        a, (bv, bt) = _both(p)
        assert np.array_equal(bt, want) and np.array_equal(a.tris, want)
        assert np.array_equal(bv, np.asarray(verts, np.float32)) and np.array_equal(a.verts, bv)


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_round_trip_all_types(tmp_path, fmt):
    """Every scalar type, extra properties before and after x/y/z, an unknown
    element in between, list properties that are not vertex_indices."""
    rng = np.random.default_rng(3)
    nv = 257
    xyz = rng.normal(size=(nv, 3)).astype(np.float32) * np.float32(100)
    vprops = [("nx", "double"), ("x", "float"), ("red", "uchar"), ("y", "float32"), ("s", "short"),
              ("z", "float"), ("u16", "ushort"), ("c", "char"), ("u32", "uint"), ("i32", "int")]
    vrows = [(float(rng.normal()), xyz[i, 0], int(rng.integers(0, 256)), xyz[i, 1], int(rng.integers(-300, 300)),
              xyz[i, 2], int(rng.integers(0, 65536)), int(rng.integers(-128, 128)), int(rng.integers(0, 1 << 31)),
              int(rng.integers(-(1 << 31), 1 << 31))) for i in range(nv)]
    faces, want = [], []
    for _ in range(300):
        k = int(rng.choice([3, 4, 5, 3, 4]))
        f = [int(x) for x in rng.integers(0, nv, size=k)]
        faces.append((int(rng.integers(0, 9)), f, [float(x) for x in rng.random(2)]))
        if k == 3:
            want.append(f)
        elif k == 4:
            want += [[f[0], f[1], f[2]], [f[0], f[2], f[3]]]
    elements = [
        ("vertex", vprops, vrows),
        ("edge", [("vertex1", "int"), ("vertex2", "int")], [(1, 2), (3, 4)]),
        ("face", [("flags", "int"), ("vertex_indices", ("list", "uchar", "uint")),
                  ("texcoord", ("list", "uchar", "float"))], faces),
    ]
    p = str(tmp_path / "all.ply")
    write_ply(p, fmt, elements)
    a, (bv, bt) = _both(p)
    assert np.array_equal(bv, xyz) and np.array_equal(bt, np.asarray(want, np.uint32))
    assert a.verts.tobytes() == bv.tobytes() and a.tris.tobytes() == bt.tobytes()


def test_ply_double_coordinates_round_to_float(tmp_path):
    x = np.array([0.1, 1.0 / 3.0, 1e-30, -7.000000000000001], np.float64)
    for fmt in ("ascii", "binary_big_endian"):
        p = str(tmp_path / f"d_{fmt}.ply")
        write_ply(p, fmt, [("vertex", [("x", "double"), ("y", "double"), ("z", "double")],
                            [(v, -v, 2 * v) for v in x]),
                           ("face", [("vertex_indices", ("list", "uchar", "int"))], [([0, 1, 2],)])])
        a, (bv, _) = _both(p)
        want = np.stack([x, -x, 2 * x], 1).astype(np.float32)
        assert np.array_equal(bv, want) and np.array_equal(a.verts, want)


def test_ply_binary_float_nan_reads_zero(tmp_path):
    """rply.cpp:1411-1419: a binary float32 NaN or Inf is read as 0."""
    p = str(tmp_path / "nan.ply")
    write_ply(p, "binary_little_endian",
              [("vertex", [("x", "float"), ("y", "float"), ("z", "float")],
                [(float("nan"), 1.0, float("inf")), (2.0, float("-inf"), 3.0), (4.0, 5.0, 6.0)]),
               ("face", [("vertex_indices", ("list", "uchar", "int"))], [([0, 1, 2],)])])
    a, (bv, _) = _both(p)
    want = np.array([[0, 1, 0], [2, 0, 3], [4, 5, 6]], np.float32)
    assert np.array_equal(bv, want) and np.array_equal(a.verts, want)


@pytest.mark.parametrize("case", ["missing", "no_faces", "no_vertex_x", "truncated", "bad_header", "bad_index",
                                  "ascii_range", "ascii_token"])
def test_ply_errors(tmp_path, case):
    p = str(tmp_path / f"{case}.ply")
    V = [("x", "float"), ("y", "float"), ("z", "float")]
    F = [("vertex_indices", ("list", "uchar", "int"))]
    tri = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)]
    if case == "no_faces":
        write_ply(p, "ascii", [("vertex", V, tri)])
    elif case == "no_vertex_x":
        write_ply(p, "ascii", [("vertex", [("a", "float"), ("y", "float"), ("z", "float")], tri),
                               ("face", F, [([0, 1, 2],)])])
    elif case == "truncated":
        write_ply(p, "binary_little_endian", [("vertex", V, tri), ("face", F, [([0, 1, 2],)] * 5)])
        with open(p, "r+b") as f:
            f.truncate(os.path.getsize(p) - 7)
    elif case == "bad_header":
        with open(p, "wb") as f:
            f.write(b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nbogus line\nend_header\n")
    elif case == "bad_index":
        write_ply(p, "ascii", [("vertex", V, tri), ("face", F, [([0, 1, 3],)])])
    elif case == "ascii_range":
        write_ply(p, "ascii", [("vertex", V, tri), ("face", [("vertex_indices", ("list", "uchar", "int"))],
                                                           [([0, 1, 2],)])])
        txt = open(p).read().replace("\n3 0 1 2", "\n300 0 1 2")
        open(p, "w").write(txt)
    elif case == "ascii_token":
        write_ply(p, "ascii", [("vertex", V, tri), ("face", F, [([0, 1, 2],)])])
        txt = open(p).read().replace("1.0 0.0 0.0", "1.0 0.0x 0.0")
        open(p, "w").write(txt)
    with pytest.raises(P.IngestError) as e:
        P.load_ply(p)
    assert e.value.code == P.LRH_EIO
    with pytest.raises((O.IngestError, OSError)):
        O.read_ply(p)


def test_reference_ascii_ply_against_text_parse(tmp_path):
    """bigmonkey.ply (ASCII, 9 vertex columns, triangle and quad faces) read
    line by line with numpy, independently of both readers."""
    d = unpack("bigmonkey", tmp_path)
    path = os.path.join(d, "bigmonkey.ply")
    lines = open(path).read().split("\n")
    end = lines.index("end_header")
    nv = int([ln for ln in lines[:end] if ln.startswith("element vertex")][0].split()[2])
    nf = int([ln for ln in lines[:end] if ln.startswith("element face")][0].split()[2])
    body = lines[end + 1:]
    v = np.array([[float(t) for t in ln.split()[:3]] for ln in body[:nv]], np.float64).astype(np.float32)
    tris = []
    for ln in body[nv:nv + nf]:
        f = [int(t) for t in ln.split()]
        assert f[0] == len(f) - 1
        if f[0] == 3:
            tris.append(f[1:])
        elif f[0] == 4:
            tris += [[f[1], f[2], f[3]], [f[1], f[3], f[4]]]
    a, (bv, bt) = _both(path)
    assert bv.shape == (nv, 3) and np.array_equal(bv, v) and np.array_equal(a.verts, v)
    assert np.array_equal(bt, np.asarray(tris, np.uint32)) and np.array_equal(a.tris, bt)


def test_reference_binary_ply_against_structured_read(tmp_path):
    """lightinstances mesh-00001.ply (binary little endian, 8 float vertex
    properties, uchar/uint face lists) through numpy structured dtypes."""
    d = unpack("lightinstances", tmp_path)
    path = os.path.join(d, "mesh-00001.ply")
    raw = open(path, "rb").read()
    hend = raw.index(b"end_header\n") + len(b"end_header\n")
    header = raw[:hend].decode().split("\n")
    nv = int([ln for ln in header if ln.startswith("element vertex")][0].split()[2])
    nf = int([ln for ln in header if ln.startswith("element face")][0].split()[2])
    vprops = [ln.split()[2] for ln in header if ln.startswith("property float")]
    assert "property list uchar uint vertex_indices" in header
    vert = np.frombuffer(raw, dtype=np.dtype([(n, "<f4") for n in vprops]), count=nv, offset=hend)
    faces = np.frombuffer(raw, dtype=np.dtype([("n", "u1"), ("i", "<u4", (3,))]), count=nf,
                          offset=hend + vert.nbytes)
    assert np.all(faces["n"] == 3)
    xyz = np.stack([vert["x"], vert["y"], vert["z"]], 1)
    a, (bv, bt) = _both(path)
    assert np.array_equal(bv, xyz) and np.array_equal(a.verts, xyz)
    assert np.array_equal(bt, faces["i"]) and np.array_equal(a.tris, bt)


# ---------------------------------------------------------------------------
# scene files
# ---------------------------------------------------------------------------
def _same_scene(a, b):
    assert len(a.meshes) == len(b.meshes)
    for x, y in zip(a.meshes, b.meshes):
        assert x.verts.tobytes() == y.verts.tobytes() and x.tris.tobytes() == y.tris.tobytes()
    assert len(a.objects) == len(b.objects)
    for x, y in zip(a.objects, b.objects):
        assert (x["name"], x["mesh"], x["kind"]) == (y["name"], y["mesh"], y["kind"])
        assert np.asarray(x["m"], np.float32).tobytes() == np.asarray(y["m"], np.float32).tobytes()
        assert np.asarray(x["times"], np.float32).tobytes() == np.asarray(y["times"], np.float32).tobytes()
        assert np.asarray(x["keys"], np.float32).tobytes() == np.asarray(y["keys"], np.float32).tobytes()
        assert x["kd"].tobytes() == y["kd"].tobytes() and x["le"].tobytes() == y["le"].tobytes()
    for u, v in zip(a.camera[:3], b.camera[:3]):
        assert np.asarray(u, np.float32).tobytes() == np.asarray(v, np.float32).tobytes()
    assert np.float32(a.camera[3]) == np.float32(b.camera[3])


@pytest.mark.parametrize("scene,name", [(s, n) for s, names in SCENES.items() for n in names])
def test_fixture_scenes_product_equals_oracle(tmp_path, scene, name):
    d = unpack(scene, tmp_path)
    path = os.path.join(d, name)
    _same_scene(P.load_scene(path), O.load_scene(path))


def test_bigmonkey_instances_closed_form(tmp_path):
    d = unpack("bigmonkey", tmp_path)
    sf = P.load_scene(os.path.join(d, "bigmonkey-instances.scn"))
    names = [o["name"] for o in sf.objects]
    assert names == ["back"] + [f"m{i}" for i in range(1, 11)] + ["l1"]
    assert [o["kind"] for o in sf.objects] == [0] + [1] * 11
    assert sf.two_level
    # one shared monkey mesh for the ten instances
    assert len({sf.objects[i]["mesh"] for i in range(1, 11)}) == 1 and len(sf.meshes) == 3
    m1, m2 = sf.objects[1]["m"], sf.objects[2]["m"]
    assert np.array_equal(m1 @ np.array([0, 0, 0, 1], np.float32), np.array([-8, 2, 0, 1], np.float32))
    assert np.array_equal(np.diag(m2), np.array([0.5, 0.5, 0.25, 1], np.float32))
    assert sf.info["nWorldTris"] == sum(sf.meshes[o["mesh"]].ntris for o in sf.objects)
    o, t, u, fov = sf.camera
    assert o == (8.0, -11.0, 5.0) and t == (1.0, 0.0, 1.0) and u == (0.0, 0.0, 1.0) and fov == 45.0


def test_bigmonkey_motion_keys_closed_form(tmp_path):
    """Keys are global->local (parseobjects.cpp:154-159): key k maps the
    object's placed origin back to the origin; motion wins over a static
    transformation given next to it (m7)."""
    d = unpack("bigmonkey", tmp_path)
    sf = P.load_scene(os.path.join(d, "bigmonkey-motion.scn"))
    by = {o["name"]: o for o in sf.objects}
    m1 = by["m1"]
    assert m1["kind"] == 2 and list(m1["times"]) == [0.0, 1.0]
    # Gauss-Jordan with full pivoting (matrix4x4.cpp:118-175) pivots on the
    # translation when it is the largest entry: exact for -8 (a power of two),
    # within an ulp of 14 for -7
    k0, k1 = m1["keys"]
    assert np.array_equal(k0 @ np.array([-8, 2, 0, 1], np.float32), np.array([0, 0, 0, 1], np.float32))
    assert np.allclose(k1 @ np.array([-7, 2, 0, 1], np.float32), [0, 0, 0, 1], rtol=0, atol=4e-6)
    assert abs(float(k1[0, 3]) - 14.0) <= 2 * np.spacing(np.float32(14.0))
    for key in (k0, k1):
        assert np.array_equal(np.diag(key), np.array([2, 2, 2, 1], np.float32))
    assert by["m7"]["kind"] == 2


def test_shapes_objects_and_properties_rules(tmp_path):
    text = "\r\n".join([
        "scene.camera.lookat.orig = 1 2 3",
        "  # a comment after blanks",
        "scene.materials.mat.type = matte",
        "scene.shapes.tri.type = inlinedmesh",
        "scene.shapes.tri.vertices = 0 0 0\t 1 0 0  0 1 0",
        "scene.shapes.tri.faces = 0 1 2",
        "scene.shapes.tri.transformation = 2 0 0 0  0 2 0 0  0 0 2 0  1 2 3 1",
        "scene.shapes.tri2.type = harlequin",
        "scene.shapes.tri2.source = tri",
        "scene.objects.a.material = \"mat\"",
        "scene.objects.a.shape = tri",
        "scene.objects.b.material = 'mat'",
        "scene.objects.b.vertices = 0 0 5 1 0 5 0 1 5",
        "scene.objects.b.faces = 2 1 0",
        "scene.objects.c.material = mat",
        "scene.objects.c.shape = tri2",
        "scene.objects.c.motion.0.time = 0",
        "scene.objects.c.motion.1.time = 0.5",
        "scene.objects.c.motion.1.transformation = 1 0 0 0 0 1 0 0 0 0 1 0 0 0 4 1",
        "scene.objects.a.transformation = 1 0 0 0 0 1 0 0 0 0 1 0 5 0 0 1",
        "scene.objects.a.material = mat",
        "",
    ])
    a = P.parse_scene(text, str(tmp_path))
    assert [o["name"] for o in a.objects] == ["a", "b", "c"]
    assert [o["kind"] for o in a.objects] == [1, 0, 2]
    tri = a.meshes[a.objects[0]["mesh"]]
    assert np.array_equal(tri.verts, np.array([[1, 2, 3], [3, 2, 3], [1, 4, 3]], np.float32))
    assert np.array_equal(a.meshes[a.objects[2]["mesh"]].verts, tri.verts)   # harlequin keeps geometry
    assert np.array_equal(a.meshes[a.objects[1]["mesh"]].tris, np.array([[2, 1, 0]], np.uint32))
    assert a.objects[1]["mesh_name"] == "InlinedMesh-b"
    assert np.array_equal(a.objects[0]["m"][:, 3], np.array([5, 0, 0, 1], np.float32))
    k0, k1 = a.objects[2]["keys"]
    assert np.array_equal(k0, np.eye(4, dtype=np.float32))   # a missing key matrix is the identity
    assert np.array_equal(k1[:, 3], np.array([0, 0, -4, 1], np.float32))
    assert a.camera[0] == (1.0, 2.0, 3.0) and a.camera[2] == (0.0, 0.0, 1.0)
    p = tmp_path / "t.scn"
    p.write_bytes(text.encode())
    _same_scene(P.load_scene(str(p)), O.load_scene(str(p)))


@pytest.mark.parametrize("bad,code", [
    ("scene.objects.a.material = nomat\nscene.objects.a.vertices = 0 0 0 1 0 0 0 1 0\nscene.objects.a.faces = 0 1 2",
     -1),
    ("scene.objects.a.vertices = 0 0 0 1 0 0 0 1 0\nscene.objects.a.faces = 0 1 2", -1),
    ("scene.objects.a.material = mat\nscene.objects.a.shape = nothere", -1),
    ("scene.objects.a.material = mat", -1),
    ("scene.objects.a.material = mat\nscene.objects.a.vertices = 0 0 0 1 0 0 0 1 0\nscene.objects.a.faces = 0 1 3",
     -1),
    ("scene.objects.a.material = mat\nscene.objects.a.vertices = 0 0 0 1 0 0 0 1 0\nscene.objects.a.faces = 0 1 2\n"
     "scene.objects.a.motion.0.time = 1\nscene.objects.a.motion.1.time = 1", -1),
    ("scene.shapes.s.type = subdiv\nscene.shapes.s.source = x", -1),
    ("scene.objects.a.material = mat\nscene.objects.a.ply = nowhere.ply", P.LRH_EIO),
    ("scene.objects.a.material = \"mat", -1),
    ("just text without an equals sign", -1),
    ("scene.objects.a.material = mat\nscene.objects.a.vertices = 0 0 0 1 0 0 0 1 0\nscene.objects.a.faces = 0 1 2\n"
     "scene.objects.a.transformation = 1 0 0", -1),
])
def test_scene_errors(tmp_path, bad, code):
    text = "scene.materials.mat.type = matte\n" + bad + "\n"
    with pytest.raises(P.IngestError) as e:
        P.parse_scene(text, str(tmp_path))
    assert e.value.code == code
    p = tmp_path / "bad.scn"
    p.write_text(text)
    with pytest.raises((O.IngestError, OSError, ValueError, KeyError)):
        O.load_scene(str(p))


def test_cfg_names_scene_file(tmp_path):
    d = unpack("simple", tmp_path)
    cfg = os.path.join(d, "render.cfg")
    with open(cfg, "w") as f:
        f.write("renderengine.type = PATHCPU\nscene.file = scenes/simple/simple.scn\n")
    a = P.load_scene(cfg)
    assert [o["name"] for o in a.objects] == [o["name"] for o in P.load_scene(os.path.join(d, "simple.scn")).objects]
    _same_scene(a, O.load_scene(cfg))


def test_ingest_library_exports_every_declared_function():
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "include", "lrh_ingest.h")) as f:
        names = set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(lrh_[a-z_0-9]+)\(", f.read(), re.M))
    assert len(names) == 11
    L = P.lib()
    for n in names:
        assert hasattr(L, n), n
    assert names == set(P._SIGNATURES), names ^ set(P._SIGNATURES)
    # host code only: the reader does not pull in the HIP runtime
    out = os.popen(f"ldd {P.LIB_PATH}").read()
    assert "amdhip" not in out and "hsa" not in out
    assert L.lrh_scenefile_get_info(None, None) == -1
    assert L.lrh_scenefile_destroy(None) == 0


def test_materials_reduced_for_shading(tmp_path):
    """kd / emission as lrh_shade uses them: numbers, constant textures,
    emission x gain, the 0.75 default for anything else."""
    d = unpack("bigmonkey", tmp_path)
    sf = P.load_scene(os.path.join(d, "bigmonkey-instances.scn"))
    by = {o["name"]: o for o in sf.objects}
    assert np.array_equal(by["l1"]["le"], [8, 8, 8]) and np.array_equal(by["l1"]["kd"], [0, 0, 0])
    assert np.array_equal(by["m1"]["kd"], np.float32([0.75, 0, 0])) and not by["m1"]["le"].any()
    text = "\n".join([
        "scene.textures.t3.type = constfloat3", "scene.textures.t3.value = 0.1 0.2 0.3",
        "scene.textures.t1.type = constfloat1", "scene.textures.t1.value = 0.4",
        "scene.textures.img.type = imagemap",
        "scene.materials.a.type = matte", "scene.materials.a.kd = t3",
        "scene.materials.b.type = glossy2", "scene.materials.b.kd = t1",
        "scene.materials.c.type = matte", "scene.materials.c.kd = img",
        "scene.materials.c.emission = 2 3 4", "scene.materials.c.emission.gain = 0.5",
        "scene.objects.a.material = a", "scene.objects.a.vertices = 0 0 0 1 0 0 0 1 0", "scene.objects.a.faces = 0 1 2",
        "scene.objects.b.material = b", "scene.objects.b.vertices = 0 0 0 1 0 0 0 1 0", "scene.objects.b.faces = 0 1 2",
        "scene.objects.c.material = c", "scene.objects.c.vertices = 0 0 0 1 0 0 0 1 0", "scene.objects.c.faces = 0 1 2",
        ""])
    a = P.parse_scene(text, str(tmp_path))
    kd = [o["kd"].tolist() for o in a.objects]
    assert kd == [np.float32([0.1, 0.2, 0.3]).tolist(), [np.float32(0.4)] * 3, [0.75] * 3]
    assert a.objects[2]["le"].tolist() == [1.0, 1.5, 2.0]
    p = tmp_path / "m.scn"
    p.write_text(text)
    _same_scene(a, O.load_scene(str(p)))
    m = a.materials()
    assert m["kd"].shape == (3, 3) and m["le"][2].tolist() == [1.0, 1.5, 2.0]
    assert len(a.light_triangles()) == 1
