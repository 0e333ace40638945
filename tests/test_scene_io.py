"""Scene ingestion on the host (libddgi's loaders, no GPU): glTF 2.0 (.gltf with an external or a
base64 buffer, .glb), the .scene container (load and save), Wavefront OBJ and .mat material files,
flattened into the ddgi_set_scene tables.

Pinned against data the loaders do not produce themselves: the reference's own glTF fixture
(data/debug-models/sphere.gltf + sphere.bin, copied to tests/golden/gltf) decoded independently
with numpy from its bufferViews, the fixture's accessor min/max, the unit-sphere geometry, scipy's
quaternion-to-matrix conversion for node rotations, and Python's json for the .scene chunk."""
import base64
import json
import struct

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from vulkanexp_amd import _lib, scenes

GOLD = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "gltf")
SPHERE = f"{GOLD}/sphere.gltf"


def _views(doc, blob, acc):
    a = doc["accessors"][acc]
    bv = doc["bufferViews"][a["bufferView"]]
    off = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
    dt = {5126: np.float32, 5123: np.uint16, 5125: np.uint32}[a["componentType"]]
    n = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}[a["type"]]
    return np.frombuffer(blob, dt, a["count"] * n, off).reshape(a["count"], n).squeeze()


def test_sphere_fixture_tables():
    doc = json.load(open(SPHERE))
    blob = open(f"{GOLD}/sphere.bin", "rb").read()
    s = scenes.load(SPHERE)
    prim = doc["meshes"][0]["primitives"][0]
    pos = _views(doc, blob, prim["attributes"]["POSITION"])
    nrm = _views(doc, blob, prim["attributes"]["NORMAL"])
    uv = _views(doc, blob, prim["attributes"]["TEXCOORD_0"])
    idx = _views(doc, blob, prim["indices"])
    assert len(s.vertices) == len(pos) == 559 and len(s.indices) == len(idx) == 2880
    assert np.array_equal(s.vertices["pos"], pos)
    assert np.array_equal(s.vertices["normal"], nrm)
    assert np.array_equal(s.vertices["tex_coord"], uv)
    assert np.array_equal(s.indices, idx.astype(np.uint32))
    assert np.all(s.vertices["color"] == 1.0) and np.all(s.vertices["tangent"] == 1.0)  # loader defaults
    # geometry: a tessellated unit sphere with outward unit normals
    r = np.linalg.norm(s.vertices["pos"].astype(np.float64), axis=1)
    assert np.all(np.abs(r - 1) < 1e-5)
    n = s.vertices["normal"].astype(np.float64)
    assert np.all(np.abs(np.linalg.norm(n, axis=1) - 1) < 1e-4)
    assert np.all(np.sum(n * s.vertices["pos"], axis=1) / r > 0.99)
    # one mesh, one identity instance of 960 triangles; no material in the file -> the default one
    assert s.offsets.tolist() == [(0, 0, 0)]
    assert len(s.instances) == 1 and s.instances[0]["triangle_count"] == 960
    assert np.array_equal(s.instances[0]["transform"], scenes.IDENTITY_3x4)
    assert s.instances[0]["mask"] == scenes.MASK_STATIC and s.instances[0]["flags"] == 0
    m = s.materials[0]
    assert len(s.materials) == 1 and m["metallic"] == 1 and m["roughness"] == 1 and np.all(m["base_color"] == 1)
    # bounds: the accessor's min/max
    acc = doc["accessors"][prim["attributes"]["POSITION"]]
    assert np.array_equal(s.bounds_min, np.float32(acc["min"])) and np.array_equal(s.bounds_max, np.float32(acc["max"]))


def _same_tables(a, b):
    for f in ("vertices", "indices", "offsets", "instances", "materials", "bounds_min", "bounds_max"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_glb_and_base64_buffers_load_the_same(tmp_path):
    doc = json.load(open(SPHERE))
    blob = open(f"{GOLD}/sphere.bin", "rb").read()
    ref = scenes.load(SPHERE)
    # .gltf with the buffer inlined as a base64 data URI
    d2 = json.loads(json.dumps(doc))
    d2["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
    (tmp_path / "inline.gltf").write_text(json.dumps(d2))
    _same_tables(scenes.load(str(tmp_path / "inline.gltf")), ref)
    # .glb: 12-byte header, JSON chunk (space padded), BIN chunk (zero padded)
    d3 = json.loads(json.dumps(doc))
    del d3["buffers"][0]["uri"]
    js = json.dumps(d3).encode()
    js += b" " * (-len(js) % 4)
    bn = blob + b"\0" * (-len(blob) % 4)
    glb = struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(js) + 8 + len(bn))
    glb += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bn), 0x004E4942) + bn
    (tmp_path / "sphere.glb").write_bytes(glb)
    _same_tables(scenes.load(str(tmp_path / "sphere.glb")), ref)


def _gltf_hierarchy(tmp_path):
    """Two meshes (mesh 0: two primitives with materials 1 and 0; mesh 1: one primitive without a
    material), a TRS parent with a matrix child, a multi-primitive node, two scenes."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    quad = np.array([[0, 0, 0], [2, 0, 0], [2, 0, 2], [0, 0, 2]], np.float32)
    qidx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    blob = tri.tobytes() + quad.tobytes() + qidx.tobytes()
    q = Rotation.from_euler("xyz", [0.3, -0.7, 1.1]).as_quat()  # (x, y, z, w)
    matrix = np.eye(4)
    matrix[:3, :3] = Rotation.from_euler("z", 0.5).as_matrix() * 0.5
    matrix[:3, 3] = [1.0, -2.0, 3.0]
    doc = {
        "asset": {"version": "2.0"},
        "buffers": [{"byteLength": len(blob), "uri": "h.bin"}],
        "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36},
                        {"buffer": 0, "byteOffset": 36, "byteLength": 48},
                        {"buffer": 0, "byteOffset": 84, "byteLength": 24}],
        "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                      {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC3",
                       "min": [0, 0, 0], "max": [2, 0, 2]},
                      {"bufferView": 2, "componentType": 5125, "count": 6, "type": "SCALAR"}],
        "materials": [{"name": "red", "pbrMetallicRoughness": {"baseColorFactor": [1, 0, 0, 0.5],
                                                               "metallicFactor": 0.0, "roughnessFactor": 0.25},
                       "emissiveFactor": [0.1, 0.2, 0.3]},
                      {"name": "metal", "pbrMetallicRoughness": {"metallicFactor": 1.0, "roughnessFactor": 0.5}}],
        "meshes": [{"name": "two", "primitives": [{"attributes": {"POSITION": 0}, "material": 1},
                                                 {"attributes": {"POSITION": 1}, "indices": 2, "material": 0}]},
                   {"name": "one", "primitives": [{"attributes": {"POSITION": 1}, "indices": 2}]}],
        "nodes": [{"name": "parent", "translation": [5, 1, -2], "rotation": list(q), "scale": [2, 3, 4],
                   "children": [1]},
                  {"name": "child", "matrix": list(matrix.T.reshape(-1)), "mesh": 1},
                  {"name": "multi", "mesh": 0, "translation": [0, 10, 0]}],
        "scenes": [{"nodes": [0]}, {"nodes": [2]}],
    }
    (tmp_path / "h.bin").write_bytes(blob)
    (tmp_path / "h.gltf").write_text(json.dumps(doc))
    trs = np.eye(4)
    trs[:3, :3] = Rotation.from_quat(q).as_matrix() @ np.diag([2.0, 3.0, 4.0])
    trs[:3, 3] = [5, 1, -2]
    return str(tmp_path / "h.gltf"), trs, matrix


def test_gltf_hierarchy_transforms_and_order(tmp_path):
    path, trs, matrix = _gltf_hierarchy(tmp_path)
    s = scenes.load(path)
    # meshes: two_Unamed (tri, material 1), two_Unamed (quad, material 0), one_Unamed (quad, default 0)
    assert s.offsets.tolist() == [(1, 0, 0), (0, 3, 3), (0, 7, 9)]
    assert len(s.vertices) == 11 and len(s.indices) == 15
    # instances by (material, mesh): quad of mesh 0 (material 0, mesh 1), quad of mesh 1 (material
    # 0, mesh 2), triangle of mesh 0 (material 1, mesh 0)
    assert s.instances["custom_index"].tolist() == [1, 2, 0]
    assert s.instances["triangle_count"].tolist() == [2, 2, 1]
    multi = np.eye(4)
    multi[:3, 3] = [0, 10, 0]
    want = [multi, trs @ matrix, multi]
    for inst, M in zip(s.instances, want):
        assert np.allclose(inst["transform"].reshape(3, 4), M[:3], atol=2e-6), (inst["transform"], M[:3])
    # materials: baseColorFactor's alpha dropped, factors and emissive read, no textures
    red, metal = s.materials[0], s.materials[1]
    assert red["base_color"].tolist() == [1, 0, 0] and red["metallic"] == 0 and red["roughness"] == 0.25
    assert np.allclose(red["emissive"], [0.1, 0.2, 0.3])
    assert metal["metallic"] == 1 and metal["roughness"] == 0.5 and metal["base_color"].tolist() == [1, 1, 1]
    assert len(s.materials) == 2 and red["albedo_texture"] == 0xFFFFFFFF
    # bounds: the triangle's and the quads' corner boxes under their transforms (two corners each)
    corners = []
    for inst in s.instances:
        oe = s.offsets[inst["custom_index"]]
        lo = s.vertices["pos"][oe["vertex_offset"]:oe["vertex_offset"] + (3 if inst["triangle_count"] == 1 else 4)]
        M = inst["transform"].reshape(3, 4).astype(np.float64)
        for c in (lo.min(0), lo.max(0)):
            corners.append(M[:, :3] @ c + M[:, 3])
    corners = np.array(corners)
    assert np.allclose(s.bounds_min, corners.min(0), atol=1e-5) and np.allclose(s.bounds_max, corners.max(0), atol=1e-5)


def test_scene_container_round_trip(tmp_path):
    sc, _ = scenes.config_small(seed=3)
    path = str(tmp_path / "room.scene")
    scenes.save(sc, path)
    raw = open(path, "rb").read()
    magic, version, length = struct.unpack_from("<III", raw, 0)
    assert (magic, version, length) == (0x4E454353, 0, len(raw))
    jlen, jtype = struct.unpack_from("<II", raw, 12)
    assert jtype == 0x4E4F534A
    doc = json.loads(raw[20:20 + jlen])
    assert set(doc) >= {"materials", "entities", "meshes", "textures"}
    assert len(doc["meshes"]) == len(sc.offsets) and len(doc["materials"]) == len(sc.materials)
    off, chunks = 20 + jlen, 0
    while off < len(raw):
        n, t = struct.unpack_from("<II", raw, off)
        assert t == 0x004E4942
        off += 8 + n
        chunks += 1
    assert off == len(raw) and chunks == 2 * len(doc["meshes"])
    back = scenes.load(path)
    assert np.array_equal(back.vertices, sc.vertices) and np.array_equal(back.indices, sc.indices)
    assert np.array_equal(back.offsets, sc.offsets) and np.array_equal(back.materials, sc.materials)
    # instances come back ordered by (material, mesh); as a set they are the originals
    key = lambda a: sorted((int(i["custom_index"]), tuple(i["transform"].tolist()), int(i["triangle_count"])) for i in a)
    assert key(back.instances) == key(sc.instances)
    assert np.array_equal(back.bounds_min, sc.bounds_min) and np.array_equal(back.bounds_max, sc.bounds_max)
    # saving the loaded scene again gives the same tables
    scenes.save(back, str(tmp_path / "again.scene"))
    again = scenes.load(str(tmp_path / "again.scene"))
    _same_tables(again, back)


def test_material_file_adds_materials_and_texture_records():
    s = scenes.load(SPHERE, materials=[f"{GOLD}/cavern-deposits.mat"])
    # the file's material comes first; the mesh (no material) still refers to index 0
    assert len(s.materials) == 1
    m = s.materials[0]
    assert m["metallic"] == 1 and m["roughness"] == 1 and np.all(m["base_color"] == 1)
    assert (m["albedo_texture"], m["normal_texture"], m["metallic_roughness_texture"], m["emissive_texture"]) == \
        (0, 1, 2, 0xFFFFFFFF)


_CUBE_OBJ = """# unit cube faces and a triangle; a group's vertices follow its `g`, as the reference expects
o cube
g bottom_and_top
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
vt 0.25 0.75
vn 0 0 1
f 1 3 2
f 1 4 3
f 5/1 6 7
f 5 7 8//1
g sides
v 2 0 0
v 3 0 0
v 3 1 0
f 9 10 11
usemtl ignored
"""


def test_obj_groups_indices_and_recomputed_normals(tmp_path):
    """OBJ as the reference loads it: a `g` with vertices already read closes the sub-mesh (its
    normals recomputed as the average of the adjacent triangles' normalize(cross(p0 - p1, p2 - p0)),
    inward for counter-clockwise faces) and opens the next; face indices are absolute in the file;
    `vt`/`vn` indices set that vertex's attributes (the normal is recomputed afterwards anyway)."""
    path = tmp_path / "cube.obj"
    path.write_text(_CUBE_OBJ)
    s = scenes.load(str(path))
    # two meshes: the 8-vertex cube faces (4 triangles) and the 3-vertex triangle
    assert s.offsets.tolist() == [(0, 0, 0), (0, 8, 12)]
    assert len(s.vertices) == 11 and s.indices.tolist() == [0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 2]
    assert len(s.instances) == 2 and s.instances["triangle_count"].tolist() == [4, 1]
    assert np.all(s.instances["transform"] == scenes.IDENTITY_3x4)
    pos = s.vertices["pos"].astype(np.float64)
    # recomputed normals: the bottom face (z = 0, CCW seen from -z... i.e. faces 1 3 2 / 1 4 3) and
    # the top face (z = 1): each triangle's normalize(cross(p0 - p1, p2 - p0))
    want = np.zeros((8, 3))
    cnt = np.zeros(8)
    idx = s.indices[:12].reshape(-1, 3)
    for a, b, c in idx:
        n = np.cross(pos[a] - pos[b], pos[c] - pos[a])
        n /= np.linalg.norm(n)
        for v in (a, b, c):
            want[v] += n
            cnt[v] += 1
    used = cnt > 0
    assert np.allclose(s.vertices["normal"][:8][used], want[used] / cnt[used, None], atol=1e-7)
    # the bottom face's triangles wind 1 3 2 (clockwise seen from +z): cross(p0 - p1, p2 - p0) = +z
    assert np.allclose(s.vertices["normal"][0], [0, 0, 1]) and np.allclose(s.vertices["normal"][4], [0, 0, -1])
    assert np.all(s.vertices["tangent"][:8][used, 3] == 1)
    # vt index 1 on vertex 5 of the file; colours default to white
    assert s.vertices["tex_coord"][4].tolist() == [0.25, 0.75] and np.all(s.vertices["color"] == 1)
    assert len(s.materials) == 1  # the default material: OBJ materials are not read
    assert np.allclose(s.bounds_min, [0, 0, 0]) and np.allclose(s.bounds_max, [3, 1, 1])


def test_loader_errors(tmp_path):
    with pytest.raises(_lib.DdgiError):
        scenes.load(str(tmp_path / "missing.gltf"))
    (tmp_path / "x.ply").write_text("ply\n")
    with pytest.raises(_lib.DdgiError):
        scenes.load(str(tmp_path / "x.ply"))
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")  # vertex 3 does not exist
    with pytest.raises(_lib.DdgiError):
        scenes.load(str(tmp_path / "bad.obj"))
    (tmp_path / "bad.gltf").write_text('{"meshes": [')
    with pytest.raises(_lib.DdgiError):
        scenes.load(str(tmp_path / "bad.gltf"))
    (tmp_path / "bad.scene").write_bytes(struct.pack("<III", 0x12345678, 0, 12))
    with pytest.raises(_lib.DdgiError):
        scenes.load(str(tmp_path / "bad.scene"))
    doc = json.load(open(SPHERE))
    doc["accessors"][3]["count"] = 100000  # runs past the buffer
    doc["buffers"][0]["uri"] = GOLD + "/sphere.bin"
    (tmp_path / "long.gltf").write_text(json.dumps(doc))
    with pytest.raises(_lib.DdgiError):
        scenes.load(str(tmp_path / "long.gltf"))
