"""glTF 2.0 ingest (mipt_scene_load_gltf) on the host, against the oracle's
Python restatement (oracle/gltf.py) and against closed forms of the scenes the
reference ships (tests/golden/gltf/, copied from the reference's data/GLTFs/):

- the PBR Cornell box: 8,096 triangles; the box node's quaternion is a 90
  degree turn about +X, so every box vertex is (x, -z, y) of its local
  position; the sphere node scales by 0.2 about its translation; the two
  triangles of the light are the only emitters, emission 1 x strength 100;
  the materials hold the JSON's factors over CPUMaterial's defaults (the
  short box metallic 1 by glTF's default, specular 0 from
  KHR_materials_specular; the sphere transmissive with IOR 1.45); the
  camera at its node's translation looking down -Z with yfov 0.4039;
- the nested-dielectrics scene: its camera node turns -45 degrees about +X;
- every array equal, bit for bit, to the restatement's, for both scenes and
  for variants written here: the same scene as .glb, with a base64 data:
  buffer, with a "matrix" node instead of T R S, under a parent node, with
  a primitive without material (the default glTF material, last), and as
  triangle strips and fans;
- errors: missing file, bad JSON, out-of-range indices, an accessor reading
  past its buffer view, a sparse accessor.
"""
import base64
import copy
import json
import math
import os
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CORNELL = os.path.join(HERE, "golden", "gltf", "cornell_pbr.gltf")
NESTED = os.path.join(HERE, "golden", "gltf", "nested-dielectrics.gltf")
FIELDS = ["emission", "emission_strength", "base_color", "roughness", "metallic", "anisotropy", "specular",
          "specular_tint", "specular_color", "coat", "coat_roughness", "sheen", "sheen_roughness", "sheen_color",
          "ior", "specular_transmission", "absorption_at_distance", "absorption_color"]


@pytest.fixture(scope="module")
def gltf_oracle():
    import gltf

    return gltf


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def check_against_oracle(mipt_mod, gltf_oracle, path, W=320, H=180):
    s = mipt_mod.load_gltf(path, W, H)
    ref = gltf_oracle.load(path)
    assert _same(s.vertices, ref["vertices"])
    assert _same(s.indices, ref["indices"])
    assert _same(s.material_indices, ref["material_indices"])
    assert _same(s.emissive, ref["emissive"])
    assert _same(s.vertex_normals, ref["normals"]) and _same(s.has_vertex_normals, ref["has_normals"])
    assert len(s.materials) == len(ref["materials"])
    for m, r in zip(s.materials, ref["materials"]):
        for f in FIELDS:
            got = getattr(m, f)
            got = tuple(got) if not isinstance(got, float) else got
            want = r[f]
            assert np.array_equal(np.float32(got), np.float32(want)), (f, got, want)
    cam = ref["camera"]
    if cam is not None:
        assert np.array_equal(np.float32(list(s.camera.position)), np.float32(cam["eye"]))
        assert s.camera.vertical_fov == np.float32(cam["yfov"])
        iv = np.array([list(r) for r in s.camera.inverse_view], np.float64)
        fwd = np.array(cam["target"]) - np.array(cam["eye"])
        assert np.allclose(-iv[:3, 2], fwd / np.linalg.norm(fwd), atol=1e-6)
    return s, ref


def test_cornell_closed_forms(mipt_mod, gltf_oracle):
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, CORNELL)
    doc, buffers = gltf_oracle.read_document(CORNELL)
    assert s.triangle_count == 8096 and len(s.materials) == 10
    # the box: a 90-degree turn about +X maps local (x, y, z) to (x, -z, y)
    box_local = np.concatenate([np.array(gltf_oracle.accessor(doc, buffers, p["attributes"]["POSITION"]))
                                for p in doc["meshes"][0]["primitives"]])
    box_world = s.vertices[: box_local.shape[0]]
    mat_of_vertex = np.full(s.vertices.shape[0], -1)
    mat_of_vertex[s.indices.reshape(-1)] = np.repeat(s.material_indices, 3)
    # box vertices come material by material; compare as sets per material
    want = np.stack([box_local[:, 0], -box_local[:, 2], box_local[:, 1]], axis=1)
    assert np.abs(np.sort(box_world, axis=0) - np.sort(want, axis=0)).max() < 2e-6
    # the sphere: 0.2 x its local radius about the node's translation
    t = np.array(doc["nodes"][2]["translation"])
    sph = s.vertices[mat_of_vertex == 8].astype(np.float64)
    sph_local = np.array(gltf_oracle.accessor(doc, buffers, doc["meshes"][1]["primitives"][0]["attributes"]["POSITION"]))
    assert np.allclose(np.sort(np.linalg.norm(sph - t, axis=1)), np.sort(0.2 * np.linalg.norm(sph_local, axis=1)),
                       atol=2e-6)
    # emitters: the light's two triangles
    assert list(s.emissive) == list(np.nonzero(s.material_indices == 2)[0]) and s.emissive.size == 2
    light = s.materials[2]
    assert tuple(light.emission) == (1.0, 1.0, 1.0) and light.emission_strength == 100.0
    short_box, left, glass, default = s.materials[0], s.materials[1], s.materials[8], s.materials[9]
    assert short_box.metallic == 1.0 and short_box.specular == 0.0 and short_box.roughness == 0.0
    assert left.specular == 1.0 and left.metallic == 0.0 and left.roughness == 0.5 and tuple(left.base_color) == (1, 0, 0)
    assert glass.specular_transmission == 1.0 and glass.ior == np.float32(1.4500000476837158)
    assert default.metallic == 1.0 and default.roughness == 1.0 and tuple(default.base_color) == (1, 1, 1)
    assert all(m.coat == 0.0 for m in s.materials)
    # the camera: translation (-0.1, 1, 6.5), looking down -Z, yfov 0.4039 rad
    assert np.allclose(list(s.camera.position), [-0.1, 1.0, 6.5], atol=1e-6)
    iv = np.array([list(r) for r in s.camera.inverse_view])
    assert np.allclose(-iv[:3, 2], [0, 0, -1], atol=1e-6) and np.allclose(iv[:3, 1], [0, 1, 0], atol=1e-6)
    assert s.camera.vertical_fov == np.float32(0.4038851857185364)
    assert (s.has_vertex_normals == 1).all()
    assert np.allclose(np.linalg.norm(s.vertex_normals, axis=1), 1.0, atol=1e-6)


def test_nested_dielectrics(mipt_mod, gltf_oracle):
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, NESTED)
    iv = np.array([list(r) for r in s.camera.inverse_view])
    h = math.sqrt(0.5)
    assert np.allclose(-iv[:3, 2], [0, -h, -h], atol=1e-6)
    assert s.emissive.size > 0 and sum(m.specular_transmission == 1.0 for m in s.materials) == 4


def _write_variant(tmp_path, name, doc, bin_bytes=None):
    path = os.path.join(str(tmp_path), name)
    if bin_bytes is not None and name.endswith(".glb"):
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        bb = bin_bytes + b"\0" * (-len(bin_bytes) % 4)
        total = 12 + 8 + len(js) + 8 + len(bb)
        with open(path, "wb") as f:
            f.write(b"glTF" + struct.pack("<II", 2, total))
            f.write(struct.pack("<II", len(js), 0x4E4F534A) + js)
            f.write(struct.pack("<II", len(bb), 0x004E4942) + bb)
        return path
    with open(path, "w") as f:
        json.dump(doc, f)
    if bin_bytes is not None:
        with open(os.path.join(str(tmp_path), doc["buffers"][0]["uri"]), "wb") as f:
            f.write(bin_bytes)
    return path


def _cornell():
    doc = json.load(open(CORNELL))
    data = open(CORNELL.replace(".gltf", ".bin"), "rb").read()
    return doc, data


def test_container_variants(mipt_mod, gltf_oracle, tmp_path):
    base = mipt_mod.load_gltf(CORNELL, 64, 64)
    doc, data = _cornell()
    glb = copy.deepcopy(doc)
    del glb["buffers"][0]["uri"]
    p = _write_variant(tmp_path, "c.glb", glb, data)
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, p, 64, 64)
    assert _same(s.vertices, base.vertices) and _same(s.indices, base.indices)
    uri = copy.deepcopy(doc)
    uri["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(data).decode()
    p = _write_variant(tmp_path, "d.gltf", uri)
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, p, 64, 64)
    assert _same(s.vertices, base.vertices)


def test_node_variants(mipt_mod, gltf_oracle, tmp_path):
    doc, data = _cornell()
    # the box node as a column-major matrix: the same turn
    m = copy.deepcopy(doc)
    x, y, z, w = m["nodes"][0].pop("rotation")
    R = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    m["nodes"][0]["matrix"] = [R[0][0], R[1][0], R[2][0], 0, R[0][1], R[1][1], R[2][1], 0,
                               R[0][2], R[1][2], R[2][2], 0, 0, 0, 0, 1]
    p = _write_variant(tmp_path, "m.gltf", m, data)
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, p, 64, 64)
    base = mipt_mod.load_gltf(CORNELL, 64, 64)
    assert np.abs(s.vertices - base.vertices).max() < 1e-6
    # everything under a parent that moves by (1, 2, 3): positions move, normals do not
    h = copy.deepcopy(doc)
    h["nodes"].append({"name": "parent", "translation": [1.0, 2.0, 3.0], "children": [0, 1, 2]})
    h["scenes"][0]["nodes"] = [3]
    p = _write_variant(tmp_path, "h.gltf", h, data)
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, p, 64, 64)
    assert np.abs(s.vertices - (base.vertices + np.float32([1, 2, 3]))).max() < 1e-5
    assert np.abs(s.vertex_normals - base.vertex_normals).max() < 1e-6
    assert np.allclose(list(s.camera.position), [0.9, 3.0, 9.5], atol=1e-5)
    # a primitive without material: the default glTF material, last
    d = copy.deepcopy(doc)
    del d["meshes"][1]["primitives"][0]["material"]
    p = _write_variant(tmp_path, "n.gltf", d, data)
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, p, 64, 64)
    assert (s.material_indices[-8064:] == 9).all() and s.materials[9].metallic == 1.0


def test_strips_and_fans(mipt_mod, gltf_oracle, tmp_path):
    # a unit quad as a 4-vertex strip and as a fan, one triangle list each
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    data = pos.tobytes()
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 1]}],
           "nodes": [{"mesh": 0}, {"mesh": 1, "translation": [0, 0, 1]}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "mode": 5}]},
                      {"primitives": [{"attributes": {"POSITION": 0}, "mode": 6}]}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"}],
           "bufferViews": [{"buffer": 0, "byteLength": 48}],
           "buffers": [{"byteLength": 48, "uri": "q.bin"}]}
    p = _write_variant(tmp_path, "q.gltf", doc, data)
    s, _ = check_against_oracle(mipt_mod, gltf_oracle, p, 64, 64)
    assert s.indices.tolist() == [[0, 1, 2], [2, 1, 3], [5, 6, 4], [6, 7, 4]]
    assert (s.has_vertex_normals == 0).all() and s.material_indices.tolist() == [0, 0, 0, 0]
    # no camera: the origin looking down -Z
    assert list(s.camera.position) == [0.0, 0.0, 0.0]


def test_errors(mipt_mod, tmp_path):
    doc, data = _cornell()
    with pytest.raises(mipt_mod.MiptError, match="cannot open"):
        mipt_mod.load_gltf(os.path.join(str(tmp_path), "missing.gltf"), 64, 64)
    bad = os.path.join(str(tmp_path), "bad.gltf")
    open(bad, "w").write('{"asset": {"version": "2.0"}, "nodes": [')
    with pytest.raises(mipt_mod.MiptError, match="JSON"):
        mipt_mod.load_gltf(bad, 64, 64)
    for name, edit, msg in [
        ("i.gltf", lambda d: d["scenes"][0]["nodes"].append(17), "node index"),
        ("a.gltf", lambda d: d["accessors"][0].__setitem__("count", 10 ** 6), "outside its bufferView"),
        ("s.gltf", lambda d: d["accessors"][0].__setitem__("sparse", {"count": 1}), "sparse"),
        ("m.gltf", lambda d: d["meshes"][0]["primitives"][0].__setitem__("material", 99), "material index"),
    ]:
        d = copy.deepcopy(doc)
        edit(d)
        p = _write_variant(tmp_path, name, d, data)
        with pytest.raises(mipt_mod.MiptError, match=msg):
            mipt_mod.load_gltf(p, 64, 64)
