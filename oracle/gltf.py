"""Test infrastructure (the checker, never the product): a Python restatement
of the glTF 2.0 ingest, which the reference performs in
SceneParser::parse_scene_file (src/Scene/SceneParser.cpp:19-226) through
ASSIMP with aiProcess_PreTransformVertices | aiProcess_Triangulate.

It reads the same files as libmipt's loader and states what the arrays must
hold, from the glTF 2.0 specification and the reference's parser:
- nodes depth first from the scene's roots, world = parent x local, local =
  the column-major "matrix" or T x R x S (section 3.5.3);
- positions as world points, normals by the inverse transpose of the 3x3
  (cofactors over the determinant) renormalised, all in double and rounded
  once to binary32;
- triangles grouped as PreTransformVertices joins meshes: material by
  material (the default glTF material, which the importer appends, last),
  within a material in node order; strips and fans expanded as section
  3.7.2.1 says;
- materials as SceneParser::read_material_properties (:392-436) reads the
  importer's keys, over CPUMaterial's defaults;
- emitters: the triangles of materials with a non-zero emission colour
  (ThreadFunctions.cpp:136-190);
- the first camera at its node's origin, looking down the node's -Z, +Y up.

Arithmetic is done in the same order as the library's so the results agree
bit for bit; what pins the restatement itself is the closed forms of the
tests (known node transforms, the JSON's material values).
"""
import base64
import json
import math
import os
import struct

import numpy as np

CPU_MATERIAL_DEFAULTS = dict(
    emission=(0.0, 0.0, 0.0), emission_strength=1.0, base_color=(1.0, 1.0, 1.0), roughness=0.3, metallic=0.0,
    anisotropy=0.0, specular=1.0, specular_tint=1.0, specular_color=(1.0, 1.0, 1.0), coat=0.0, coat_roughness=0.0,
    sheen=0.0, sheen_roughness=0.5, sheen_color=(1.0, 1.0, 1.0), ior=1.4, specular_transmission=0.0,
    absorption_at_distance=1.0, absorption_color=(1.0, 1.0, 1.0))

COMPONENT = {5120: ("b", 1), 5121: ("B", 1), 5122: ("h", 2), 5123: ("H", 2), 5125: ("I", 4), 5126: ("f", 4)}
WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def _f32(x):
    return float(np.float32(x))


def read_document(path):
    data = open(path, "rb").read()
    bin_chunk = None
    if data[:4] == b"glTF":
        total = min(struct.unpack_from("<I", data, 8)[0], len(data))
        pos, js = 12, None
        while pos + 8 <= total:
            length, kind = struct.unpack_from("<II", data, pos)
            pos += 8
            if kind == 0x4E4F534A and js is None:
                js = data[pos:pos + length]
            elif kind == 0x004E4942 and bin_chunk is None:
                bin_chunk = data[pos:pos + length]
            pos += length
        doc = json.loads(js.decode("utf-8"))
    else:
        doc = json.loads(data.decode("utf-8"))
    folder = os.path.dirname(path)
    buffers = []
    for i, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            buffers.append(bin_chunk)
        elif uri.startswith("data:"):
            buffers.append(base64.b64decode(uri[uri.index(",") + 1:]))
        else:
            from urllib.parse import unquote
            buffers.append(open(os.path.join(folder, unquote(uri)), "rb").read())
    return doc, buffers


def accessor(doc, buffers, index):
    a = doc["accessors"][index]
    n = WIDTH[a["type"]]
    fmt, size = COMPONENT[a["componentType"]]
    count = a["count"]
    out = [[0.0] * n for _ in range(count)]
    if "bufferView" not in a:
        return out
    v = doc["bufferViews"][a["bufferView"]]
    buf = buffers[v["buffer"]]
    stride = v.get("byteStride", size * n)
    base = v.get("byteOffset", 0) + a.get("byteOffset", 0)
    norm = a.get("normalized", False)
    for i in range(count):
        for c in range(n):
            x = float(struct.unpack_from("<" + fmt, buf, base + i * stride + c * size)[0])
            if norm:
                x = {"B": lambda t: t / 255.0, "H": lambda t: t / 65535.0,
                     "b": lambda t: max(t / 127.0, -1.0), "h": lambda t: max(t / 32767.0, -1.0)}[fmt](x)
            out[i][c] = x
    return out


def identity():
    return [[1.0 if i == j else 0.0 for j in range(4)] for i in range(4)]


def mul(a, b):
    r = [[0.0] * 4 for _ in range(4)]
    for i in range(4):
        for j in range(4):
            acc = 0.0
            for k in range(4):
                acc = acc + a[i][k] * b[k][j]
            r[i][j] = acc
    return r


def local_matrix(node):
    if "matrix" in node:
        m = node["matrix"]
        return [[float(m[4 * c + r]) for c in range(4)] for r in range(4)]
    t = node.get("translation", [0.0, 0.0, 0.0])
    q = node.get("rotation", [0.0, 0.0, 0.0, 1.0])
    s = node.get("scale", [1.0, 1.0, 1.0])
    ln = math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
    x, y, z, w = q[0] / ln, q[1] / ln, q[2] / ln, q[3] / ln
    R = [[1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - z * w), 2.0 * (x * z + y * w)],
         [2.0 * (x * y + z * w), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - x * w)],
         [2.0 * (x * z - y * w), 2.0 * (y * z + x * w), 1.0 - 2.0 * (x * x + y * y)]]
    r = identity()
    for i in range(3):
        for j in range(3):
            r[i][j] = R[i][j] * float(s[j])
        r[i][3] = float(t[i])
    return r


def normal_matrix(m):
    cof = [[m[1][1] * m[2][2] - m[1][2] * m[2][1], m[1][2] * m[2][0] - m[1][0] * m[2][2],
            m[1][0] * m[2][1] - m[1][1] * m[2][0]],
           [m[0][2] * m[2][1] - m[0][1] * m[2][2], m[0][0] * m[2][2] - m[0][2] * m[2][0],
            m[0][1] * m[2][0] - m[0][0] * m[2][1]],
           [m[0][1] * m[1][2] - m[0][2] * m[1][1], m[0][2] * m[1][0] - m[0][0] * m[1][2],
            m[0][0] * m[1][1] - m[0][1] * m[1][0]]]
    det = m[0][0] * cof[0][0] + m[0][1] * cof[0][1] + m[0][2] * cof[0][2]
    return [[cof[i][j] / det for j in range(3)] for i in range(3)]


def material(mat):
    """CPUMaterial fields (a dict) of one glTF material."""
    out = dict(CPU_MATERIAL_DEFAULTS)
    pbr = mat.get("pbrMetallicRoughness", {})
    out["base_color"] = tuple(_f32(c) for c in pbr.get("baseColorFactor", [1, 1, 1, 1])[:3])
    out["emission"] = tuple(_f32(c) for c in mat.get("emissiveFactor", [0, 0, 0]))
    out["metallic"] = _f32(pbr.get("metallicFactor", 1.0))
    out["roughness"] = _f32(pbr.get("roughnessFactor", 1.0))
    ext = mat.get("extensions", {})
    if "KHR_materials_emissive_strength" in ext:
        out["emission_strength"] = _f32(ext["KHR_materials_emissive_strength"].get("emissiveStrength", 1.0))
    if "KHR_materials_anisotropy" in ext:
        out["anisotropy"] = _f32(ext["KHR_materials_anisotropy"].get("anisotropyStrength", 0.0))
    if "KHR_materials_sheen" in ext:
        e = ext["KHR_materials_sheen"]
        out["sheen_color"] = tuple(_f32(c) for c in e.get("sheenColorFactor", [0, 0, 0]))
        out["sheen_roughness"] = _f32(e.get("sheenRoughnessFactor", 0.0))
        out["sheen"] = 1.0
    if "KHR_materials_specular" in ext:
        out["specular"] = _f32(ext["KHR_materials_specular"].get("specularFactor", 1.0))
        out["specular_tint"] = 1.0
        out["specular_color"] = (1.0, 1.0, 1.0)
    if "KHR_materials_clearcoat" in ext:
        e = ext["KHR_materials_clearcoat"]
        out["coat"] = _f32(e.get("clearcoatFactor", 0.0))
        out["coat_roughness"] = _f32(e.get("clearcoatRoughnessFactor", 0.0))
    if "KHR_materials_ior" in ext:
        out["ior"] = _f32(ext["KHR_materials_ior"].get("ior", 1.5))
    if "KHR_materials_transmission" in ext:
        out["specular_transmission"] = _f32(ext["KHR_materials_transmission"].get("transmissionFactor", 0.0))
    if "KHR_materials_volume" in ext:
        e = ext["KHR_materials_volume"]
        out["absorption_color"] = tuple(_f32(c) for c in e.get("attenuationColor", [1, 1, 1]))
        if "attenuationDistance" in e:
            out["absorption_at_distance"] = _f32(e["attenuationDistance"])
    return out


def load(path):
    """dict of the scene arrays and camera placement (eye, target, up, yfov)."""
    doc, buffers = read_document(path)
    mats = doc.get("materials", [])
    nmat = len(mats)
    materials = [material(m) for m in mats] + [material({})]
    instances = []
    cam = {"node": None, "world": identity()}

    def visit(i, parent):
        node = doc["nodes"][i]
        world = mul(parent, local_matrix(node))
        if node.get("camera") == 0 and cam["node"] is None:
            cam["node"], cam["world"] = i, world
        if "mesh" in node:
            for prim in doc["meshes"][node["mesh"]]["primitives"]:
                if prim.get("mode", 4) in (4, 5, 6):
                    instances.append((prim.get("material", nmat), world, prim))
        for c in node.get("children", []):
            visit(c, world)

    scenes = doc.get("scenes", [])
    if scenes:
        for r in scenes[doc.get("scene", 0)].get("nodes", []):
            visit(r, identity())
    verts, normals, has_n, idx, mat_idx, emissive = [], [], [], [], [], []
    for m in range(nmat + 1):
        for mi, world, prim in instances:
            if mi != m:
                continue
            attrs = prim["attributes"]
            pos = accessor(doc, buffers, attrs["POSITION"])
            nrm = accessor(doc, buffers, attrs["NORMAL"]) if "NORMAL" in attrs else None
            N = normal_matrix(world) if nrm is not None else None
            base = len(verts)
            for v, (x, y, z) in enumerate(pos):
                verts.append([_f32(world[r][0] * x + world[r][1] * y + world[r][2] * z + world[r][3])
                              for r in range(3)])
                if nrm is None:
                    normals.append([0.0, 0.0, 0.0])
                    has_n.append(0)
                    continue
                a, b, c = nrm[v]
                t = [N[r][0] * a + N[r][1] * b + N[r][2] * c for r in range(3)]
                ln = math.sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2])
                normals.append([_f32(t[r] / ln) if ln > 0 else 0.0 for r in range(3)])
                has_n.append(1)
            ind = [int(i[0]) for i in accessor(doc, buffers, prim["indices"])] if "indices" in prim \
                else list(range(len(pos)))
            mode = prim.get("mode", 4)
            if mode == 4:
                tris = [ind[i:i + 3] for i in range(0, len(ind) - 2, 3)]
            elif mode == 5:
                tris = [[ind[i], ind[i + 1], ind[i + 2]] if i % 2 == 0 else [ind[i + 1], ind[i], ind[i + 2]]
                        for i in range(len(ind) - 2)]
            else:
                tris = [[ind[i], ind[i + 1], ind[0]] for i in range(1, len(ind) - 1)]
            emits = any(c != 0.0 for c in materials[m]["emission"])
            for t in tris:
                idx.append([base + t[0], base + t[1], base + t[2]])
                mat_idx.append(m)
                if emits:
                    emissive.append(len(mat_idx) - 1)
    cams = doc.get("cameras", [])
    camera = None
    if cams and "perspective" in cams[0]:
        W = cam["world"]
        eye = [W[r][3] for r in range(3)]
        fwd = [-W[r][2] for r in range(3)]
        fl = math.sqrt(fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2])
        camera = dict(eye=eye, target=[eye[r] + fwd[r] / fl for r in range(3)], up=[W[r][1] for r in range(3)],
                      yfov=cams[0]["perspective"]["yfov"])
    return dict(vertices=np.array(verts, np.float32).reshape(-1, 3), normals=np.array(normals, np.float32).reshape(-1, 3),
                has_normals=np.array(has_n, np.uint8), indices=np.array(idx, np.int32).reshape(-1, 3),
                material_indices=np.array(mat_idx, np.int32), emissive=np.array(emissive, np.int32),
                materials=materials, camera=camera)
