"""CPU-side checks of libmipt: the procedural scenes (golden hashes, counts,
invariants) and the C ABI (library loads, exports every symbol declared in
include/mipt.h, struct layouts agree with the header)."""
import ctypes
import json
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mipt.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mipt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(mipt_mod):
    lib = mipt_mod.load_library()
    declared = _declared_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), f"libmipt.so does not export {name}"
    assert sorted(mipt_mod.exported_symbols()) == declared


def test_struct_layouts_match_header(mipt_mod):
    assert ctypes.sizeof(mipt_mod.Camera) == 3 * 64 + 12 + 4 + 12
    text = open(HEADER).read()
    body = re.search(r"typedef struct mipt_material \{(.*?)\} mipt_material;", text, re.S).group(1)
    fields = re.findall(r"(float|int32_t)\s+(\w+)(\[3\])?;", body)
    assert [f[1] for f in fields] == [f[0] for f in mipt_mod.Material._fields_]
    assert ctypes.sizeof(mipt_mod.Material) == sum(12 if f[2] else 4 for f in fields)


def test_material_defaults_follow_cpumaterial(mipt_mod):
    m = mipt_mod.default_material()
    assert m.roughness == np.float32(0.3) and m.specular == 1.0 and m.ior == np.float32(1.4)
    assert list(m.base_color) == [1.0, 1.0, 1.0] and m.alpha_opacity == 1.0
    assert m.coat_ior == 1.5 and m.thin_film_thickness == 500.0


def test_scene_hashes_are_golden(mipt_mod):
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_scene_hashes

    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "scene_hashes.json")))
    assert make_scene_hashes.scene_hash(1, 512, 512) == golden["config1_512x512"]
    assert make_scene_hashes.scene_hash(2, 1920, 1080) == golden["config2_1920x1080"]


def test_scene1_invariants(scene1):
    assert scene1.triangle_count == 24 + 4 + 20 * 5120 == 102428
    assert len(scene1.emissive) == 4
    v = scene1.vertices
    assert v.min() >= -1.0 and v.max() <= 1.0
    # Spheres do not overlap and stay inside the box.
    centres = []
    for s in range(20):
        tri = scene1.indices[28 + s * 5120: 28 + (s + 1) * 5120]
        pts = v[np.unique(tri)]
        c = pts.mean(axis=0)
        r = np.linalg.norm(pts - c, axis=1).mean()
        assert 0.08 - 1e-3 <= r <= 0.2 + 1e-3
        centres.append((c, r))
    for i in range(20):
        for j in range(i):
            assert np.linalg.norm(centres[i][0] - centres[j][0]) > centres[i][1] + centres[j][1]
    mats = scene1.materials
    for s in range(20):
        m = mats[scene1.material_indices[28 + s * 5120]]
        assert (m.metallic == 1.0 and 0.05 <= m.roughness <= 1.0) or (m.metallic == 0.0 and m.specular == 0.0)


def test_scene2_invariants(scene2):
    assert scene2.triangle_count == 524288 + 400 * 1280 + 4096 == 1040384
    assert len(scene2.emissive) == 4096
    mats = scene2.materials
    powers = []
    for e in scene2.emissive:
        tri = scene2.vertices[scene2.indices[e]].astype(np.float64)
        area = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
        em = np.array(mats[scene2.material_indices[e]].emission, np.float64)
        powers.append(em.max() * area)
    lp = np.log10(powers)
    assert lp.min() > -0.5 and lp.max() < 3.01


def test_scene_generation_is_deterministic(mipt_mod):
    a = mipt_mod.generate_scene(1, 64, 64)
    b = mipt_mod.generate_scene(1, 64, 64)
    assert a.vertices.tobytes() == b.vertices.tobytes()
    assert a.indices.tobytes() == b.indices.tobytes()


def test_invalid_scene_config_is_rejected(mipt_mod):
    import pytest

    with pytest.raises(mipt_mod.MiptError):
        mipt_mod.generate_scene(9, 64, 64)
