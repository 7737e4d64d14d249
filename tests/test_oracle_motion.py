"""ReSTIR DI temporal reuse under a moving camera (render_settings.accumulate
= 0): the oracle's find_temporal_neighbor_index restatement (UtilsTemporal.h:
37-99) against closed forms on a fronto-parallel wall seen without jitter:

- a static camera finds every pixel itself;
- a camera moved sideways by exactly three pixels' worth finds (x + 3, y),
  and the pixels whose reprojection leaves the image have no neighbour when
  the search is off (max_neighbor_search_count 0);
- with the search on, such pixels may take a pixel of the previous image
  within neighbor_search_radius / 2 of their reprojection, and the
  reprojected position itself is unchanged;
- permutation sampling moves the reprojected pixel to
  ((x + dx) ^ 3) - dx, ((y + dy) ^ 3) - dy with (dx, dy) the low bits;
- a camera that moves forward finds the neighbour at the pixel the wall
  point had before (radial scaling about the image centre);
- accumulating, the neighbour is the pixel itself whatever the camera did.
"""
import math

import numpy as np
import pytest

from conftest import look_at

W, H = 64, 48
VFOV = 60.0
DEPTH = 4.0


def wall_scene(mipt):
    """A 40 x 40 wall at z = -4 facing +z and a small emitter in front of it."""
    verts = np.array([[-20, -20, -DEPTH], [20, -20, -DEPTH], [20, 20, -DEPTH], [-20, 20, -DEPTH],
                      [-0.2, 1.5, -1.0], [0.2, 1.5, -1.0], [0.2, 1.5, -1.4], [-0.2, 1.5, -1.4]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6]], np.int32)
    mats = (mipt.Material * 2)()
    for m in mats:
        mipt.load_library().mipt_material_default(C_ref(m))
    mats[0].base_color[:] = [0.7, 0.7, 0.7]
    mats[0].specular = 0.0
    mats[1].emission[:] = [5.0, 5.0, 5.0]
    cam = look_at(mipt, (0, 0, 0), (0, 0, -1), VFOV, W, H, jitter=False)
    return mipt.Scene(verts, idx, np.array([0, 0, 1, 1], np.int32), mats, np.array([2, 3], np.int32), cam, W, H)


def C_ref(m):
    import ctypes

    return ctypes.byref(m)


def pixels_per_unit():
    """Image pixels per world unit across the wall (horizontal)."""
    tan_h = math.tan(math.radians(VFOV) / 2) * W / H
    return W / (2.0 * DEPTH * tan_h)


def run(mipt, oracle, cams, accumulate=0, **restir):
    scene = wall_scene(mipt)
    s = mipt.default_render_settings(width=W, height=H, accumulate=accumulate,
                                     direct_light_sampling_strategy=mipt.MIPT_LSS_RESTIR_DI)
    rs = mipt.default_restir_di_settings(**restir)
    R = oracle.Renderer(scene, s, rs)
    for f, cam in enumerate(cams):
        R.set_camera(cam)
        R.camera_pass(f, mipt.host_random_number(7 * f))
    return R, R.temporal_neighbors(len(cams) - 1, mipt.host_random_number(1234))


def hits(nb):
    """Pixels whose camera ray hit: the wall's diagonal seam, where the
    reference's (not watertight) triangle test misses both triangles, loses
    a few."""
    hit = nb[:, 0] != -2
    assert hit.mean() > 0.99
    return hit


def test_static_camera(mipt_mod, oracle_mod):
    cam = look_at(mipt_mod, (0, 0, 0), (0, 0, -1), VFOV, W, H, jitter=False)
    R, nb = run(mipt_mod, oracle_mod, [cam, cam], max_neighbor_search_count=0)
    hit = hits(nb)
    assert (nb[hit, 0] == np.arange(W * H)[hit]).all()


def test_sideways_move(mipt_mod, oracle_mod):
    dx = 3.0 / pixels_per_unit()
    c0 = look_at(mipt_mod, (0, 0, 0), (0, 0, -1), VFOV, W, H, jitter=False)
    c1 = look_at(mipt_mod, (dx, 0, 0), (dx, 0, -1), VFOV, W, H, jitter=False)
    _, nb = run(mipt_mod, oracle_mod, [c0, c1], max_neighbor_search_count=0)
    hit = hits(nb)
    x, y = np.meshgrid(np.arange(W), np.arange(H))
    x, y = x.reshape(-1), y.reshape(-1)
    # the previous image's seam pixels fail the similarity test
    assert (nb[hit, 1] == (x + 3)[hit]).all() and (nb[hit, 2] == y[hit]).all()
    inside = (x + 3 < W) & hit
    target = x + 3 + y * W
    seam_before = np.isin(target, np.nonzero(~hits_prev(mipt_mod, oracle_mod, c0))[0])
    assert (nb[inside & ~seam_before, 0] == target[inside & ~seam_before]).all()
    assert (nb[inside & seam_before, 0] == -1).all()
    assert (nb[(x + 3 >= W) & hit, 0] == -1).all()
    # with the search: the reprojection is the same, new neighbours lie in
    # the image within radius / 2 of it
    _, nb2 = run(mipt_mod, oracle_mod, [c0, c1], max_neighbor_search_count=8, neighbor_search_radius=4)
    assert (nb2[:, 1:] == nb[:, 1:]).all()
    keep = inside & ~seam_before
    assert (nb2[keep, 0] == nb[keep, 0]).all()
    out = (x + 3 >= W) & hit
    found = nb2[out, 0]
    fx, fy = found % W, found // W
    ok = found >= 0
    assert ok.sum() > 0
    assert (np.abs(fx[ok] - (x[out] + 3)[ok]) <= 2).all() and (np.abs(fy[ok] - y[out][ok]) <= 2).all()


def hits_prev(mipt, oracle, cam):
    """Hit mask of one frame seen through `cam`."""
    _, nb = run(mipt, oracle, [cam], max_neighbor_search_count=0)
    return nb[:, 0] != -2


def test_permutation_sampling(mipt_mod, oracle_mod):
    cam = look_at(mipt_mod, (0, 0, 0), (0, 0, -1), VFOV, W, H, jitter=False)
    for bits in (0, 5, 11, 15):
        scene = wall_scene(mipt_mod)
        s = mipt_mod.default_render_settings(width=W, height=H, accumulate=0,
                                             direct_light_sampling_strategy=mipt_mod.MIPT_LSS_RESTIR_DI)
        rs = mipt_mod.default_restir_di_settings(max_neighbor_search_count=0, use_permutation_sampling=1)
        R = oracle_mod.Renderer(scene, s, rs)
        R.set_permutation_bits(bits)
        R.camera_pass(0, 1)
        R.camera_pass(0, 2)
        nb = R.temporal_neighbors(0, 3)
        hit = hits(nb)
        x, y = np.meshgrid(np.arange(W), np.arange(H))
        x, y = x.reshape(-1), y.reshape(-1)
        dx, dy = bits & 3, (bits >> 2) & 3
        px, py = ((x + dx) ^ 3) - dx, ((y + dy) ^ 3) - dy
        inside = (px >= 0) & (px < W) & (py >= 0) & (py < H)
        target = np.where(inside, px + py * W, 0)
        ok = inside & hit & hit[target]
        assert (nb[ok, 0] == target[ok]).all(), bits
        assert (nb[~inside & hit, 0] == -1).all()
        assert (nb[hit, 1] == x[hit]).all() and (nb[hit, 2] == y[hit]).all()


def test_forward_move(mipt_mod, oracle_mod):
    c0 = look_at(mipt_mod, (0, 0, 0), (0, 0, -1), VFOV, W, H, jitter=False)
    c1 = look_at(mipt_mod, (0, 0, -1.0), (0, 0, -2.0), VFOV, W, H, jitter=False)
    _, nb = run(mipt_mod, oracle_mod, [c0, c1], max_neighbor_search_count=0)
    # a wall point at pixel offset u from the centre at distance 3 was at
    # u * 3 / 4 at distance 4
    x, y = np.meshgrid(np.arange(W), np.arange(H))
    cx, cy = (W - 1) / 2.0, (H - 1) / 2.0
    ex = (x.reshape(-1) - cx) * 0.75 + cx
    ey = (y.reshape(-1) - cy) * 0.75 + cy
    near_half = (np.abs(ex - np.floor(ex) - 0.5) < 1e-3) | (np.abs(ey - np.floor(ey) - 0.5) < 1e-3)
    hit = hits(nb)
    sel = hit & ~near_half
    assert sel.mean() > 0.5
    assert (nb[sel, 1] == np.floor(ex[sel] + 0.5)).all()
    assert (nb[sel, 2] == np.floor(ey[sel] + 0.5)).all()
    assert (nb[hit, 0] >= 0).mean() > 0.99


def test_accumulating_ignores_motion(mipt_mod, oracle_mod):
    c0 = look_at(mipt_mod, (0, 0, 0), (0, 0, -1), VFOV, W, H, jitter=False)
    c1 = look_at(mipt_mod, (0.7, 0.2, 0), (0.7, 0.2, -1), VFOV, W, H, jitter=False)
    _, nb = run(mipt_mod, oracle_mod, [c0, c1], accumulate=1)
    hit = hits(nb)
    assert (nb[hit, 0] == np.arange(W * H)[hit]).all()
