"""Pins the oracle's emissive power sampling, alias table and ReSTIR DI
passes with properties the mathematics fixes:

- the alias table reproduces the pmf exactly (enumeration of every
  (bucket, accept/alias) outcome), and the library's host build is the
  same table bit for bit;
- light sampling by power and uniformly are both unbiased estimators of the
  same direct lighting, so their means agree;
- ReSTIR DI with the initial candidates only, and with spatial reuse in a
  scene without occluders, converges to the plain NEE + MIS mean: a wrong
  candidate weight, MIS term or Jacobian shows up as a bias of several %;
- reservoir invariants: M capped, UCW >= 0 or -1 (killed by visibility
  reuse), target >= 0, flags within the ReSTIR DI flag set.

The remaining bias of temporal + spatial reuse with occluders (about 0.6 %
darker for one spatial pass at 1024 frames) is the algorithm's visibility
reuse, measured separately; it is not asserted here."""
import numpy as np
import pytest

W = H = 48


def _pmf_of(probas, alias):
    n = probas.size
    pmf = probas.astype(np.float64).copy()
    np.add.at(pmf, alias, 1.0 - probas.astype(np.float64))
    return pmf / n


def _float_sum(w):
    s = np.float32(0.0)
    for x in np.asarray(w, np.float32):
        s = np.float32(s + x)
    return s


@pytest.mark.parametrize("case", ["random", "lognormal", "uniform", "one_hot", "two"])
def test_alias_table_reproduces_pmf(mipt_mod, oracle_mod, case):
    rng = np.random.default_rng(3)
    w = {"random": rng.random(1000), "lognormal": np.exp(rng.normal(0, 3, 4096)), "uniform": np.ones(64),
         "one_hot": np.eye(1, 17, 5)[0], "two": np.array([1.0, 3.0])}[case].astype(np.float32)
    total = _float_sum(w)
    p, a = oracle_mod.alias_table(w, total)
    pm, am = mipt_mod.compute_alias_table(w, total)
    assert np.array_equal(p.view(np.uint32), pm.view(np.uint32)) and np.array_equal(a, am)
    assert ((p >= 0) & (p <= 1)).all() and ((a >= 0) & (a < w.size)).all()
    pmf = _pmf_of(p, a)
    want = w.astype(np.float64) / w.astype(np.float64).sum()
    assert np.abs(pmf - want).max() < 2e-6 * max(1.0, want.max() * w.size)
    if case == "uniform":
        assert (p == 1.0).all()


def test_alias_table_empty(mipt_mod, oracle_mod):
    p, a = mipt_mod.compute_alias_table(np.zeros(0, np.float32), 0.0)
    assert p.size == 0 and a.size == 0


def _draws(s, rs):
    """Host random numbers per sample, as mipt_draws_per_sample counts them."""
    if s.direct_light_sampling_strategy != 5:
        return 2
    if rs.do_fused_spatiotemporal:
        return 3 + 3 + rs.spatial_number_of_passes - 1
    return 3 + 2 * rs.do_temporal_reuse_pass + rs.spatial_number_of_passes * rs.do_spatial_reuse_pass


def _render(mipt_mod, oracle_mod, scene, frames, strategy=3, base=0, **restir):
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=0, direct_light_sampling_strategy=strategy,
                                         direct_light_sampling_base_strategy=base)
    rs = mipt_mod.default_restir_di_settings(**restir)
    R = oracle_mod.Renderer(scene, s, rs)
    draws = _draws(s, rs)
    for f in range(frames):
        R.render_frame(f, [mipt_mod.host_random_number(f * draws + k) for k in range(draws)])
    return R.fb64.reshape(-1, 3) / frames, R


@pytest.fixture(scope="module")
def scene48(mipt_mod):
    return mipt_mod.generate_scene(1, W, H)


@pytest.fixture(scope="module")
def open_scene48(mipt_mod):
    """Config 1 without the spheres: nothing occludes the lights."""
    full = mipt_mod.generate_scene(1, W, H)
    return mipt_mod.Scene(full.vertices, full.indices[:28], full.material_indices[:28], full.materials,
                          full.emissive, full.camera, W, H)


@pytest.fixture(scope="module")
def mis_reference(mipt_mod, oracle_mod, scene48):
    img, _ = _render(mipt_mod, oracle_mod, scene48, 512)
    return img.mean(axis=0)


def test_power_sampling_is_unbiased(mipt_mod, oracle_mod, scene48, mis_reference):
    img, R = _render(mipt_mod, oracle_mod, scene48, 512, base=1)
    assert np.allclose(img.mean(axis=0), mis_reference, rtol=0.01)
    assert R.rejected == 0


def test_restir_initial_candidates_unbiased(mipt_mod, oracle_mod, scene48, mis_reference):
    img, R = _render(mipt_mod, oracle_mod, scene48, 256, strategy=5, do_temporal_reuse_pass=0,
                     do_spatial_reuse_pass=0)
    assert np.allclose(img.mean(axis=0), mis_reference, rtol=0.01)
    res = R.read_reservoirs(3)
    hit = R.gb_prims >= 0
    assert (res["M"][hit] == 1).all()


def test_restir_spatial_unbiased_without_occluders(mipt_mod, oracle_mod, open_scene48):
    ref, _ = _render(mipt_mod, oracle_mod, open_scene48, 512)
    img, R = _render(mipt_mod, oracle_mod, open_scene48, 256, strategy=5, do_temporal_reuse_pass=0)
    assert np.allclose(img.mean(axis=0), ref.mean(axis=0), rtol=0.01)


def test_restir_reservoir_invariants(mipt_mod, oracle_mod, scene48):
    _, R = _render(mipt_mod, oracle_mod, scene48, 6, strategy=5, spatial_number_of_passes=2)
    hit = R.gb_prims >= 0
    for which in (0, 1, 2, 3):
        res = R.read_reservoirs(which)
        assert (res["M"] >= 0).all() and (res["M"] <= 3).all()
        assert ((res["ucw"] >= 0) | (res["ucw"] == -1)).all()
        assert (res["target"] >= 0).all()
        assert (res["flags"] & ~0xff == 0).all()
    out = R.read_reservoirs(3)
    # shaded reservoirs of hit pixels point at emissive triangles
    sel = hit & (out["ucw"] > 0)
    assert sel.any()
    assert np.isin(out["tri"][sel], R.emissive).all()


def test_backfacing_bsdf_candidate_weight(mipt_mod, oracle_mod):
    """Pixels that see an emissive triangle can send a BSDF candidate into the
    back of another emitter: the reference's weight is 0/0 = NaN
    (InitialCandidates.h:402-410) and the reservoir's UCW becomes NaN. With
    guard_backfacing_bsdf_candidates the candidate weighs 0 and every
    reservoir stays finite."""
    W2, H2 = 240, 135
    scene = mipt_mod.generate_scene(2, W2, H2)
    for m in scene.materials:
        m.coat = m.sheen = m.specular_transmission = m.diffuse_transmission = m.thin_film = m.specular = 0.0
    s = mipt_mod.default_render_settings(width=W2, height=H2, nb_bounces=1, direct_light_sampling_strategy=5,
                                         direct_light_sampling_base_strategy=1)
    counts = {}
    for guard in (0, 1):
        rs = mipt_mod.default_restir_di_settings(number_of_initial_light_candidates=32, m_cap=20,
                                                 spatial_number_of_passes=2, guard_backfacing_bsdf_candidates=guard)
        R = oracle_mod.Renderer(scene, s, rs)
        bad = 0
        for f in range(2):
            R.render_frame(f, [mipt_mod.host_random_number(f * 7 + k) for k in range(7)])
            for which in (0, 1, 2):
                bad += int((~np.isfinite(R.read_reservoirs(which)["ucw"])).sum())
        counts[guard] = (bad, R.rejected)
    assert counts[0][0] > 0
    assert counts[1] == (0, 0)


def test_temporal_reuse_of_a_surface_that_missed(mipt_mod, oracle_mod, open_back_box_scene):
    """A pixel whose camera ray hits in frame f-2, misses in f-1 and hits in
    f: the temporal pass of frame f evaluates the initial candidate at the
    previous-frame surface, which has no primary hit (prim -1) but keeps the
    material of the older hit. Its volume state is the empty stack
    (RayVolumeState.h:37-39); reading the material index of primitive -1
    was an out-of-bounds access. This pins that the open-back box reaches
    the case within 4 frames: such pixels exist with a temporal reservoir
    (M > 0) and an initial candidate on a light, and the oracle's temporal
    output for them is a finite reservoir."""
    Wt = Ht = 24
    scene = open_back_box_scene(mipt_mod, Wt, Ht)
    s = mipt_mod.default_render_settings(width=Wt, height=Ht, nb_bounces=1, direct_light_sampling_strategy=5)
    rs = mipt_mod.default_restir_di_settings()
    R = oracle_mod.Renderer(scene, s, rs)
    hist = []
    reached = 0
    for f in range(4):
        rn = [mipt_mod.host_random_number(f * 6 + k) for k in range(6)]
        R.camera_pass(f, rn[0])
        hist.append(R.gb_prims.copy())
        tn = R.read_reservoirs(1)  # the last frame's output: this frame's temporal input
        R.restir_initial(f, rn[1])
        if f >= 2:
            sel = (hist[-3] >= 0) & (hist[-2] < 0) & (hist[-1] >= 0)
            sel &= (tn["M"] > 0) & (R.read_reservoirs(0)["tri"] >= 0)
            reached += int(sel.sum())
        R.restir_temporal(f, rn[2])
        if f >= 2:
            out = R.read_reservoirs(0)
            assert np.isfinite(out["ucw"][sel]).all() and (out["M"][sel] >= 1).all()
        R.restir_spatial(0, f, rn[4])
        R.path_trace(f, rn[5])
    assert reached >= 4


def _frame_numbers(mipt_mod, draws):
    return lambda f: [mipt_mod.host_random_number(f * draws + k) for k in range(draws)]




def test_windowed_frames_equal_full_frames_inside_the_tile(mipt_mod, oracle_mod):
    """The windowed frame renderer (each pass over the pixels the tile's
    final values depend on) gives, inside the tile, the full-frame
    renderer's reservoirs and framebuffer bit for bit, over 3 frames of
    temporal + 2 spatial passes. A halo one pixel narrower than the reuse
    radius changes some of them: the window does leave pixels out, and the
    margin is the radius."""
    Wt = Ht = 128
    scene = mipt_mod.generate_scene(1, Wt, Ht)
    s = mipt_mod.default_render_settings(width=Wt, height=Ht, nb_bounces=2, direct_light_sampling_strategy=5)
    rs = mipt_mod.default_restir_di_settings(spatial_reuse_radius=6, spatial_number_of_passes=2, m_cap=20)
    draws = _draws(s, rs)
    frames = 3
    tile = (52, 76, 40, 64)
    full = oracle_mod.Renderer(scene, s, rs)
    for f in range(frames):
        full.render_frame(f, _frame_numbers(mipt_mod, draws)(f))
    sel = np.zeros((Ht, Wt), bool)
    sel[tile[2]:tile[3], tile[0]:tile[1]] = True
    sel = sel.reshape(-1)
    assert (full.gb_prims[sel] >= 0).all()
    diffs = {}
    for halo in (None, rs.spatial_reuse_radius - 1):
        win = oracle_mod.Renderer(scene, s, rs)
        win.render_frames_in_window(frames, tile, _frame_numbers(mipt_mod, draws), halo=halo)
        # the first window is the tile grown by frames x passes halos and
        # stays inside the image: the pixels outside it were never computed
        assert (win.read_reservoirs(0)["M"][~sel] == 0).any()
        n = 0
        for which in (0, 1, 2, 3):
            a, b = win.read_reservoirs(which)[sel], full.read_reservoirs(which)[sel]
            n += int((a.view(np.uint32).reshape(len(a), -1) != b.view(np.uint32).reshape(len(b), -1)).any(axis=1).sum())
        n += int((win.fb.reshape(-1, 3)[sel] != full.fb.reshape(-1, 3)[sel]).any(axis=1).sum())
        diffs[halo] = n
    assert diffs[None] == 0
    assert diffs[rs.spatial_reuse_radius - 1] > 0


PRESAMPLED = dict(do_light_presampling=1, light_presampling_number_of_subsets=16, light_presampling_subset_size=64)


@pytest.mark.parametrize("scene_name", ["open", "full"])
def test_restir_fused_spatiotemporal_unbiased(mipt_mod, oracle_mod, scene48, open_scene48, scene_name):
    """The fused spatiotemporal pass (temporal neighbour, spatial neighbours
    of last frame's reservoirs, initial candidates, pairwise MIS with
    confidence weights) over 256 frames averages to the MIS light + BSDF
    image within 1.5 % per channel, with and without occluders. Dropping the
    temporal neighbour's canonical-weight term moves the mean by 5-30 %;
    leaving its M out of the confidence sum makes it diverge."""
    scene = open_scene48 if scene_name == "open" else scene48
    ref, _ = _render(mipt_mod, oracle_mod, scene, 512)
    img, R = _render(mipt_mod, oracle_mod, scene, 256, strategy=5, do_fused_spatiotemporal=1)
    assert np.allclose(img.mean(axis=0), ref.mean(axis=0), rtol=0.015)
    out = R.read_reservoirs(3)
    hit = R.gb_prims >= 0
    assert (out["M"][hit] >= 1).all() and (out["M"] <= 20).all()
    sel = hit & (out["ucw"] > 0)
    assert np.isin(out["tri"][sel], R.emissive).all()


def test_light_presampling_unbiased(mipt_mod, oracle_mod, scene48, mis_reference):
    """Initial candidates drawn from 16 presampled subsets of 64 lights
    (tile 8x8 per subset) average to the MIS light + BSDF image within 1 %."""
    img, R = _render(mipt_mod, oracle_mod, scene48, 256, strategy=5, do_temporal_reuse_pass=0,
                     do_spatial_reuse_pass=0, **PRESAMPLED)
    assert np.allclose(img.mean(axis=0), mis_reference, rtol=0.01)
    assert (R.read_reservoirs(3)["M"][R.gb_prims >= 0] == 1).all()


def test_fused_first_frame_keeps_the_initial_candidates(mipt_mod, oracle_mod, scene48):
    """On the first frame last frame's reservoirs are empty (M = 0), so the
    fused pass resamples only the initial candidates: the output reservoir
    holds the initial one's sample, M, target and UCW, except that an initial
    reservoir that visibility reuse emptied (UCW -1) is not selected (sample
    cleared, M kept) and the final visibility reuse empties the occluded
    ones (none here: the initial pass's own visibility reuse already did)."""
    s = mipt_mod.default_render_settings(width=W, height=H, nb_bounces=0, direct_light_sampling_strategy=5)
    rs = mipt_mod.default_restir_di_settings(do_fused_spatiotemporal=1)
    R = oracle_mod.Renderer(scene48, s, rs)
    R.render_frame(0, _frame_numbers(mipt_mod, _draws(s, rs))(0))
    init, out = R.read_reservoirs(0), R.read_reservoirs(3)
    hit = R.gb_prims >= 0
    live = hit & (init["ucw"] > 0)
    dead = hit & ~live
    assert dead.any() and live.sum() > 0.5 * hit.sum()
    assert np.array_equal(init["M"][hit], out["M"][hit])
    for k in ("tri", "target", "point"):
        assert np.array_equal(init[k][live], out[k][live]), k
    assert (out["tri"][dead] == -1).all() and (out["target"][dead] == 0).all()
    # the initial pass's visibility reuse left only unoccluded samples, so
    # the final one keeps them all; UCW = (1 * target * UCW) / target
    assert np.allclose(init["ucw"][live], out["ucw"][live], rtol=2e-7, atol=0)


@pytest.mark.parametrize("passes", [1, 2])
def test_windowed_fused_frames_equal_full_frames(mipt_mod, oracle_mod, passes):
    """The windowed renderer routes the presampling and fused passes as the
    full-frame one does: inside the tile both agree bit for bit over 3
    frames."""
    Wt = Ht = 96
    scene = mipt_mod.generate_scene(1, Wt, Ht)
    s = mipt_mod.default_render_settings(width=Wt, height=Ht, nb_bounces=1, direct_light_sampling_strategy=5)
    rs = mipt_mod.default_restir_di_settings(spatial_reuse_radius=5, spatial_number_of_passes=passes, m_cap=20,
                                             do_fused_spatiotemporal=1, **PRESAMPLED)
    draws = _draws(s, rs)
    tile = (40, 60, 36, 56)
    full = oracle_mod.Renderer(scene, s, rs)
    for f in range(3):
        full.render_frame(f, _frame_numbers(mipt_mod, draws)(f))
    win = oracle_mod.Renderer(scene, s, rs)
    win.render_frames_in_window(3, tile, _frame_numbers(mipt_mod, draws))
    sel = np.zeros((Ht, Wt), bool)
    sel[tile[2]:tile[3], tile[0]:tile[1]] = True
    sel = sel.reshape(-1)
    assert (win.read_reservoirs(0)["M"][~sel] == 0).any()
    for which in (0, 1, 2, 3):
        a, b = win.read_reservoirs(which)[sel], full.read_reservoirs(which)[sel]
        assert a.tobytes() == b.tobytes(), which
    assert win.fb.reshape(-1, 3)[sel].tobytes() == full.fb.reshape(-1, 3)[sel].tobytes()


def test_zero_xorshift_state_pixel_under_freeze(mipt_mod, oracle_mod):
    """Regression: with freeze_random the seed of pixel 60 is wang_hash(61)
    = 0, a xorshift state that stays 0, so each of its draws is 0.0. The light
    candidates' `draw > envmap probability` test then fails with probability
    0 too; without an envmap that draw must leave no candidate rather than
    sample an envmap that does not exist (the read of an empty alias table
    that faulted). Two frames of ReSTIR DI render, and that pixel's initial
    reservoir holds no envmap sample."""
    assert oracle_mod.lib().orc_wang_hash(61) == 0
    scene = mipt_mod.generate_scene(1, 64, 64)
    s = mipt_mod.default_render_settings(width=64, height=64, direct_light_sampling_strategy=5, freeze_random=1)
    rs = mipt_mod.default_restir_di_settings()
    R = oracle_mod.Renderer(scene, s, rs)
    draws = _draws(s, rs)
    for f in range(2):
        R.render_frame(f, _frame_numbers(mipt_mod, draws)(f))
    init = R.read_reservoirs(0)
    assert R.gb_prims[60] >= 0 and (init["flags"][60] & 1) == 0
    assert np.isfinite(R.fb).all()
