"""GPU parity of the learned radiance cache through the C ABI, against the
CPU restatement in oracle/nrc.py.

Bars:
  - encoding (hash grid + one-blob, bfloat16 rows): bit-exact;
  - inference: with s = |oracle| + mean |oracle|, |gpu - oracle| <= 1e-4 s on
    >= 99 % of rows and <= 2e-2 s on every row. The oracle sums in float64 and
    the MFMA units in float32; the hidden activations are rounded to bfloat16
    on both sides, so now and then one lands on the other side of a rounding
    boundary (a 2^-8 relative step that the following layers carry);
  - gradients: relative Frobenius error <= 5e-3 per weight matrix and for the
    grid (same reason, plus float32 atomics in an unspecified order);
  - Adam: masters within 1e-3 * lr of the oracle's step on the GPU's own
    gradient; the bfloat16 / FP16 copies equal the rounded masters bit for bit;
  - training reduces the loss of a smooth synthetic radiance field by >= 10x.
"""
import numpy as np
import pytest

import nrc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _bounds(scene):
    return nrc.bounds_transform(scene.vertices.min(axis=0), scene.vertices.max(axis=0))


def _queries(rng, scene, n, edge=True):
    lo, hi = scene.vertices.min(axis=0), scene.vertices.max(axis=0)
    ext = hi - lo
    q = np.zeros((n, 16), np.float32)
    q[:, 0:3] = rng.uniform(lo - 0.05 * ext, hi + 0.05 * ext, size=(n, 3))
    for base in (3, 6):
        d = rng.normal(size=(n, 3))
        q[:, base:base + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    q[:, 9] = rng.uniform(0, 1, n)
    q[:, 10:13] = rng.uniform(0, 1, (n, 3))
    if edge and n >= 16:
        q[0, 0:3] = lo
        q[1, 0:3] = hi
        q[2, 0:3] = hi + ext  # clamped to the far corner
        q[3, 3:6] = [0, 0, -1]
        q[4, 3:6] = [0, 0, 1]
        q[5, 6:9] = [1, 0, 0]
        q[6, 6:9] = [0, -1, 0]
        q[7, 9] = 0.0
        q[8, 9] = 1.0
    return q


def _levels(cs):
    return nrc.grid_levels(cs.grid_levels, cs.grid_log2_table, cs.grid_base_resolution, cs.grid_per_level_scale)


def _make(mipt_mod, scene, **over):
    ctx = mipt_mod.Context(0)
    ctx.set_scene(scene)
    cs = mipt_mod.default_radiance_cache_settings(**over)
    ctx.set_radiance_cache(cs)
    return ctx, cs


def _randomize_grid(ctx, cs, rng, scale=1.0):
    n = cs.grid_levels * (1 << cs.grid_log2_table) * 2
    g = (rng.uniform(-1, 1, n) * scale).astype(np.float32)
    ctx.set_cache_params(grid_master=g)
    return g


@pytest.mark.parametrize("over", [{}, {"grid_log2_table": 12, "grid_levels": 7},
                                  {"grid_log2_table": 14, "grid_levels": 16, "grid_base_resolution": 8}])
def test_encode_bit_exact(mipt_mod, torch_mod, scene1, over):
    rng = np.random.default_rng(11)
    ctx, cs = _make(mipt_mod, scene1, **over)
    _randomize_grid(ctx, cs, rng)
    n = 3001
    q = _queries(rng, scene1, n)
    d_q = mipt_mod.to_device(q)
    d_x = mipt_mod.empty_device(2 * 64 * n)
    ctx.cache_encode(d_q, n, d_x)
    torch_mod.cuda.synchronize()
    got = mipt_mod.from_device(d_x, np.uint16)[:64 * n].reshape(n, 64)
    _, _, _, gh = ctx.cache_params()
    bmin, inv = _bounds(scene1)
    want = nrc.encode(q, gh, _levels(cs), cs.grid_log2_table, bmin, inv)
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} encoded values differ, first {bad[:5].tolist()}"


def _infer(mipt_mod, torch_mod, ctx, q):
    n = q.shape[0]
    d_q = mipt_mod.to_device(q)
    d_o = mipt_mod.empty_device(16 * n)
    ctx.cache_infer(d_q, n, d_o)
    torch_mod.cuda.synchronize()
    return mipt_mod.from_device(d_o, np.float32)[:4 * n].reshape(n, 4)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 4099])
def test_infer_matches_oracle(mipt_mod, torch_mod, scene1, n):
    rng = np.random.default_rng(12 + n)
    ctx, cs = _make(mipt_mod, scene1)
    _randomize_grid(ctx, cs, rng)
    q = _queries(rng, scene1, n)
    out = _infer(mipt_mod, torch_mod, ctx, q)
    _, wb, _, gh = ctx.cache_params()
    bmin, inv = _bounds(scene1)
    x0 = nrc.bf16_value(nrc.encode(q, gh, _levels(cs), cs.grid_log2_table, bmin, inv))
    ref = nrc.forward(x0, nrc.bf16_value(wb).astype(np.float64))
    assert (out[:, 3] == 0).all()
    s = np.abs(ref) + np.abs(ref).mean()
    r = (np.abs(out[:, :3] - ref) / s).max(axis=1)
    print(f"rows {n}: max rel {r.max():.3g}, rows above 1e-4: {(r > 1e-4).sum()}")
    assert (r <= 2e-2).all(), f"max {r.max():.3g}, worst rows {np.argsort(-r)[:5]}"
    assert (r <= 1e-4).mean() >= 0.99


def _targets(rng, q):
    t = np.zeros((q.shape[0], 4), np.float32)
    t[:, :3] = rng.uniform(0, 2, (q.shape[0], 3))
    return t


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("n,batch", [(5000, 65536), (3000, 1024), (33, 65536)])
def test_gradients_match_oracle(mipt_mod, torch_mod, scene1, n, batch):
    rng = np.random.default_rng(13 + n)
    ctx, cs = _make(mipt_mod, scene1, batch_size=batch, grid_log2_table=16)
    _randomize_grid(ctx, cs, rng, 0.5)
    q = _queries(rng, scene1, n)
    t = _targets(rng, q)
    wg, gg, loss = ctx.cache_gradients(mipt_mod.to_device(q), mipt_mod.to_device(t), n)
    rows = min(n, batch)
    q, t = q[:rows], t[:rows]
    _, wb, _, gh = ctx.cache_params()
    bmin, inv = _bounds(scene1)
    levels = _levels(cs)
    x0 = nrc.bf16_value(nrc.encode(q, gh, levels, cs.grid_log2_table, bmin, inv))
    loss_sum, dW, dx0 = nrc.backward(x0, nrc.bf16_value(wb).astype(np.float64), t)
    dG = nrc.grid_gradient(q, dx0, levels, cs.grid_log2_table, bmin, inv)
    assert abs(loss - loss_sum / (3 * rows)) <= 1e-4 * abs(loss)
    for l in range(nrc.LAYERS):
        k, m = nrc.layer_shape(l)
        sl = slice(nrc.woff(l), nrc.woff(l) + k * m)
        assert _rel(wg[sl], dW[sl]) <= 5e-3, (l, _rel(wg[sl], dW[sl]))
    assert _rel(gg, dG) <= 5e-3, _rel(gg, dG)
    # entries no query touches stay exactly zero
    assert ((gg != 0) <= (dG != 0)).all()


def test_adam_step_matches_oracle(mipt_mod, torch_mod, scene1):
    rng = np.random.default_rng(14)
    ctx, cs = _make(mipt_mod, scene1, grid_log2_table=14)
    _randomize_grid(ctx, cs, rng, 0.5)
    n = 2048
    q = _queries(rng, scene1, n)
    t = _targets(rng, q)
    d_q, d_t = mipt_mod.to_device(q), mipt_mod.to_device(t)
    wm0, _, gm0, _ = ctx.cache_params()
    mw, vw = np.zeros_like(wm0), np.zeros_like(wm0)
    mg, vg = np.zeros_like(gm0), np.zeros_like(gm0)
    lr = cs.learning_rate
    for step in (1, 2):
        wg, gg, _ = ctx.cache_gradients(d_q, d_t, n)
        ctx.cache_train_step(d_q, d_t, n)
        wm1, wb1, gm1, gh1 = ctx.cache_params()
        want_w, mw, vw = nrc.adam_step(wm0, mw, vw, wg, step, lr, cs.adam_beta1, cs.adam_beta2, cs.adam_eps)
        want_g, mg, vg = nrc.adam_step(gm0, mg, vg, gg, step, lr, cs.adam_beta1, cs.adam_beta2, cs.adam_eps)
        assert np.abs(wm1 - want_w).max() <= 1e-3 * lr
        assert np.abs(gm1 - want_g).max() <= 1e-3 * lr
        np.testing.assert_array_equal(wb1, nrc.bf16_bits(wm1))
        np.testing.assert_array_equal(gh1, gm1.astype(np.float16).view(np.uint16))
        # every weight moved by about lr on the first step (sign of its gradient)
        if step == 1:
            moved = np.abs(wm1 - wm0)[wg != 0]
            assert np.median(moved) > 0.5 * lr
        wm0, gm0 = wm1, gm1


def test_training_reduces_loss(mipt_mod, torch_mod, scene1):
    torch = torch_mod
    ctx, cs = _make(mipt_mod, scene1, batch_size=16384)
    lo = torch.tensor(scene1.vertices.min(axis=0), device="cuda")
    hi = torch.tensor(scene1.vertices.max(axis=0), device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    n = 16384

    def batch():
        q = torch.zeros((n, 16), device="cuda")
        p = lo + (hi - lo) * torch.rand((n, 3), device="cuda", generator=g)
        d = torch.randn((n, 3), device="cuda", generator=g)
        d = d / d.norm(dim=1, keepdim=True)
        q[:, 0:3] = p
        q[:, 3:6] = d
        q[:, 6:9] = d
        q[:, 9] = 0.5
        q[:, 10:13] = 0.5
        u = (p - lo) / (hi - lo).max()
        t = torch.zeros((n, 4), device="cuda")
        t[:, 0] = 1.0 + 0.8 * torch.sin(6.0 * u[:, 0]) * torch.cos(4.0 * u[:, 1])
        t[:, 1] = 0.5 + 0.4 * u[:, 2] + 0.3 * d[:, 2].clamp(min=0)
        t[:, 2] = 0.2 + (u[:, 0] - 0.5) ** 2
        return q.contiguous(), t.contiguous()

    losses = []
    for step in range(300):
        q, t = batch()
        losses.append(ctx.cache_train_step(q, t, n, want_loss=(step < 5 or step >= 290)))
    first = np.mean([x for x in losses[:5]])
    last = np.mean([x for x in losses[290:]])
    print(f"relative L2 loss {first:.4g} -> {last:.4g}")
    assert last < 0.1 * first
    # the trained cache predicts fresh points of the field
    q, t = batch()
    d_o = torch.empty((n, 4), device="cuda")
    ctx.cache_infer(q, n, d_o)
    torch.cuda.synchronize()
    rel = ((d_o[:, :3] - t[:, :3]) ** 2 / (t[:, :3] ** 2 + 0.01)).mean().item()
    assert rel < 0.1 * first


def test_cache_errors(mipt_mod, torch_mod, scene1):
    ctx = mipt_mod.Context(0)
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_radiance_cache(mipt_mod.default_radiance_cache_settings())  # no scene
    ctx.set_scene(scene1)
    d = mipt_mod.empty_device(64 * 16)
    with pytest.raises(mipt_mod.MiptError):
        ctx.cache_infer(d, 16, d)  # not configured
    with pytest.raises(mipt_mod.MiptError):
        ctx.set_radiance_cache(mipt_mod.default_radiance_cache_settings(grid_levels=17))
