"""Pins of the radiance-cache oracle (oracle/nrc.py) on the CPU.

Each check is independent of the oracle's own formulas:
  - trilinear interpolation on a dense level reproduces a linear function of
    the grid coordinates exactly (catches swapped weights, wrong corner
    offsets or strides);
  - the spatial hash reduces to closed forms on the axes (prime 1 on x);
  - one-blob bins form a partition of unity and reproduce x linearly;
  - the octahedral map hits its closed-form values on the axes and inverts
    through a separately written decoder;
  - the MLP forward matches torch (float64 and with bfloat16 activations),
    and the bfloat16 rounding matches torch's conversion bit for bit;
  - the backward matches torch autograd for the weights and central finite
    differences of the loss for weights and grid features;
  - Adam matches torch.optim.Adam and the first-step closed form.
"""
import math

import numpy as np
import pytest

import nrc


def _levels_dense():
    # 3 dense levels in a 2^12 table: scale 3, 5, 8 -> res 5, 7, 10 points per axis
    return nrc.grid_levels(3, 12, 4, 1.5)


def test_level_schedule():
    lv = nrc.grid_levels(16, 19, 16, 1.3819)
    # scale_l = 16 * b^l - 1; level 15 spans about 2048 cells
    assert float(lv[0][0]) == 15.0 and lv[0][1] == 17
    assert abs(float(lv[15][0]) - (16 * 1.3819 ** 15 - 1)) < 1e-2
    assert 2000 < float(lv[15][0]) < 2100
    # dense exactly while res^3 <= 2^19
    for scale, res, dense in lv:
        assert dense == (res ** 3 <= 1 << 19)
    assert lv[0][2] and not lv[15][2]


def test_dense_level_reproduces_linear_function():
    rng = np.random.default_rng(3)
    levels = _levels_dense()
    assert all(d for _, _, d in levels)
    table = 1 << 12
    grid = np.zeros((len(levels), table, 2))
    coef = rng.normal(size=(len(levels), 2, 4))
    for l, (scale, res, _) in enumerate(levels):
        z, y, x = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
        idx = (x + y * res + z * res * res).reshape(-1)
        for f in range(2):
            c = coef[l, f]
            grid[l, idx, f] = (c[0] * x + c[1] * y + c[2] * z + c[3]).reshape(-1)
    n = 500
    q = np.zeros((n, 16), np.float32)
    q[:, 0:3] = rng.uniform(0.0, 1.0, size=(n, 3))
    q[:5, 0:3] = 1.0  # the far corner of the unit cube
    q[5:10, 0:3] = 0.0
    q[:, 3:6] = [0, 0, 1]
    q[:, 6:9] = [0, 0, 1]
    bmin, inv = nrc.bounds_transform([0, 0, 0], [1, 1, 1])
    x0 = nrc.encode_float64(q, grid.reshape(-1), levels, 12, bmin, inv)
    for l, (scale, res, _) in enumerate(levels):
        # corner g holds c.g + c3; the weights of the cell around s = p * scale + 0.5
        # interpolate it to c.s + c3 (sum_c w_c g_c = floor(s) + frac(s) = s)
        s = q[:, 0:3].astype(np.float64) * float(scale) + 0.5
        for f in range(2):
            c = coef[l, f]
            want = c[0] * s[:, 0] + c[1] * s[:, 1] + c[2] * s[:, 2] + c[3]
            np.testing.assert_allclose(x0[:, 2 * l + f], want, rtol=1e-5, atol=1e-4)


def test_hash_closed_forms_and_spread():
    levels = nrc.grid_levels(16, 19, 16, 1.3819)
    hashed = [L for L in levels if not L[2]][0]
    tmask = (1 << 19) - 1
    g = np.zeros((1000, 3), np.uint32)
    g[:, 0] = np.arange(1000)
    assert (nrc.hash_index(hashed, g, tmask) == np.arange(1000)).all()  # prime 1 on x
    g = np.zeros((5, 3), np.uint32)
    g[:, 1] = np.arange(5)
    want = [(y * 2654435761) % 2 ** 32 & tmask for y in range(5)]
    assert list(nrc.hash_index(hashed, g, tmask)) == want
    rng = np.random.default_rng(0)
    g = rng.integers(0, 2000, size=(1 << 16, 3)).astype(np.uint32)
    counts = np.bincount(nrc.hash_index(hashed, g, tmask) >> 13, minlength=64)  # 64 buckets
    exp = (1 << 16) / 64
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert chi2 < 130  # 63 degrees of freedom; p ~ 1e-6 bound


def test_one_blob_partition_of_unity():
    x = np.linspace(0.125, 0.875, 301).astype(np.float32)
    b = nrc.one_blob(x).astype(np.float64)
    np.testing.assert_allclose(b.sum(axis=1), 1.0, atol=1e-6)
    centers = (np.arange(nrc.BLOB) + 0.5) / nrc.BLOB
    np.testing.assert_allclose(b @ centers, x, atol=1e-6)
    peaks = nrc.one_blob(centers.astype(np.float32))
    np.testing.assert_array_equal(peaks, np.eye(nrc.BLOB, dtype=np.float32))


def _oct_decode(u, v):
    # inverse octahedral map written from its geometric definition
    x = 2.0 * u - 1.0
    y = 2.0 * v - 1.0
    z = 1.0 - np.abs(x) - np.abs(y)
    t = np.maximum(-z, 0.0)
    x = x + np.where(x >= 0, -t, t)
    y = y + np.where(y >= 0, -t, t)
    d = np.stack([x, y, z], axis=1)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_octahedral_map():
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    u, v = nrc.oct_uv(axes)
    np.testing.assert_array_equal(u, [1.0, 0.0, 0.5, 0.5, 0.5, 1.0])
    np.testing.assert_array_equal(v, [0.5, 0.5, 1.0, 0.0, 0.5, 1.0])
    rng = np.random.default_rng(1)
    d = rng.normal(size=(2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    u, v = nrc.oct_uv(d.astype(np.float32))
    assert (u >= 0).all() and (u <= 1).all() and (v >= 0).all() and (v <= 1).all()
    np.testing.assert_allclose(_oct_decode(u.astype(np.float64), v.astype(np.float64)), d, atol=2e-6)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=10000) * 10.0 ** rng.integers(-30, 30, size=10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.0078125]]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(nrc.bf16_bits(x), want)


def _random_weights(rng, gain=1.0):
    w = np.zeros(nrc.WEIGHTS)
    for l in range(nrc.LAYERS):
        k, n = nrc.layer_shape(l)
        w[nrc.woff(l):nrc.woff(l) + k * n] = rng.uniform(-1, 1, size=k * n) * gain * math.sqrt(6.0 / (k + n))
    return w


def _torch_forward(torch, x0, w, round_bf16):
    h = torch.from_numpy(np.asarray(x0, np.float64))
    W = [torch.from_numpy(np.ascontiguousarray(m)) for m in nrc.split_weights(w)]
    for l in range(nrc.LAYERS - 1):
        h = torch.relu(h @ W[l])
        if round_bf16:
            h = h.float().to(torch.bfloat16).double()
    return (h @ W[-1])[:, :3].numpy()


def test_forward_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    x0 = nrc.bf16_round(rng.uniform(-1, 1, size=(300, nrc.WIDTH)))
    w = nrc.bf16_round(_random_weights(rng, 2.0)).astype(np.float64)
    np.testing.assert_allclose(nrc.forward(x0, w, rounding=False), _torch_forward(torch, x0, w, False),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(nrc.forward(x0, w, rounding=True), _torch_forward(torch, x0, w, True),
                               rtol=1e-12, atol=1e-12)
    # ReLU MLPs without biases are positively homogeneous
    np.testing.assert_allclose(nrc.forward(2.5 * x0, w, rounding=False), 2.5 * nrc.forward(x0, w, rounding=False),
                               rtol=1e-10, atol=1e-12)


def _problem(seed=5, n=37):
    rng = np.random.default_rng(seed)
    levels = nrc.grid_levels(3, 10, 4, 2.0)  # dense, dense, hashed
    table = 1 << 10
    grid = rng.uniform(-1, 1, size=len(levels) * table * 2)
    q = np.zeros((n, 16), np.float32)
    q[:, 0:3] = rng.uniform(-1.0, 3.0, size=(n, 3))
    for base in (3, 6):
        d = rng.normal(size=(n, 3))
        q[:, base:base + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    q[:, 9] = rng.uniform(0, 1, n)
    q[:, 10:13] = rng.uniform(0, 1, (n, 3))
    targets = np.zeros((n, 4))
    targets[:, :3] = rng.uniform(0, 2, (n, 3))
    bmin, inv = nrc.bounds_transform([-1, -1, -1], [3, 2.5, 3])
    w = _random_weights(rng, 1.5)
    return levels, grid, q, targets, bmin, inv, w


def test_backward_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    levels, grid, q, targets, bmin, inv, w = _problem()
    x0 = nrc.encode_float64(q, grid, levels, 10, bmin, inv)
    loss_sum, dW, dx0 = nrc.backward(x0, w, targets, rounding=False)
    X = torch.from_numpy(x0).requires_grad_(True)
    Ws = [torch.from_numpy(np.ascontiguousarray(m)).clone().requires_grad_(True) for m in nrc.split_weights(w)]
    h = X
    for l in range(nrc.LAYERS - 1):
        h = torch.relu(h @ Ws[l])
    p = (h @ Ws[-1])[:, :3]
    y = torch.from_numpy(targets[:, :3])
    den = p.detach() ** 2 + 0.01
    loss = ((p - y) ** 2 / den).sum() / (3 * q.shape[0])
    loss.backward()
    want = np.concatenate([m.grad.numpy().reshape(-1) for m in Ws])
    np.testing.assert_allclose(dW, want, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(dx0 / (3 * q.shape[0]), X.grad.numpy(), rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(loss_sum / (3 * q.shape[0]), loss.item(), rtol=1e-12)


def test_gradients_match_finite_differences():
    levels, grid, q, targets, bmin, inv, w = _problem(seed=6, n=29)
    x0 = nrc.encode_float64(q, grid, levels, 10, bmin, inv)
    pred = nrc.forward(x0, w, rounding=False)
    den = pred * pred + 0.01
    _, dW, dx0 = nrc.backward(x0, w, targets, rounding=False)
    dG = nrc.grid_gradient(q, dx0, levels, 10, bmin, inv)

    def L(wv, gv):
        return nrc.surrogate_loss(q, targets, wv, gv, levels, 10, bmin, inv, den)

    rng = np.random.default_rng(7)
    h = 1e-6
    checked = 0
    for i in rng.choice(np.flatnonzero(np.abs(dW) > 1e-6), 25, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (L(wp, grid) - L(wm, grid)) / (2 * h)
        assert abs(fd - dW[i]) <= 1e-5 * abs(dW[i]) + 1e-9, (i, fd, dW[i])
        checked += 1
    for i in rng.choice(np.flatnonzero(np.abs(dG) > 1e-7), 25, replace=False):
        gp, gm = grid.copy(), grid.copy()
        gp[i] += h
        gm[i] -= h
        fd = (L(w, gp) - L(w, gm)) / (2 * h)
        assert abs(fd - dG[i]) <= 1e-5 * abs(dG[i]) + 1e-9, (i, fd, dG[i])
        checked += 1
    assert checked == 50
    # untouched grid entries have exactly zero gradient
    assert (dG != 0).sum() <= q.shape[0] * 8 * len(levels) * 2


def test_adam_matches_torch_and_first_step():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(8)
    p0 = rng.normal(size=1000).astype(np.float32)
    grads = [rng.normal(size=1000).astype(np.float32) * s for s in (1.0, 0.3, 2.0, 1e-3)]
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    p, m, v = p0.copy(), np.zeros(1000, np.float32), np.zeros(1000, np.float32)
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([tp], lr=lr, betas=(b1, b2), eps=eps)
    for step, g in enumerate(grads, start=1):
        p, m, v = nrc.adam_step(p, m, v, g, step, lr, b1, b2, eps)
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-6, atol=1e-7)
        if step == 1:
            # bias-corrected first step: p0 - lr * g / (|g| + eps)
            np.testing.assert_allclose(p, p0 - lr * grads[0] / (np.abs(grads[0]) + eps), rtol=1e-6, atol=1e-7)
