"""CPU restatement of the learned radiance cache -- the checker of the HIP kernels.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.

The cache has no counterpart in the reference renderer. The algorithms it
restates are published ones:
  - multiresolution hash encoding: Mueller, Evans, Schied, Keller, "Instant
    Neural Graphics Primitives with a Multiresolution Hash Encoding", SIGGRAPH
    2022, section 3 (level scales b^l * N_min, trilinear interpolation of the
    8 cell corners, spatial hash XOR of coordinates times the primes 1,
    2654435761, 805459861, dense indexing where the level fits the table);
  - one-blob input encoding and the relative L2 loss with a stop-gradient
    denominator: Mueller, Rousselle, Keller, Novak, "Real-time Neural Radiance
    Caching for Path Tracing", SIGGRAPH 2021, sections 5-6;
  - Adam with bias correction: Kingma & Ba, ICLR 2015, Algorithm 1.

Precision contract (the HIP kernels follow it and these functions mirror it):
  - encoding in float32 with the exact operation order of the device code
    (bit-exact parity: the result is rounded to bfloat16, round to nearest even);
  - the MLP consumes bfloat16 inputs and weights; its hidden activations are
    rounded to bfloat16 after the ReLU; products are summed in float32 on the
    device (float64 here, so parity is within a stated tolerance);
  - backward: per-sample loss gradient 2(p - y) / (p^2 + 0.01) rounded to
    bfloat16; hidden-layer gradients masked by (activation > 0) and rounded to
    bfloat16; weight and grid gradients in float32 (float64 here), scaled by
    1 / (3 * rows);
  - Adam in float32 with the device's operation order (bit-exact given the
    same gradient).

rounding=False switches every bfloat16 rounding off (pure float64), the form
the finite-difference and torch checks pin.
"""
from __future__ import annotations

import math

import numpy as np

WIDTH = 64
OUT = 16
LAYERS = 6
BLOB = 4
GRID_MAX = 32
WEIGHTS = 5 * WIDTH * WIDTH + WIDTH * OUT
PRIMES = (1, 2654435761, 805459861)


def woff(layer: int) -> int:
    return layer * WIDTH * WIDTH


def layer_shape(layer: int):
    return WIDTH, (OUT if layer == LAYERS - 1 else WIDTH)


def split_weights(flat):
    """Flat [K][N]-per-layer array -> list of 6 (K, N) matrices (views)."""
    out = []
    for l in range(LAYERS):
        k, n = layer_shape(l)
        out.append(np.asarray(flat)[woff(l):woff(l) + k * n].reshape(k, n))
    return out


# ---------------------------------------------------------------------------
# Number formats
# ---------------------------------------------------------------------------
def bf16_bits(x) -> np.ndarray:
    """float32 -> bfloat16 bits, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return ((u >> 16) & 0xFFFF).astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(x) -> np.ndarray:
    """Rounds to the nearest bfloat16 (via float32), returned as float32."""
    return bf16_value(bf16_bits(np.asarray(x, np.float32)))


def f16_value(bits) -> np.ndarray:
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float32)


# ---------------------------------------------------------------------------
# Settings-derived grid description (host code of the library, in double)
# ---------------------------------------------------------------------------
def grid_levels(levels: int, log2_table: int, base_resolution: int, per_level_scale: float):
    """[(scale float32, res, dense)] per level: scale_l = N_min * b^l - 1;
    corners reach floor(scale + 0.5) + 1 along an axis, so res = floor(scale
    + 0.5) + 2 points per axis; dense when res^3 fits the 2^log2_table table."""
    b = float(np.float32(per_level_scale))
    table = math.ldexp(1.0, log2_table)
    out = []
    for l in range(levels):
        scale = float(base_resolution) * math.pow(b, l) - 1.0
        res = int(math.floor(float(np.float32(scale)) + 0.5)) + 2
        out.append((np.float32(scale), res, float(res) ** 3 <= table))
    return out


def bounds_transform(bmin, bmax):
    """(bmin float32[3], inv_extent float32): positions map to
    clamp((p - bmin) * inv_extent, 0, 1), one scale for all axes."""
    bmin = np.asarray(bmin, np.float32)
    bmax = np.asarray(bmax, np.float32)
    ext = np.float32(0.0)
    for a in range(3):
        ext = np.float32(max(ext, np.float32(bmax[a] - bmin[a])))
    inv = np.float32(np.float32(1.0) / ext) if ext > 0 else np.float32(1.0)
    return bmin, inv


# ---------------------------------------------------------------------------
# Encoding
# ---------------------------------------------------------------------------
def normalized_positions(q, bmin, inv):
    q = np.asarray(q, np.float32)
    p = (q[:, 0:3] - bmin[None, :]).astype(np.float32) * inv
    return np.minimum(np.float32(1.0), np.maximum(np.float32(0.0), p.astype(np.float32)))


def hash_index(level, g, tmask):
    """Table index of integer corner coordinates g (uint32 [n, 3])."""
    scale, res, dense = level
    g = g.astype(np.uint64)
    if dense:
        r = np.uint64(res)
        idx = (g[:, 0] + g[:, 1] * r + g[:, 2] * r * r) & np.uint64(0xFFFFFFFF)
    else:
        idx = np.zeros(g.shape[0], np.uint64)
        for a in range(3):
            idx ^= (g[:, a] * np.uint64(PRIMES[a])) & np.uint64(0xFFFFFFFF)
    return (idx & np.uint64(tmask)).astype(np.int64)


def level_corners(level, pos, tmask):
    """For one level: [(index int64[n], weight float32[n])] for the 8 corners
    c (bit a of c = +1 along axis a)."""
    scale = level[0]
    s = (pos * scale).astype(np.float32) + np.float32(0.5)
    f = np.floor(s).astype(np.float32)
    fr = (s - f).astype(np.float32)
    base = f.astype(np.uint32)
    out = []
    for c in range(8):
        bits = np.array([(c >> a) & 1 for a in range(3)], np.uint32)
        g = base + bits[None, :]
        wk = np.where(bits[None, :] == 1, fr, np.float32(1.0) - fr).astype(np.float32)
        w = ((wk[:, 0] * wk[:, 1]).astype(np.float32) * wk[:, 2]).astype(np.float32)
        out.append((hash_index(level, g, tmask), w))
    return out


def oct_uv(d):
    """Octahedral map of unit vectors to [0,1]^2 (device operation order)."""
    d = np.asarray(d, np.float32)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    s = ((np.abs(x) + np.abs(y)).astype(np.float32) + np.abs(z)).astype(np.float32)
    px = (x / s).astype(np.float32)
    py = (y / s).astype(np.float32)
    sx = np.where(px >= 0, np.float32(1.0), np.float32(-1.0))
    sy = np.where(py >= 0, np.float32(1.0), np.float32(-1.0))
    ox = ((np.float32(1.0) - np.abs(py)) * sx).astype(np.float32)
    oy = ((np.float32(1.0) - np.abs(px)) * sy).astype(np.float32)
    neg = z < 0
    px = np.where(neg, ox, px)
    py = np.where(neg, oy, py)
    u = (px * np.float32(0.5)).astype(np.float32) + np.float32(0.5)
    v = (py * np.float32(0.5)).astype(np.float32) + np.float32(0.5)
    return u.astype(np.float32), v.astype(np.float32)


def one_blob(x):
    """[n, BLOB] hat-kernel bins: max(0, 1 - |x * BLOB - (i + 0.5)|)."""
    t = (np.asarray(x, np.float32) * np.float32(BLOB)).astype(np.float32)
    cols = []
    for i in range(BLOB):
        cols.append(np.maximum(np.float32(0.0),
                               np.float32(1.0) - np.abs((t - np.float32(i + 0.5)).astype(np.float32))))
    return np.stack(cols, axis=1).astype(np.float32)


def encode(queries, grid_f16_bits, levels, log2_table, bmin, inv):
    """Encoded rows (bfloat16 bits [n, 64]) of 16-float query records:
    position[3], direction[3], normal[3], roughness, albedo[3], unused[3].
    grid_f16_bits: [levels, 2^log2_table, 2] FP16 bits (or flat)."""
    q = np.asarray(queries, np.float32).reshape(-1, 16)
    n = q.shape[0]
    table = 1 << log2_table
    grid = f16_value(np.asarray(grid_f16_bits).reshape(-1)).reshape(len(levels), table, 2)
    tmask = table - 1
    pos = normalized_positions(q, bmin, inv)
    out = np.zeros((n, WIDTH), np.float32)
    for l, L in enumerate(levels):
        f0 = np.zeros(n, np.float32)
        f1 = np.zeros(n, np.float32)
        for idx, w in level_corners(L, pos, tmask):
            f0 = (f0 + (w * grid[l, idx, 0]).astype(np.float32)).astype(np.float32)
            f1 = (f1 + (w * grid[l, idx, 1]).astype(np.float32)).astype(np.float32)
        out[:, 2 * l] = f0
        out[:, 2 * l + 1] = f1
    k = 2 * len(levels)
    for base in (3, 6):
        u, v = oct_uv(q[:, base:base + 3])
        out[:, k:k + BLOB] = one_blob(u)
        out[:, k + BLOB:k + 2 * BLOB] = one_blob(v)
        k += 2 * BLOB
    out[:, k:k + BLOB] = one_blob(q[:, 9])
    k += BLOB
    out[:, k:k + 3] = q[:, 10:13]
    out[:, k + 3] = 1.0
    return bf16_bits(out)


# ---------------------------------------------------------------------------
# MLP
# ---------------------------------------------------------------------------
def _round_act(x, rounding):
    return bf16_round(x.astype(np.float32)).astype(np.float64) if rounding else x


def forward(x0, weights, rounding=True, keep=False):
    """x0: [n, 64] input values; weights: flat [K][N]-per-layer values.
    Returns predicted RGB [n, 3] (float64) and, with keep, the layer inputs."""
    W = [w.astype(np.float64) for w in split_weights(np.asarray(weights, np.float64))]
    h = np.asarray(x0, np.float64)
    xs = []
    for l in range(LAYERS - 1):
        xs.append(h)
        h = _round_act(np.maximum(h @ W[l], 0.0), rounding)
    xs.append(h)
    pred = (h @ W[LAYERS - 1])[:, :3]
    return (pred, xs) if keep else pred


def loss_terms(pred, targets, den=None):
    """Relative L2 terms (p - y)^2 / den with den = p^2 + 0.01 unless given."""
    y = np.asarray(targets, np.float64).reshape(-1, 4)[:, :3]
    if den is None:
        den = pred * pred + 0.01
    d = pred - y
    return d * d / den, d, den


def backward(x0, weights, targets, rounding=True):
    """Loss sum and gradients of mean over rows and RGB of the relative L2
    loss (denominator held constant): (loss_sum, dW flat [K][N]-per-layer,
    dX0 [n, 64]) -- dX0 unscaled, as the grid scatter applies 1 / (3 rows)."""
    W = [w.astype(np.float64) for w in split_weights(np.asarray(weights, np.float64))]
    pred, xs = forward(x0, weights, rounding, keep=True)
    if rounding:
        pred = pred.astype(np.float32).astype(np.float64)
    terms, d, den = loss_terms(pred, targets)
    rows = pred.shape[0]
    scale = 1.0 / (3.0 * rows)
    g = 2.0 * d / den
    dy = np.zeros((rows, OUT))
    dy[:, :3] = _round_act(g, rounding)
    dW = [None] * LAYERS
    dx0 = None
    for l in range(LAYERS - 1, -1, -1):
        dW[l] = (xs[l].T @ dy) * scale
        dx = dy @ W[l].T
        if l > 0:
            dy = _round_act(np.where(xs[l] > 0.0, dx, 0.0), rounding)
        else:
            dx0 = dx
    flat = np.zeros(WEIGHTS)
    for l in range(LAYERS):
        k, n = layer_shape(l)
        flat[woff(l):woff(l) + k * n] = dW[l].reshape(-1)
    return terms.sum(), flat, dx0


def grid_gradient(queries, dx0, levels, log2_table, bmin, inv):
    """Scatter of the feature gradients dX0[:, :2L] * 1 / (3 rows) to the 8
    corners of every level: flat [levels * 2^log2_table * 2]."""
    q = np.asarray(queries, np.float32).reshape(-1, 16)
    rows = q.shape[0]
    table = 1 << log2_table
    pos = normalized_positions(q, bmin, inv)
    out = np.zeros((len(levels), table, 2))
    scale = 1.0 / (3.0 * rows)
    for l, L in enumerate(levels):
        for idx, w in level_corners(L, pos, table - 1):
            for f in range(2):
                np.add.at(out[l, :, f], idx, w.astype(np.float64) * dx0[:, 2 * l + f] * scale)
    return out.reshape(-1)


def surrogate_loss(queries, targets, weights, grid, levels, log2_table, bmin, inv, den):
    """Float64 loss with a frozen denominator, as a function of the FP32
    parameters (no rounding anywhere): mean over rows and RGB of
    (p - y)^2 / den. Its gradient is what backward + grid_gradient compute."""
    q = np.asarray(queries, np.float32).reshape(-1, 16)
    x0 = encode_float64(q, grid, levels, log2_table, bmin, inv)
    pred = forward(x0, weights, rounding=False)
    terms, _, _ = loss_terms(pred, targets, den)
    return terms.sum() / (3.0 * pred.shape[0])


def encode_float64(q, grid, levels, log2_table, bmin, inv):
    """encode() without rounding: grid features from FP32/64 values
    (grid: flat float [levels * table * 2]), float64 output."""
    n = q.shape[0]
    table = 1 << log2_table
    g = np.asarray(grid, np.float64).reshape(len(levels), table, 2)
    pos = normalized_positions(q, bmin, inv)
    out = np.zeros((n, WIDTH))
    for l, L in enumerate(levels):
        for idx, w in level_corners(L, pos, table - 1):
            out[:, 2 * l] += w * g[l, idx, 0]
            out[:, 2 * l + 1] += w * g[l, idx, 1]
    k = 2 * len(levels)
    for base in (3, 6):
        u, v = oct_uv(q[:, base:base + 3])
        out[:, k:k + BLOB] = one_blob(u)
        out[:, k + BLOB:k + 2 * BLOB] = one_blob(v)
        k += 2 * BLOB
    out[:, k:k + BLOB] = one_blob(q[:, 9])
    k += BLOB
    out[:, k:k + 3] = q[:, 10:13]
    out[:, k + 3] = 1.0
    return out


# ---------------------------------------------------------------------------
# Adam (float32, device operation order)
# ---------------------------------------------------------------------------
def adam_step(master, m, v, grad, step, lr, beta1, beta2, eps):
    """One Adam update in place-free form; returns (master, m, v). step is the
    1-based step number; the bias corrections are 1 - beta^step computed in
    double and rounded to float32, as the library's host code does."""
    f = np.float32
    g = np.asarray(grad, np.float32)
    b1, b2 = f(beta1), f(beta2)
    b1t = f(1.0 - math.pow(float(b1), step))
    b2t = f(1.0 - math.pow(float(b2), step))
    m = (b1 * np.asarray(m, np.float32)).astype(np.float32) + ((f(1.0) - b1) * g).astype(np.float32)
    m = m.astype(np.float32)
    v = (b2 * np.asarray(v, np.float32)).astype(np.float32) + (((f(1.0) - b2) * g).astype(np.float32) * g)
    v = v.astype(np.float32)
    mh = (m / b1t).astype(np.float32)
    vh = (v / b2t).astype(np.float32)
    upd = ((f(lr) * mh).astype(np.float32) / (np.sqrt(vh).astype(np.float32) + f(eps))).astype(np.float32)
    p = (np.asarray(master, np.float32) - upd).astype(np.float32)
    return p, m, v
