// nrc.hip -- learned radiance cache: a multiresolution hash-grid encoding of
// the position plus one-blob encodings of direction, normal, roughness and
// albedo (64 inputs), a 6-matrix ReLU MLP of width 64 with BF16 weights on the
// MFMA units (FP32 accumulation), a relative-L2 loss and Adam, all trained on
// the GPU from the tracer's own paths.
//
// MFMA: v_mfma_f32_16x16x32_bf16, one wave per 16-row tile. Operand layout
// of lane l (0..63): A[m = l%16][k = 8*(l/16) + j], B[k = 8*(l/16) + j][n = l%16]
// for j < 8; D[m = 4*(l/16) + i][n = l%16] for i < 4. Activations move between
// layers through a padded LDS tile (D layout -> row-major -> A layout).
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nrc.h"

namespace mipt {
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int TILE_LD = NRC_WIDTH + 8; // LDS row stride (bf16) of an activation tile

// binary32 -> bfloat16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ v8bf load_v8(const uint16_t* p)
{
    const uint4 q = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(v8bf, q);
}

__device__ __forceinline__ v4f mfma(v8bf a, v8bf b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Encoding
// ---------------------------------------------------------------------------
__device__ __forceinline__ void oct_uv(float x, float y, float z, float& u, float& v)
{
    const float s = fabsf(x) + fabsf(y) + fabsf(z);
    float px = x / s, py = y / s;
    if (z < 0.0f)
    {
        const float ox = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        const float oy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = ox;
        py = oy;
    }
    u = px * 0.5f + 0.5f;
    v = py * 0.5f + 0.5f;
}

// One-blob with NRC_BLOB bins: bin i holds max(0, 1 - |x * NRC_BLOB - (i + 0.5)|).
__device__ __forceinline__ void one_blob(float x, uint16_t* out)
{
    const float t = x * (float)NRC_BLOB;
    for (int i = 0; i < NRC_BLOB; i++)
        out[i] = to_bf16(fmaxf(0.0f, 1.0f - fabsf(t - ((float)i + 0.5f))));
}

__device__ __forceinline__ uint32_t grid_index(const NrcGridLevel& L, uint32_t x, uint32_t y, uint32_t z, uint32_t tmask)
{
    if (L.dense) // res^3 fits the table; the mask keeps any index inside it
        return (x + y * L.res + z * L.res * L.res) & tmask;
    return ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & tmask;
}

// Corner c of the cell containing the scaled position, its weight and index.
__device__ __forceinline__ void corner(const NrcGridLevel& L, const float pos[3], int c, uint32_t tmask, uint32_t& idx,
                                       float& w)
{
    uint32_t g[3];
    float wk[3];
    for (int a = 0; a < 3; a++)
    {
        const float s = pos[a] * L.scale + 0.5f;
        const float f = floorf(s);
        const float fr = s - f;
        const uint32_t gi = (uint32_t)f + (((c >> a) & 1) ? 1u : 0u);
        g[a] = gi;
        wk[a] = ((c >> a) & 1) ? fr : 1.0f - fr;
    }
    w = (wk[0] * wk[1]) * wk[2];
    idx = grid_index(L, g[0], g[1], g[2], tmask);
}

__device__ __forceinline__ void normalized_position(const NrcDev& P, const float* q, float pos[3])
{
    for (int a = 0; a < 3; a++)
        pos[a] = fminf(1.0f, fmaxf(0.0f, (q[a] - P.bmin[a]) * P.inv_extent));
}

// Query record (16 floats): position, direction, normal, roughness, albedo.
// Encodes the selected rows into x0 rows [0, rows) and zero-fills the rows up
// to the next multiple of pad (at most n rows in all).
//
// Launch: NRC_XCDS x chunks blocks of 128 threads (encode_launch). Workgroups
// go to the 8 XCDs round-robin, so block b runs on XCD b % 8 and interpolates
// the levels l with l % 8 == b % 8 for rows [128 (b / 8), +128): each XCD's
// L2 (4 MB) then holds the tables of two levels (2 MB of FP16 pairs each at
// 2^19 entries) instead of missing on all sixteen. The block whose chunk
// index matches its XCD (chunk % 8 == b % 8) also writes the other inputs
// and the zero padding of its rows.
constexpr uint32_t NRC_XCDS = 8;

__global__ __launch_bounds__(128) void k_nrc_encode(NrcDev P, const float4* __restrict__ queries, RowSel sel,
                                                    uint32_t n, uint32_t pad, uint16_t* __restrict__ x0)
{
    const uint32_t xcd = blockIdx.x % NRC_XCDS;
    const uint32_t chunk = blockIdx.x / NRC_XCDS;
    const uint32_t row = chunk * blockDim.x + threadIdx.x;
    const uint32_t rows = sel.rows();
    const uint32_t padded = (rows + pad - 1) / pad * pad;
    if (row >= n || row >= padded)
        return;
    uint16_t* out = x0 + (size_t)row * NRC_WIDTH;
    const bool misc = chunk % NRC_XCDS == xcd;
    if (row >= rows)
    {
        if (misc)
            for (int k = 0; k < NRC_WIDTH; k += 8)
                *reinterpret_cast<uint4*>(out + k) = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const size_t rec = sel.record(row);
    const float4 a = queries[4 * rec];
    const float q3[3] = {a.x, a.y, a.z};
    float pos[3];
    normalized_position(P, q3, pos);
    const uint32_t tmask = (1u << P.log2_table) - 1u;
    for (int l = (int)xcd; l < P.levels; l += (int)NRC_XCDS)
    {
        const NrcGridLevel& L = P.level[l];
        float f0 = 0.0f, f1 = 0.0f;
        for (int cc = 0; cc < 8; cc++)
        {
            uint32_t idx;
            float w;
            corner(L, pos, cc, tmask, idx, w);
            const size_t e = ((size_t)l << P.log2_table) + idx;
            const uint32_t pair = reinterpret_cast<const uint32_t*>(P.grid_f16)[e];
            f0 += w * (float)__builtin_bit_cast(_Float16, (uint16_t)(pair & 0xffffu));
            f1 += w * (float)__builtin_bit_cast(_Float16, (uint16_t)(pair >> 16));
        }
        *reinterpret_cast<uint32_t*>(out + 2 * l) = (uint32_t)to_bf16(f0) | ((uint32_t)to_bf16(f1) << 16);
    }
    if (!misc)
        return;
    const float4 b = queries[4 * rec + 1], c = queries[4 * rec + 2], d = queries[4 * rec + 3];
    const float q[16] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
    int k = 2 * P.levels;
    float u, v;
    oct_uv(q[3], q[4], q[5], u, v);
    one_blob(u, out + k);
    one_blob(v, out + k + NRC_BLOB);
    k += 2 * NRC_BLOB;
    oct_uv(q[6], q[7], q[8], u, v);
    one_blob(u, out + k);
    one_blob(v, out + k + NRC_BLOB);
    k += 2 * NRC_BLOB;
    one_blob(q[9], out + k);
    k += NRC_BLOB;
    out[k++] = to_bf16(q[10]);
    out[k++] = to_bf16(q[11]);
    out[k++] = to_bf16(q[12]);
    out[k++] = to_bf16(1.0f);
    for (; k < NRC_WIDTH; k++)
        out[k] = 0;
}

void encode_launch(const NrcDev& P, const float* d_queries, const RowSel& sel, uint32_t n, uint32_t pad,
                   uint16_t* x0, hipStream_t s)
{
    const unsigned chunks = (unsigned)((n + 127u) / 128u);
    hipLaunchKernelGGL(k_nrc_encode, dim3(NRC_XCDS * chunks), dim3(128), 0, s, P,
                       reinterpret_cast<const float4*>(d_queries), sel, n, pad, x0);
}

// ---------------------------------------------------------------------------
// Fused MLP
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_weights_lds(const uint16_t* __restrict__ src, uint16_t* dst, int count)
{
    for (int i = threadIdx.x * 8; i < count; i += blockDim.x * 8)
        *reinterpret_cast<uint4*>(dst + i) = *reinterpret_cast<const uint4*>(src + i);
}

// One layer on a 16-row tile: A fragments (two k-steps) times W^T (n-major
// [N][64] in LDS) -> D for the n-block nb.
__device__ __forceinline__ v4f layer_block(const v8bf a[2], const uint16_t* wt, int nb, int lane)
{
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int n = 16 * nb + (lane & 15);
    for (int s = 0; s < 2; s++)
        acc = mfma(a[s], load_v8(wt + n * NRC_WIDTH + 32 * s + 8 * (lane >> 4)), acc);
    return acc;
}

// Inference: out[row] = MLP(x0[row]) (RGB + 0), rows < count.
__global__ __launch_bounds__(256) void k_nrc_forward(const uint16_t* __restrict__ wt_global,
                                                     const uint16_t* __restrict__ x0, const uint32_t* count,
                                                     uint32_t n, float4* __restrict__ out)
{
    __shared__ uint16_t wt[NRC_WEIGHTS];
    __shared__ uint16_t tiles[4][16 * TILE_LD];
    load_weights_lds(wt_global, wt, NRC_WEIGHTS);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t rows = *count < n ? *count : n;
    uint16_t* tile = tiles[wave];
    const uint32_t tiles_total = (rows + 15) / 16;
    for (uint32_t t = blockIdx.x * 4 + wave; t < tiles_total; t += gridDim.x * 4)
    {
        const uint32_t r0 = t * 16;
        v8bf a[2];
        for (int s = 0; s < 2; s++)
            a[s] = load_v8(x0 + (size_t)(r0 + (lane & 15)) * NRC_WIDTH + 32 * s + 8 * (lane >> 4));
        for (int layer = 0; layer < NRC_LAYERS; layer++)
        {
            const uint16_t* w = wt + NRC_WOFF(layer);
            if (layer < NRC_LAYERS - 1)
            {
                for (int nb = 0; nb < 4; nb++)
                {
                    const v4f acc = layer_block(a, w, nb, lane);
                    for (int i = 0; i < 4; i++)
                        tile[(4 * (lane >> 4) + i) * TILE_LD + 16 * nb + (lane & 15)] = to_bf16(fmaxf(acc[i], 0.0f));
                }
                __builtin_amdgcn_s_waitcnt(0xc07f);
                __builtin_amdgcn_wave_barrier();
                for (int s = 0; s < 2; s++)
                    a[s] = load_v8(tile + (lane & 15) * TILE_LD + 32 * s + 8 * (lane >> 4));
                __builtin_amdgcn_wave_barrier();
            }
            else
            {
                const v4f acc = layer_block(a, w, 0, lane);
                const int col = lane & 15;
                if (col < 4)
                    for (int i = 0; i < 4; i++)
                    {
                        const uint32_t row = r0 + 4 * (lane >> 4) + i;
                        if (row < rows)
                            reinterpret_cast<float*>(out)[4 * (size_t)row + col] = col < 3 ? acc[i] : 0.0f;
                    }
            }
        }
    }
}

// Training: forward, loss gradient and backward of one 16-row tile per wave.
// Writes the layer inputs X_l and output gradients dY_l column-major
// ([64][batch_stride], bf16) for the weight-gradient GEMMs, dX0 (float,
// [batch][2 * levels]) for the grid, and adds the loss to *loss_sum. Tiles
// cover the rows rounded up to 32 (the k-step of the dW GEMM); the encoded
// rows past the count are zero, so their activations and gradients are zero.
__global__ __launch_bounds__(256) void k_nrc_train_fwd_bwd(NrcTrainDev T)
{
    __shared__ uint16_t wt[NRC_WEIGHTS];   // W^T, [N][K]
    __shared__ uint16_t wk[NRC_WEIGHTS];   // W, [K][N]
    __shared__ uint16_t acts[4][NRC_LAYERS][16 * TILE_LD];
    load_weights_lds(T.wt, wt, NRC_WEIGHTS);
    load_weights_lds(T.wk, wk, NRC_WEIGHTS);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t rows = T.sel.rows();
    const uint32_t tiles_total = ((rows + 31) & ~31u) / 16;
    const uint32_t S = T.batch_stride;
    for (uint32_t t = blockIdx.x * 4 + wave; t < tiles_total; t += gridDim.x * 4)
    {
        const uint32_t r0 = t * 16;
        // layer-0 input tile (row-major in LDS) from the encoded rows
        uint16_t* x_in = acts[wave][0];
        for (int s = 0; s < 2; s++)
        {
            const int r = lane & 15, k0 = 32 * s + 8 * (lane >> 4);
            *reinterpret_cast<uint4*>(x_in + r * TILE_LD + k0) =
                *reinterpret_cast<const uint4*>(T.x0 + (size_t)(r0 + r) * NRC_WIDTH + k0);
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        v4f pred = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int layer = 0; layer < NRC_LAYERS; layer++)
        {
            const uint16_t* xin = acts[wave][layer];
            // column-major copy of this layer's input for dW: lane = column
            {
                uint16_t* dst = T.xcm + (size_t)layer * NRC_WIDTH * S + (size_t)lane * S + r0;
                for (int r = 0; r < 16; r++)
                    dst[r] = xin[r * TILE_LD + lane];
            }
            v8bf a[2];
            for (int s = 0; s < 2; s++)
                a[s] = load_v8(xin + (lane & 15) * TILE_LD + 32 * s + 8 * (lane >> 4));
            const uint16_t* w = wt + NRC_WOFF(layer);
            if (layer < NRC_LAYERS - 1)
            {
                uint16_t* xo = acts[wave][layer + 1];
                for (int nb = 0; nb < 4; nb++)
                {
                    const v4f acc = layer_block(a, w, nb, lane);
                    for (int i = 0; i < 4; i++)
                        xo[(4 * (lane >> 4) + i) * TILE_LD + 16 * nb + (lane & 15)] = to_bf16(fmaxf(acc[i], 0.0f));
                }
                __builtin_amdgcn_s_waitcnt(0xc07f);
                __builtin_amdgcn_wave_barrier();
            }
            else
                pred = layer_block(a, w, 0, lane);
        }
        // relative L2: sum over rows and RGB of (p - y)^2 / (p^2 + 0.01); the
        // denominator is held constant in the gradient.
        uint16_t* dy = acts[wave][0]; // reuse: dY of the current layer, row-major
        float lsum = 0.0f;
        {
            const int col = lane & 15;
            for (int i = 0; i < 4; i++)
            {
                const uint32_t row = r0 + 4 * (lane >> 4) + i;
                float g = 0.0f;
                if (col < 3 && row < rows)
                {
                    const float p = pred[i];
                    const float y = T.targets[4 * T.sel.record(row) + col];
                    const float den = p * p + 0.01f;
                    const float d = p - y;
                    lsum += d * d / den;
                    g = 2.0f * d / den;
                }
                dy[(4 * (lane >> 4) + i) * TILE_LD + col] = to_bf16(g);
                dy[(4 * (lane >> 4) + i) * TILE_LD + 16 + col] = 0;
            }
        }
        for (int off = 32; off > 0; off >>= 1)
            lsum += __shfl_down(lsum, off, 64);
        if (lane == 0)
            atomicAdd(T.loss_sum, lsum);
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        for (int layer = NRC_LAYERS - 1; layer >= 0; layer--)
        {
            const int N = layer == NRC_LAYERS - 1 ? NRC_OUT : NRC_WIDTH;
            // column-major copy of dY_layer for dW (lanes < N are columns)
            if (lane < N)
            {
                uint16_t* dst = T.dycm + (size_t)layer * NRC_WIDTH * S + (size_t)lane * S + r0;
                for (int r = 0; r < 16; r++)
                    dst[r] = dy[r * TILE_LD + lane];
            }
            // dX = dY W^T: A = dY (k' = n), B[k' = n][col = k] = W[k][n] from the [K][N] layout
            v8bf a[2];
            for (int s = 0; s < 2; s++)
            {
                const int n0 = 32 * s + 8 * (lane >> 4);
                if (n0 < N)
                    a[s] = load_v8(dy + (lane & 15) * TILE_LD + n0);
                else
                    a[s] = __builtin_bit_cast(v8bf, make_uint4(0u, 0u, 0u, 0u));
            }
            const uint16_t* w = wk + NRC_WOFF(layer);
            __builtin_amdgcn_s_waitcnt(0xc07f);
            __builtin_amdgcn_wave_barrier();
            for (int kb = 0; kb < 4; kb++)
            {
                v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
                const int k = 16 * kb + (lane & 15);
                for (int s = 0; s < 2; s++)
                {
                    const int n0 = 32 * s + 8 * (lane >> 4);
                    const v8bf b = n0 < N ? load_v8(w + k * N + n0) : __builtin_bit_cast(v8bf, make_uint4(0u, 0u, 0u, 0u));
                    acc = mfma(a[s], b, acc);
                }
                if (layer > 0)
                {
                    // ReLU mask of the layer input (the previous layer's output)
                    const uint16_t* xin = acts[wave][layer];
                    for (int i = 0; i < 4; i++)
                    {
                        const int r = 4 * (lane >> 4) + i;
                        const float m = from_bf16(xin[r * TILE_LD + k]) > 0.0f ? acc[i] : 0.0f;
                        dy[r * TILE_LD + k] = to_bf16(m);
                    }
                }
                else
                {
                    // gradient of the grid features (columns < 2 * levels)
                    if (k < 2 * T.levels)
                        for (int i = 0; i < 4; i++)
                        {
                            const uint32_t row = r0 + 4 * (lane >> 4) + i;
                            T.dx0[(size_t)row * NRC_GRID_MAX + k] = row < rows ? acc[i] : 0.0f;
                        }
                }
            }
            __builtin_amdgcn_s_waitcnt(0xc07f);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// dW_l[k][n] += scale * sum_r X_l[r][k] dY_l[r][n] over one chunk of rows.
// Block = 4 waves = the 4 k-blocks of 16; each wave keeps the n-blocks of its
// k-block in registers and adds them to the FP32 gradient once.
__global__ __launch_bounds__(256) void k_nrc_weight_grad(NrcTrainDev T, uint32_t chunk)
{
    const int layer = blockIdx.y;
    const int lane = threadIdx.x & 63, kb = threadIdx.x >> 6;
    const uint32_t rows = T.sel.rows();
    const uint32_t rows_pad = (rows + 31) & ~31u;
    const uint32_t begin = blockIdx.x * chunk;
    if (begin >= rows_pad)
        return;
    const uint32_t end = begin + chunk < rows_pad ? begin + chunk : rows_pad;
    const int N = layer == NRC_LAYERS - 1 ? NRC_OUT : NRC_WIDTH;
    const int nbs = N / 16;
    const uint32_t S = T.batch_stride;
    const uint16_t* X = T.xcm + (size_t)layer * NRC_WIDTH * S;
    const uint16_t* D = T.dycm + (size_t)layer * NRC_WIDTH * S;
    v4f acc[4];
    for (int nb = 0; nb < 4; nb++)
        acc[nb] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const int k = 16 * kb + (lane & 15);
    for (uint32_t r = begin; r < end; r += 32)
    {
        const uint32_t rr = r + 8 * (lane >> 4);
        const v8bf a = load_v8(X + (size_t)k * S + rr);
        for (int nb = 0; nb < nbs; nb++)
        {
            const v8bf b = load_v8(D + (size_t)(16 * nb + (lane & 15)) * S + rr);
            acc[nb] = mfma(a, b, acc[nb]);
        }
    }
    const float scale = 1.0f / (3.0f * (float)rows);
    float* g = T.wgrad + NRC_WOFF(layer);
    for (int nb = 0; nb < nbs; nb++)
        for (int i = 0; i < 4; i++)
        {
            const int kk = 16 * kb + 4 * (lane >> 4) + i;
            const int nn = 16 * nb + (lane & 15);
            atomicAdd(g + kk * N + nn, acc[nb][i] * scale);
        }
}

// Grid gradient: scatters the feature gradients to the 8 corners of a level.
__global__ __launch_bounds__(128) void k_nrc_grid_grad(NrcDev P, NrcTrainDev T, const float4* __restrict__ queries)
{
    // One thread per (row, level): a batch has few rows, and a row's 16 levels
    // x 8 corners x 2 atomics in one thread left the GPU latency-bound. Blocks
    // map to XCDs as in k_nrc_encode (levels l % 8 == block % 8), so each
    // XCD's atomics stay within the tables of two levels.
    const uint32_t xcd = blockIdx.x % NRC_XCDS;
    const uint32_t row = (blockIdx.x / NRC_XCDS) * blockDim.x + threadIdx.x;
    const uint32_t rows = T.sel.rows();
    if (row >= rows)
        return;
    const float4 a = queries[4 * T.sel.record(row)];
    const float q[3] = {a.x, a.y, a.z};
    float pos[3];
    normalized_position(P, q, pos);
    const uint32_t tmask = (1u << P.log2_table) - 1u;
    const float scale = 1.0f / (3.0f * (float)rows);
    for (int l = (int)xcd; l < P.levels; l += (int)NRC_XCDS)
    {
        const NrcGridLevel& L = P.level[l];
        const float g0 = T.dx0[(size_t)row * NRC_GRID_MAX + 2 * l] * scale;
        const float g1 = T.dx0[(size_t)row * NRC_GRID_MAX + 2 * l + 1] * scale;
        for (int cc = 0; cc < 8; cc++)
        {
            uint32_t idx;
            float w;
            corner(L, pos, cc, tmask, idx, w);
            const size_t e = ((size_t)l << P.log2_table) + idx;
            atomicAdd(T.ggrad + 2 * e, w * g0);
            atomicAdd(T.ggrad + 2 * e + 1, w * g1);
        }
    }
}

// Adam (Kingma & Ba 2015) with bias correction; also writes the low-precision
// copies the kernels read and clears the gradient.
__global__ void k_nrc_adam_mlp(NrcTrainDev T, float lr1, float lr_scale_b1, float lr_scale_b2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NRC_WEIGHTS || T.sel.rows() == 0)
        return;
    const float g = T.wgrad[i];
    T.wgrad[i] = 0.0f;
    float m = T.wm[i] = T.beta1 * T.wm[i] + (1.0f - T.beta1) * g;
    float v = T.wv[i] = T.beta2 * T.wv[i] + (1.0f - T.beta2) * g * g;
    const float mh = m / lr_scale_b1, vh = v / lr_scale_b2;
    const float p = T.wmaster[i] - lr1 * mh / (sqrtf(vh) + T.eps);
    T.wmaster[i] = p;
    // i indexes the [K][N] layout of its layer; mirror into [N][K]
    int layer = i < NRC_WOFF(NRC_LAYERS - 1) ? i / (NRC_WIDTH * NRC_WIDTH) : NRC_LAYERS - 1;
    const int N = layer == NRC_LAYERS - 1 ? NRC_OUT : NRC_WIDTH;
    const int local = i - NRC_WOFF(layer);
    const int k = local / N, n = local % N;
    const uint16_t b = to_bf16(p);
    T.wk_rw[i] = b;
    T.wt_rw[NRC_WOFF(layer) + n * NRC_WIDTH + k] = b;
}

__global__ void k_nrc_adam_grid(NrcTrainDev T, size_t count, float lr1, float lr_scale_b1, float lr_scale_b2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || T.sel.rows() == 0)
        return;
    const float g = T.ggrad[i];
    T.ggrad[i] = 0.0f;
    float m = T.gm[i] = T.beta1 * T.gm[i] + (1.0f - T.beta1) * g;
    float v = T.gv[i] = T.beta2 * T.gv[i] + (1.0f - T.beta2) * g * g;
    const float mh = m / lr_scale_b1, vh = v / lr_scale_b2;
    const float p = T.gmaster[i] - lr1 * mh / (sqrtf(vh) + T.eps);
    T.gmaster[i] = p;
    T.grid_rw[i] = (_Float16)p;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

__global__ void k_nrc_zero(float* p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0.0f;
}

inline unsigned blocks_for(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

uint32_t xorshift(uint32_t& s)
{
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

uint16_t host_to_bf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

} // namespace

void nrc_set_params(Ctx& ctx, const float* mlp_master, const float* grid_master)
{
    Nrc& C = ctx.nrc;
    if (mlp_master)
    {
        std::vector<uint16_t> wk(NRC_WEIGHTS), wt(NRC_WEIGHTS);
        for (int layer = 0; layer < NRC_LAYERS; layer++)
        {
            const int N = layer == NRC_LAYERS - 1 ? NRC_OUT : NRC_WIDTH;
            for (int k = 0; k < NRC_WIDTH; k++)
                for (int n = 0; n < N; n++)
                {
                    const int i = NRC_WOFF(layer) + k * N + n;
                    wk[i] = host_to_bf16(mlp_master[i]);
                    wt[NRC_WOFF(layer) + n * NRC_WIDTH + k] = wk[i];
                }
        }
        HIP_CHECK(hipMemcpy(C.w_master.ptr, mlp_master, 4 * NRC_WEIGHTS, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(C.wk.ptr, wk.data(), 2 * NRC_WEIGHTS, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(C.wt.ptr, wt.data(), 2 * NRC_WEIGHTS, hipMemcpyHostToDevice));
    }
    if (grid_master)
    {
        std::vector<_Float16> gh(C.grid_count);
        for (size_t i = 0; i < C.grid_count; i++)
            gh[i] = (_Float16)grid_master[i];
        HIP_CHECK(hipMemcpy(C.grid_master.ptr, grid_master, 4 * C.grid_count, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(C.grid_f16.ptr, gh.data(), 2 * C.grid_count, hipMemcpyHostToDevice));
    }
}

void nrc_configure(Ctx& ctx, const mipt_radiance_cache_settings& s)
{
    Nrc& C = ctx.nrc;
    C.settings = s;
    const int L = s.grid_levels;
    const size_t T = (size_t)1 << s.grid_log2_table;
    C.grid_count = (size_t)L * T * 2;
    C.grid_f16.reserve(2 * C.grid_count);
    C.grid_master.reserve(4 * C.grid_count);
    C.grid_m.reserve(4 * C.grid_count);
    C.grid_v.reserve(4 * C.grid_count);
    C.grid_grad.reserve(4 * C.grid_count);
    C.w_master.reserve(4 * NRC_WEIGHTS);
    C.w_m.reserve(4 * NRC_WEIGHTS);
    C.w_v.reserve(4 * NRC_WEIGHTS);
    C.w_grad.reserve(4 * NRC_WEIGHTS);
    C.wt.reserve(2 * NRC_WEIGHTS);
    C.wk.reserve(2 * NRC_WEIGHTS);
    C.loss.reserve(64);
    // Host initialisation: Glorot-uniform weights, grid features U(-1e-4, 1e-4),
    // from a xorshift32 stream seeded with `seed`.
    uint32_t st = s.seed ? s.seed : 1u;
    std::vector<float> w(NRC_WEIGHTS);
    std::vector<uint16_t> wk(NRC_WEIGHTS), wt(NRC_WEIGHTS);
    for (int layer = 0; layer < NRC_LAYERS; layer++)
    {
        const int N = layer == NRC_LAYERS - 1 ? NRC_OUT : NRC_WIDTH;
        const int fan_out = layer == NRC_LAYERS - 1 ? 3 : NRC_WIDTH;
        const float lim = std::sqrt(6.0f / (float)(NRC_WIDTH + fan_out));
        for (int k = 0; k < NRC_WIDTH; k++)
            for (int n = 0; n < N; n++)
            {
                float u = (float)xorshift(st) * (1.0f / 4294967296.0f);
                float x = (2.0f * u - 1.0f) * lim;
                if (layer == NRC_LAYERS - 1 && n >= 3)
                    x = 0.0f;
                const int i = NRC_WOFF(layer) + k * N + n;
                w[i] = x;
                wk[i] = host_to_bf16(x);
                wt[NRC_WOFF(layer) + n * NRC_WIDTH + k] = wk[i];
            }
    }
    std::vector<float> gm(C.grid_count);
    std::vector<_Float16> gh(C.grid_count);
    for (size_t i = 0; i < C.grid_count; i++)
    {
        float u = (float)xorshift(st) * (1.0f / 4294967296.0f);
        gm[i] = (2.0f * u - 1.0f) * 1.0e-4f;
        gh[i] = (_Float16)gm[i];
    }
    HIP_CHECK(hipMemcpy(C.w_master.ptr, w.data(), 4 * NRC_WEIGHTS, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(C.wk.ptr, wk.data(), 2 * NRC_WEIGHTS, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(C.wt.ptr, wt.data(), 2 * NRC_WEIGHTS, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(C.grid_master.ptr, gm.data(), 4 * C.grid_count, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(C.grid_f16.ptr, gh.data(), 2 * C.grid_count, hipMemcpyHostToDevice));
    for (DeviceBuffer* b : {&C.w_m, &C.w_v, &C.w_grad})
        HIP_CHECK(hipMemset(b->ptr, 0, 4 * NRC_WEIGHTS));
    for (DeviceBuffer* b : {&C.grid_m, &C.grid_v, &C.grid_grad})
        HIP_CHECK(hipMemset(b->ptr, 0, 4 * C.grid_count));
    C.step = 0;
    C.configured = true;
}

NrcDev nrc_dev(const Ctx& ctx)
{
    const Nrc& C = ctx.nrc;
    const mipt_radiance_cache_settings& s = C.settings;
    NrcDev P{};
    for (int a = 0; a < 3; a++)
        P.bmin[a] = ctx.scene_bmin[a];
    float ext = 0.0f;
    for (int a = 0; a < 3; a++)
        ext = std::fmax(ext, ctx.scene_bmax[a] - ctx.scene_bmin[a]);
    P.inv_extent = ext > 0.0f ? 1.0f / ext : 1.0f;
    P.levels = s.grid_levels;
    P.log2_table = s.grid_log2_table;
    const double T = std::ldexp(1.0, s.grid_log2_table);
    for (int l = 0; l < s.grid_levels; l++)
    {
        // Instant-NGP resolution schedule: scale_l = base * b^l - 1. Corners
        // reach floor(scale + 0.5) + 1 along an axis, so a dense level holds
        // floor(scale + 0.5) + 2 points per axis.
        const double scale = (double)s.grid_base_resolution * std::pow((double)s.grid_per_level_scale, l) - 1.0;
        const uint32_t res = (uint32_t)std::floor((double)(float)scale + 0.5) + 2u;
        P.level[l].scale = (float)scale;
        P.level[l].res = res;
        P.level[l].dense = (double)res * res * res <= T ? 1u : 0u;
    }
    P.grid_f16 = C.grid_f16.as<_Float16>();
    return P;
}

// Spreads the 10 low bits of v to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Morton key of each valid row's position (10 bits per axis of the grid's
// cube); rows past the count get the largest key and sort last.
__global__ void k_nrc_order_keys(NrcDev P, const float4* __restrict__ queries, const uint32_t* count, uint32_t n,
                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n)
        return;
    vals[row] = row;
    if (row >= *count)
    {
        keys[row] = 0xffffffffu;
        return;
    }
    const float4 a = queries[4 * (size_t)row];
    const float q[3] = {a.x, a.y, a.z};
    float pos[3];
    normalized_position(P, q, pos);
    uint32_t c[3];
    for (int k = 0; k < 3; k++)
        c[k] = (uint32_t)fminf(pos[k] * 1024.0f, 1023.0f);
    keys[row] = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
}

__global__ void k_nrc_order_gather(const float4* __restrict__ queries, const uint32_t* __restrict__ pix,
                                   const uint32_t* __restrict__ perm, const uint32_t* count, float4* __restrict__ q_out,
                                   uint32_t* __restrict__ pix_out)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= *count)
        return;
    const uint32_t src = perm[row];
    for (int k = 0; k < 4; k++)
        q_out[4 * (size_t)row + k] = queries[4 * (size_t)src + k];
    pix_out[row] = pix[src];
}

void nrc_spatial_order(Ctx& ctx, const float4* d_queries, const uint32_t* d_pix, const uint32_t* d_count, uint32_t n,
                       float4* q_out, uint32_t* pix_out, hipStream_t s)
{
    Nrc& C = ctx.nrc;
    if (n == 0)
        return;
    C.sort_keys.reserve(4 * (size_t)n);
    C.sort_keys_out.reserve(4 * (size_t)n);
    C.sort_vals.reserve(4 * (size_t)n);
    C.sort_vals_out.reserve(4 * (size_t)n);
    const NrcDev P = nrc_dev(ctx);
    hipLaunchKernelGGL(k_nrc_order_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, d_queries, d_count, n,
                       C.sort_keys.as<uint32_t>(), C.sort_vals.as<uint32_t>());
    size_t tmp_bytes = 0;
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, C.sort_keys.as<uint32_t>(), C.sort_keys_out.as<uint32_t>(),
                                        C.sort_vals.as<uint32_t>(), C.sort_vals_out.as<uint32_t>(), (size_t)n, 0u, 32u,
                                        s));
    C.sort_tmp.reserve(tmp_bytes > 0 ? tmp_bytes : 1);
    HIP_CHECK(rocprim::radix_sort_pairs(C.sort_tmp.ptr, tmp_bytes, C.sort_keys.as<uint32_t>(),
                                        C.sort_keys_out.as<uint32_t>(), C.sort_vals.as<uint32_t>(),
                                        C.sort_vals_out.as<uint32_t>(), (size_t)n, 0u, 32u, s));
    hipLaunchKernelGGL(k_nrc_order_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, d_queries, d_pix,
                       C.sort_vals_out.as<uint32_t>(), d_count, q_out, pix_out);
    HIP_CHECK(hipGetLastError());
}

void nrc_infer(Ctx& ctx, const float* d_queries, const uint32_t* d_count, uint32_t n, float* d_out, hipStream_t s)
{
    Nrc& C = ctx.nrc;
    if (!C.configured)
        fail(MIPT_ERR_INVALID_ARGUMENT, "radiance cache not configured: call mipt_set_radiance_cache first");
    if (n == 0)
        return;
    // rows are padded to whole 16-row tiles (zero rows past the count)
    const uint32_t n16 = (n + 15) & ~15u;
    C.x0.reserve(2 * (size_t)n16 * NRC_WIDTH);
    C.count.reserve(4);
    if (!d_count)
    {
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, C.count.as<uint32_t>(), n);
        d_count = C.count.as<uint32_t>();
    }
    const NrcDev P = nrc_dev(ctx);
    encode_launch(P, d_queries, RowSel{d_count, 1u, 0u, n}, n16, 16u, C.x0.as<uint16_t>(), s);
    const unsigned grid = (unsigned)std::min<size_t>(blocks_for((n + 15) / 16, 4), (size_t)ctx.cu_count * 4);
    hipLaunchKernelGGL(k_nrc_forward, dim3(grid), dim3(256), 0, s, C.wt.as<uint16_t>(), C.x0.as<uint16_t>(),
                       d_count, n, reinterpret_cast<float4*>(d_out));
    HIP_CHECK(hipGetLastError());
}

void nrc_encode(Ctx& ctx, const float* d_queries, uint32_t n, uint16_t* d_out, hipStream_t s)
{
    Nrc& C = ctx.nrc;
    if (!C.configured)
        fail(MIPT_ERR_INVALID_ARGUMENT, "radiance cache not configured: call mipt_set_radiance_cache first");
    if (n == 0)
        return;
    C.count.reserve(4);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, C.count.as<uint32_t>(), n);
    encode_launch(nrc_dev(ctx), d_queries, RowSel{C.count.as<uint32_t>(), 1u, 0u, n}, n, 1u, d_out, s);
    HIP_CHECK(hipGetLastError());
}

void nrc_gradients(Ctx& ctx, const float* d_queries, const float* d_targets, const uint32_t* d_count, uint32_t n,
                   uint32_t rec_stride, uint32_t rec_offset, bool clear_loss, hipStream_t s)
{
    Nrc& C = ctx.nrc;
    if (!C.configured)
        fail(MIPT_ERR_INVALID_ARGUMENT, "radiance cache not configured: call mipt_set_radiance_cache first");
    // rows of this batch at most (records below n)
    const uint32_t avail = n > rec_offset ? (n - rec_offset + rec_stride - 1) / rec_stride : 0u;
    const uint32_t batch = std::min<uint32_t>(avail, (uint32_t)C.settings.batch_size);
    C.last_sel_count = nullptr;
    if (batch == 0)
        return;
    const uint32_t stride = (batch + 31) & ~31u;
    C.x0.reserve(2 * (size_t)stride * NRC_WIDTH);
    C.xcm.reserve(2 * (size_t)NRC_LAYERS * NRC_WIDTH * stride);
    C.dycm.reserve(2 * (size_t)NRC_LAYERS * NRC_WIDTH * stride);
    C.dx0.reserve(4 * (size_t)stride * NRC_GRID_MAX);
    C.count.reserve(4);
    const uint32_t* cnt = d_count;
    if (!cnt)
    {
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, C.count.as<uint32_t>(), n);
        cnt = C.count.as<uint32_t>();
    }
    const RowSel sel{cnt, rec_stride, rec_offset, batch};
    C.last_sel_count = cnt;
    C.last_sel_stride = rec_stride;
    C.last_sel_offset = rec_offset;
    C.last_sel_limit = batch;
    const NrcDev P = nrc_dev(ctx);
    // encode the batch; rows past its count up to a multiple of 32 are zero
    encode_launch(P, d_queries, sel, stride, 32u, C.x0.as<uint16_t>(), s);
    if (clear_loss)
        hipLaunchKernelGGL(k_nrc_zero, dim3(1), dim3(64), 0, s, C.loss.as<float>(), (size_t)1);
    NrcTrainDev T{};
    T.wt = C.wt.as<uint16_t>();
    T.wk = C.wk.as<uint16_t>();
    T.x0 = C.x0.as<uint16_t>();
    T.targets = d_targets;
    T.sel = sel;
    T.batch_stride = stride;
    T.levels = C.settings.grid_levels;
    T.xcm = C.xcm.as<uint16_t>();
    T.dycm = C.dycm.as<uint16_t>();
    T.dx0 = C.dx0.as<float>();
    T.loss_sum = C.loss.as<float>();
    T.wgrad = C.w_grad.as<float>();
    T.ggrad = C.grid_grad.as<float>();
    const unsigned grid = (unsigned)std::min<size_t>(blocks_for(stride / 16, 4), (size_t)ctx.cu_count);
    hipLaunchKernelGGL(k_nrc_train_fwd_bwd, dim3(grid), dim3(256), 0, s, T);
    const uint32_t chunk = 1024;
    hipLaunchKernelGGL(k_nrc_weight_grad, dim3(blocks_for(stride, chunk), NRC_LAYERS), dim3(256), 0, s, T, chunk);
    hipLaunchKernelGGL(k_nrc_grid_grad, dim3(NRC_XCDS * blocks_for(stride, 128)), dim3(128), 0, s, P, T,
                       reinterpret_cast<const float4*>(d_queries));
    HIP_CHECK(hipGetLastError());
}

void nrc_train_pass(Ctx& ctx, hipStream_t s)
{
    Nrc& C = ctx.nrc;
    if (!C.configured || C.train_capacity == 0)
        fail(MIPT_ERR_INVALID_ARGUMENT,
             "RADIANCE_CACHE_TRAIN: no training records; run a path-trace pass with radiance_cache on first");
    const uint32_t cap = C.train_capacity;
    const uint32_t batch = (uint32_t)C.settings.batch_size;
    const uint32_t steps = (cap + batch - 1) / batch;
    const uint32_t* count = C.counters.as<uint32_t>() + 1;
    for (uint32_t i = 0; i < steps; i++)
    {
        nrc_gradients(ctx, C.train_queries.as<float>(), C.train_targets.as<float>(), count, cap, steps, i, i == 0,
                      s);
        nrc_apply_adam(ctx, s);
    }
}

void nrc_apply_adam(Ctx& ctx, hipStream_t s)
{
    Nrc& C = ctx.nrc;
    if (!C.last_sel_count)
        return; // no batch
    C.step++;
    const mipt_radiance_cache_settings& st = C.settings;
    const float b1t = 1.0f - (float)std::pow((double)st.adam_beta1, (double)C.step);
    const float b2t = 1.0f - (float)std::pow((double)st.adam_beta2, (double)C.step);
    NrcTrainDev T{};
    T.beta1 = st.adam_beta1;
    T.beta2 = st.adam_beta2;
    T.eps = st.adam_eps;
    T.wgrad = C.w_grad.as<float>();
    T.wm = C.w_m.as<float>();
    T.wv = C.w_v.as<float>();
    T.wmaster = C.w_master.as<float>();
    T.wk_rw = C.wk.as<uint16_t>();
    T.wt_rw = C.wt.as<uint16_t>();
    T.ggrad = C.grid_grad.as<float>();
    T.gm = C.grid_m.as<float>();
    T.gv = C.grid_v.as<float>();
    T.gmaster = C.grid_master.as<float>();
    T.grid_rw = C.grid_f16.as<_Float16>();
    T.sel = RowSel{C.last_sel_count, C.last_sel_stride, C.last_sel_offset, C.last_sel_limit};
    hipLaunchKernelGGL(k_nrc_adam_mlp, dim3(blocks_for(NRC_WEIGHTS, 256)), dim3(256), 0, s, T, st.learning_rate, b1t,
                       b2t);
    hipLaunchKernelGGL(k_nrc_adam_grid, dim3(blocks_for(C.grid_count, 256)), dim3(256), 0, s, T, C.grid_count,
                       st.learning_rate, b1t, b2t);
    HIP_CHECK(hipGetLastError());
}

} // namespace mipt
