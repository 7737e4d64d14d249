// nrc.h -- learned radiance cache: constants, device views and entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace mipt {

constexpr int NRC_WIDTH = 64;     // input and hidden width
constexpr int NRC_OUT = 16;       // output block (RGB in the first 3 columns)
constexpr int NRC_LAYERS = 6;     // weight matrices: 5 x [64][64] + [64][16]
constexpr int NRC_BLOB = 4;       // one-blob bins per scalar
constexpr int NRC_GRID_MAX = 32;  // 2 features x at most 16 levels
constexpr int NRC_WEIGHTS = 5 * NRC_WIDTH * NRC_WIDTH + NRC_WIDTH * NRC_OUT;
#define NRC_WOFF(layer) ((layer) * NRC_WIDTH * NRC_WIDTH)

// Rows of one batch: records offset, offset + stride, ... below *count, at
// most limit of them (row j is record j * stride + offset).
struct RowSel {
    const uint32_t* count;
    uint32_t stride;
    uint32_t offset;
    uint32_t limit;
#ifdef __HIPCC__
    __device__ uint32_t rows() const
    {
        const uint32_t c = *count;
        const uint32_t r = c > offset ? (c - offset + stride - 1) / stride : 0u;
        return r < limit ? r : limit;
    }
    __device__ size_t record(uint32_t row) const { return (size_t)row * stride + offset; }
#endif
};

struct NrcGridLevel {
    float scale;
    uint32_t res;
    uint32_t dense;
};

struct NrcDev {
    float bmin[3];
    float inv_extent;
    int32_t levels;
    int32_t log2_table;
    NrcGridLevel level[16];
    const _Float16* grid_f16;
};

struct NrcTrainDev {
    const uint16_t* wt;   // W^T per layer, [N][64] bf16
    const uint16_t* wk;   // W per layer, [64][N] bf16
    const uint16_t* x0;   // encoded rows, [rows][64] bf16
    const float* targets; // [records][4]
    RowSel sel;           // the batch's rows among the records
    uint32_t batch_stride;
    int32_t levels;
    uint16_t* xcm;        // layer inputs, column-major [layer][64][stride]
    uint16_t* dycm;       // layer output gradients, column-major [layer][64][stride]
    float* dx0;           // grid-feature gradients [stride][32]
    float* loss_sum;
    float* wgrad;         // FP32 [K][N] per layer
    float* ggrad;         // FP32 grid gradient
    // Adam
    float beta1, beta2, eps;
    float* wm;
    float* wv;
    float* wmaster;
    uint16_t* wk_rw;
    uint16_t* wt_rw;
    float* gm;
    float* gv;
    float* gmaster;
    _Float16* grid_rw;
};

void nrc_configure(Ctx& ctx, const mipt_radiance_cache_settings& s);
NrcDev nrc_dev(const Ctx& ctx);
// Encodes and evaluates rows [0, min(*d_count, n)) (or [0, n) with d_count
// null) of d_queries (16 floats per row) into d_out (4 floats per row).
void nrc_infer(Ctx& ctx, const float* d_queries, const uint32_t* d_count, uint32_t n, float* d_out,
               hipStream_t stream);
// Copies the query records [0, *d_count) (16 floats each) and their pixel
// indices into q_out / pix_out sorted by the Morton code of the position in
// the grid's cube: rows next to each other then read the same hash-grid
// entries. n bounds *d_count.
void nrc_spatial_order(Ctx& ctx, const float4* d_queries, const uint32_t* d_pix, const uint32_t* d_count,
                       uint32_t n, float4* q_out, uint32_t* pix_out, hipStream_t stream);
// Writes the encoded input rows of [0, n) (64 bf16 per row) to d_out.
void nrc_encode(Ctx& ctx, const float* d_queries, uint32_t n, uint16_t* d_out, hipStream_t stream);
// Adds the loss gradients of one batch (records offset, offset + stride, ...
// below *d_count, or below n with d_count null; at most batch_size of them)
// to the FP32 gradient buffers and its loss sum to ctx.nrc.loss[0] (cleared
// first when clear_loss).
void nrc_gradients(Ctx& ctx, const float* d_queries, const float* d_targets, const uint32_t* d_count, uint32_t n,
                   uint32_t stride, uint32_t offset, bool clear_loss, hipStream_t stream);
// One Adam step from the accumulated gradients (which it clears); skipped on
// the device when the batch nrc_gradients last saw was empty.
void nrc_apply_adam(Ctx& ctx, hipStream_t stream);
// RADIANCE_CACHE_TRAIN: Adam steps over the training records of the last path
// pass, in batches of at most batch_size that interleave the records (step i
// takes records i, i + steps, ...), so each step sees the whole frame.
void nrc_train_pass(Ctx& ctx, hipStream_t stream);
// Uploads FP32 masters (either may be null) and refreshes the BF16 / FP16
// copies the kernels read; the Adam moments are kept.
void nrc_set_params(Ctx& ctx, const float* mlp_master, const float* grid_master);

} // namespace mipt
