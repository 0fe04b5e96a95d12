// Device records and launchers of the neural radiance cache.
#pragma once

#include <hip/hip_runtime.h>

#include "../core/nrc.h"

namespace btrc {

// network state: FP32 master weights and Adam moments, BF16 copies as
// [out][in] and (per layer) [in][out]
struct NrcWeights
{
    float *w;
    float *m;
    float *v;
    uint16_t *w_bf16;
    uint16_t *wt_bf16;
    float *ema;          // moving average of w (FP32) and its BF16 copy
    uint16_t *ema_bf16;
};

// a cache query made at the vertex that ends a path's prefix
// the training-path id of a path that is not one
constexpr uint32_t NO_TRAIN = 0xffffffffu;

struct __attribute__((aligned(16))) NrcQuery
{
    NrcVertex v;
    float beta[3];   // throughput the vertex was reached with
    uint32_t pixel;
    uint32_t train;  // training path whose tail this is (its sum gets the answer too), or NO_TRAIN
    uint32_t pad[3]; // 96 bytes: six 16-byte loads
};

// a training record: a vertex of an extended (training) path
struct NrcRecord
{
    NrcVertex v;
    float beta[3];      // throughput the vertex was reached with
    float gathered[3];  // the path's radiance sum when the vertex was reached (its emission included)
    uint32_t path;      // training path id: row of the radiance sums
    uint32_t valid;
};

// what the shading stages need in cache mode
struct NrcView
{
    int32_t query_depth;     // vertex depth answered by the cache (-1: no queries)
    int32_t slots_per_path;  // record slots per training path (max depth + 1, or the training depth)
    int32_t train_depth;     // training paths end here with a query (-1: traced to the end)
    NrcQuery *queries;
    int32_t *query_count;
    NrcRecord *rec;
    float *acc;              // radiance sum per training path (4 floats)
};

// A training step's samples. Direct mode (index null): samples
// [first, first + batch), sample i reading record first + i, or
// nrc_record_slot(first + i) with permute. Compacted mode (index set): the
// frame's valid records, index[0, *total), split into `steps` equal parts
// (multiples of 32, at most `batch`); this step takes part `step`. Either way `batch` is the
// row stride of the step's operand buffers (a multiple of 64, at least the
// step's sample count).
struct NrcBatch
{
    const NrcRecord *rec;
    int64_t first;
    int64_t batch;
    const float *acc;
    int32_t *valid_count;    // set by a training step
    float loss_scale;
    int32_t permute;
    const int32_t *index;    // compacted mode: record slots of the valid records
    const int32_t *total;    // compacted mode: how many
    int32_t step, steps;
};

// first sample and sample count of a step
__host__ __device__ inline void nrc_batch_range(const NrcBatch &b, int64_t *first, int64_t *count)
{
    if(!b.index)
    {
        *first = b.first;
        *count = b.batch;
        return;
    }
    const int64_t total = *b.total;
    int64_t per = ((total + b.steps - 1) / b.steps + 31) / 32 * 32;
    per = per < b.batch ? per : b.batch; // at most one stride of samples per step
    const int64_t f = (int64_t)b.step * per;
    *first = f;
    *count = total - f < 0 ? 0 : (total - f < per ? total - f : per);
}

// A frame's records sit at (training path, depth); training batches read
// them through this bijection of [0, 2^16) so that every batch mixes paths
// and depths.
__host__ __device__ inline int64_t nrc_record_slot(int64_t i)
{
    return (int64_t)(((uint32_t)i * 40503u) & 0xffffu);
}

constexpr int NRC_KSPLIT = 32;  // sample slices of the weight-gradient products

struct NrcScratch
{
    uint16_t *xt;            // [layer][64][batch] layer inputs
    uint16_t *dzt;           // [layer][64][batch] layer output gradients
    double *loss_partial;    // one per wavefront
    int32_t *count_partial;  // valid samples, one per wavefront
    float *grad_partial;     // [NRC_KSPLIT][NRC_PARAMS]
    float *step_scale;       // summed gradient -> gradient of the mean loss; 0: skip the update
    double *g2_partial;      // squared gradient norm, one per reduce workgroup
};

constexpr int NRC_REDUCE_BLOCKS = (NRC_PARAMS + 255) / 256;

struct NrcAdam
{
    float lr, beta1, beta2, eps;
    float inv_bias1, inv_bias2; // 1 / (1 - beta^t)
    float loss_scale;
    float ema_decay;            // 0: no moving average
};

struct NrcStepStats
{
    double loss;       // mean relative L2 over the batch's valid samples and channels
    double grad_norm;  // L2 norm of the unscaled gradient
    int32_t samples;
    int32_t skipped;   // 1: non-finite gradient or empty batch, no update
};

// queries [0, *count) with *count <= max_queries (which only sizes the grid),
// no two of them for the same pixel
void launch_nrc_infer(hipStream_t s, int cu_count, int64_t max_queries, const NrcWeights &net, const NrcBox &box,
                      const NrcQuery *q, const int32_t *count, float *film, double *film64, float *acc);
void launch_nrc_eval(hipStream_t s, const NrcWeights &net, const NrcBox &box, const NrcVertex *v, int32_t n,
                     float *out);
void launch_nrc_train_step(hipStream_t s, const NrcWeights &net, const NrcBox &box, const NrcBatch &b,
                           const NrcScratch &sc, const NrcAdam &opt, float *grad, NrcStepStats *stats);
void launch_nrc_gradients(hipStream_t s, const NrcWeights &net, const NrcBox &box, const NrcBatch &b,
                          const NrcScratch &sc, float *grad);
void launch_nrc_clear(hipStream_t s, NrcRecord *rec, int64_t nrec, float *acc, int64_t npaths);
// index = the valid record slots among rec[0, 2^16) in the order of
// nrc_record_slot(0, 1, ...), *total = how many (deterministic); mask: 1024
// words of scratch
void launch_nrc_compact(hipStream_t s, const NrcRecord *rec, uint64_t *mask, int32_t *index, int32_t *total);
void launch_nrc_publish(hipStream_t s, const NrcWeights &net);

} // namespace btrc
