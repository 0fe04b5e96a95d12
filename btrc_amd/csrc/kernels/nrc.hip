// Neural radiance cache kernels for gfx950: inference, fused forward +
// backward of a training batch, the weight-gradient products and Adam.
//
// All matrix products run on MFMA (v_mfma_f32_32x32x16_bf16): BF16 operands,
// FP32 accumulation. One wavefront owns 32 samples; a layer is
// Y^T = W X^T with M = output units, N = the 32 samples, K = input units, so
// the 16 results a lane holds belong to one sample (lane % 32). Weights sit
// in LDS as [out][in] (A operand) and, for the backward pass, as [in][out],
// rows padded to 72 halves.
//
// Lane layout of the 32x32x16 MFMA (checked on the device before use):
//   A (32 x 16): lane l holds row l % 32, k = 8 (l / 32) + 0..7
//   B (16 x 32): lane l holds column l % 32, k = 8 (l / 32) + 0..7
//   D (32 x 32): element i of lane l is row 8 (i / 4) + 4 (l / 32) + i % 4,
//                column l % 32
//
// Activations and gradients stay in registers from layer to layer. The A rows
// of a 32-row output tile read weight rows in the order swap23(r) (bits 2
// and 3 of r exchanged); D element i of lane l then holds unit
// 16 (i / 8) + 8 (l / 32) + i % 8 of the tile, which is exactly the next
// layer's B operand for k slice 2 tile + i / 8. The weights keep their
// natural layouts; only the order in which rows are read changes. The
// backward pass reads rows of the [in][out] copy the same way, so each
// input gradient is the next product's B operand too.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../core/nrc.h"
#include "nrc_kernels.h"

namespace btrc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int WAVE = 64;
constexpr int TILE = 32; // samples per wavefront

// FP32 -> BF16, round to nearest even (NaN stays a quiet NaN)
__device__ __forceinline__ uint16_t to_bf16(float f)
{
    const uint32_t u = __float_as_uint(f);
    if((u & 0x7fffffffu) > 0x7f800000u)
        return 0x7fc0;
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x16 mfma(u16x8 a, u16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                    0, 0, 0);
}

__device__ __forceinline__ u16x8 lds8(const uint16_t *p) { return *reinterpret_cast<const u16x8 *>(p); }

// ---- register-resident forward pass (inference)

constexpr int WS = NRC_IN + 8;                  // LDS row stride of the weights, halves (no bank conflicts)
constexpr int NRC_ROWS = NRC_PARAMS / NRC_IN;   // 352 weight rows

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// two FP32 values to packed BF16, round to nearest even: one
// v_cvt_pk_bf16_f32 from the compiler's own lowering (not inline asm, which
// would read MFMA results without the compiler's hazard waits)
__device__ __forceinline__ uint32_t pk_bf16(float lo, float hi)
{
    const f32x2 v = { lo, hi };
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// ReLU of two packed BF16 values: negative values (sign bit set) are negative
// as int16, so a signed max with 0 clears them; ReLU and rounding commute
__device__ __forceinline__ uint32_t relu2(uint32_t x)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, x), (s16x2)0));
}

__device__ __forceinline__ u16x8 words8(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = { a, b, c, d };
    return __builtin_bit_cast(u16x8, w);
}

__device__ __forceinline__ int swap23(int r) { return (r & ~12) | ((r & 4) << 1) | ((r & 8) >> 1); }

// the B operand of a 64-wide layer input: k slices 0..3, this lane's 8 values each
struct Frag
{
    u16x8 s[4];
};

// Copies N 16-byte chunks from global memory into LDS with THREADS threads,
// chunk i to halves s + dst(i): each thread issues BATCH loads before its
// LDS writes (one memory latency per BATCH chunks, not per chunk)
template <int N, int THREADS, int BATCH, class Dst>
__device__ __forceinline__ void stage_lds(const uint4 *src, uint16_t *s, Dst dst)
{
    static_assert(N % (THREADS * BATCH) == 0, "whole batches only");
#pragma unroll
    for(int i0 = 0; i0 < N; i0 += THREADS * BATCH)
    {
        uint4 v[BATCH];
#pragma unroll
        for(int k = 0; k < BATCH; ++k)
            v[k] = src[i0 + k * THREADS + (int)threadIdx.x];
#pragma unroll
        for(int k = 0; k < BATCH; ++k)
            *reinterpret_cast<uint4 *>(s + dst(i0 + k * THREADS + (int)threadIdx.x)) = v[k];
    }
}

constexpr int WEIGHT_CHUNKS = NRC_PARAMS / 8; // 16-byte chunks of one weight copy

// weights of all layers into LDS as [out][in], rows padded to WS halves, by
// a workgroup of THREADS threads
template <int THREADS>
__device__ __forceinline__ void load_weights_padded(const uint16_t *w, uint16_t *s_w)
{
    stage_lds<WEIGHT_CHUNKS, THREADS, 11>(
        reinterpret_cast<const uint4 *>(w), s_w,
        [](int i) { return (i / (NRC_IN / 8)) * WS + 8 * (i % (NRC_IN / 8)); });
}

// the encoded inputs of this lane's sample as the first layer's B operand:
// lane half h holds inputs 16 s + 8 h + 0..7 of slice s, which is what
// nrc_encode_half computes for its half (an invalid sample encodes a zero
// vertex: finite inputs whose gradients are zero)
__device__ __forceinline__ void encode_frag(const NrcVertex &v, const NrcBox &box, int h, Frag &x)
{
    float in[NRC_IN / 2];
    nrc_encode_half(v, box, h, in);
#pragma unroll
    for(int s = 0; s < 4; ++s)
        x.s[s] = words8(pk_bf16(in[8 * s + 0], in[8 * s + 1]), pk_bf16(in[8 * s + 2], in[8 * s + 3]),
                        pk_bf16(in[8 * s + 4], in[8 * s + 5]), pk_bf16(in[8 * s + 6], in[8 * s + 7]));
}

// 8 accumulator values from o on, ReLU, as one B fragment
__device__ __forceinline__ u16x8 relu_frag(const f32x16 &a, int o)
{
    return words8(relu2(pk_bf16(a[o + 0], a[o + 1])), relu2(pk_bf16(a[o + 2], a[o + 3])),
                  relu2(pk_bf16(a[o + 4], a[o + 5])), relu2(pk_bf16(a[o + 6], a[o + 7])));
}

// hidden layer l (64 x 64, ReLU) for T tiles of 32 samples, in registers:
// each weight fragment read from LDS feeds T products
template <int T>
__device__ __forceinline__ void hidden_layer_n(const uint16_t *s_w, int l, int col, int h, Frag *x)
{
    // keeps the scheduler from hoisting every layer's weight loads to the
    // top (their registers would lower occupancy)
    __builtin_amdgcn_sched_barrier(0);
    const uint16_t *r0 = s_w + (l * NRC_WIDTH + swap23(col)) * WS + 8 * h;
    const uint16_t *r1 = r0 + 32 * WS;
    f32x16 acc0[T], acc1[T];
#pragma unroll
    for(int t = 0; t < T; ++t)
        acc0[t] = acc1[t] = f32x16{};
#pragma unroll
    for(int s = 0; s < 4; ++s)
    {
        const u16x8 a0 = lds8(r0 + 16 * s), a1 = lds8(r1 + 16 * s);
#pragma unroll
        for(int t = 0; t < T; ++t)
        {
            acc0[t] = mfma(a0, x[t].s[s], acc0[t]);
            acc1[t] = mfma(a1, x[t].s[s], acc1[t]);
        }
    }
#pragma unroll
    for(int t = 0; t < T; ++t)
    {
        x[t].s[0] = relu_frag(acc0[t], 0);
        x[t].s[1] = relu_frag(acc0[t], 8);
        x[t].s[2] = relu_frag(acc1[t], 0);
        x[t].s[3] = relu_frag(acc1[t], 8);
    }
}

__device__ __forceinline__ void hidden_layer(const uint16_t *s_w, int l, int col, int h, Frag &x)
{
    hidden_layer_n<1>(s_w, l, col, h, &x);
}

// the whole network for T tiles: this lane's sample of tile t is v[t];
// y[t][0..2] valid on lanes 0..31
template <int T>
__device__ __forceinline__ void forward_regs_n(const uint16_t *s_w, const NrcVertex *v, const NrcBox &box,
                                               float (*y)[3])
{
    const int lane = threadIdx.x % WAVE, col = lane % 32, h = lane / 32;
    Frag x[T];
#pragma unroll
    for(int t = 0; t < T; ++t)
        encode_frag(v[t], box, h, x[t]);
#pragma unroll
    for(int l = 0; l < NRC_LAYERS - 1; ++l)
        hidden_layer_n<T>(s_w, l, col, h, x);
    // output layer: natural row order, rows 3..31 are zero
    const uint16_t *r = s_w + ((NRC_LAYERS - 1) * NRC_WIDTH + col) * WS + 8 * h;
    f32x16 acc[T];
#pragma unroll
    for(int t = 0; t < T; ++t)
        acc[t] = f32x16{};
#pragma unroll
    for(int s = 0; s < 4; ++s)
    {
        const u16x8 a = lds8(r + 16 * s);
#pragma unroll
        for(int t = 0; t < T; ++t)
            acc[t] = mfma(a, x[t].s[s], acc[t]);
    }
#pragma unroll
    for(int t = 0; t < T; ++t)
    {
        y[t][0] = acc[t][0];
        y[t][1] = acc[t][1];
        y[t][2] = acc[t][2];
    }
}

__device__ __forceinline__ void forward_regs(const uint16_t *s_w, const NrcVertex &v, const NrcBox &box, float *y)
{
    forward_regs_n<1>(s_w, &v, box, reinterpret_cast<float(*)[3]>(y));
}

// a query's answer: y = max(net(x), 0); the film gets beta * y at the pixel,
// and a training path's radiance sum too when the query is its tail. No two
// queries of a launch share a pixel (the cache renders one path per pixel,
// and a path makes at most one query), so both are plain read-modify-writes.
__device__ __forceinline__ void emit_answer(const NrcQuery &qi, const float *y, float *film, double *film64,
                                            float *acc)
{
    const float r[3] = { qi.beta[0] * fmaxf(y[0], 0.0f), qi.beta[1] * fmaxf(y[1], 0.0f),
                         qi.beta[2] * fmaxf(y[2], 0.0f) };
    if(!(isfinite(r[0]) && isfinite(r[1]) && isfinite(r[2])))
        return;
    float4 *p = reinterpret_cast<float4 *>(film) + qi.pixel;
    float4 f = *p;
    f.x += r[0];
    f.y += r[1];
    f.z += r[2];
    *p = f;
    if(film64)
    {
        double *d = film64 + 3 * (size_t)qi.pixel;
        d[0] += (double)r[0];
        d[1] += (double)r[1];
        d[2] += (double)r[2];
    }
    if(qi.train != NO_TRAIN)
    {
        float *a = acc + 4 * (size_t)qi.train;
        a[0] = a[0] + r[0];
        a[1] = a[1] + r[1];
        a[2] = a[2] + r[2];
    }
}

} // namespace

// ---------------------------------------------------------------- inference

// inference launch shape: one tile of 32 queries per wavefront at a time
// (two tiles sharing each weight fragment measured no faster: 170 VGPRs
// halve the resident waves), four wavefronts per workgroup, three workgroups
// per CU (three LDS copies of the weights fit)
constexpr int INFER_TILES = 1;
constexpr int INFER_THREADS = 256;
constexpr int INFER_BLOCKS_PER_CU = 3;

// queries [0, n): y = max(net(x), 0); the film gets beta * y at the pixel,
// and a training path's radiance sum too when the query is its tail (one
// query per training path, after every other addition to its sum).
// Persistent: each wavefront takes INFER_TILES tiles of 32 queries at a
// time; activations stay in registers, the weights in LDS.
__global__ void __launch_bounds__(INFER_THREADS) k_nrc_infer(NrcWeights net, NrcBox box, const NrcQuery *q,
                                                             const int32_t *count, float *film, double *film64,
                                                             float *acc)
{
    constexpr int T = INFER_TILES, WAVES = INFER_THREADS / WAVE, SPAN = T * TILE;
    __shared__ __attribute__((aligned(16))) uint16_t s_w[NRC_ROWS * WS];
    load_weights_padded<INFER_THREADS>(net.w_bf16, s_w);
    __syncthreads();
    const int n = *count;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, col = lane % 32, half = lane / 32;
    const int stride = gridDim.x * WAVES * SPAN;
    // the next span's queries are loaded while the current ones are computed
    int base = (blockIdx.x * WAVES + wave) * SPAN;
    NrcQuery next[T];
#pragma unroll
    for(int t = 0; t < T; ++t)
    {
        next[t] = NrcQuery{};
        if(base + t * TILE + col < n)
            next[t] = q[base + t * TILE + col];
    }
    for(; base < n; base += stride)
    {
        NrcQuery qi[T];
        NrcVertex v[T];
#pragma unroll
        for(int t = 0; t < T; ++t)
        {
            qi[t] = next[t];
            v[t] = qi[t].v;
            if(base + stride + t * TILE + col < n)
                next[t] = q[base + stride + t * TILE + col];
        }
        float y[T][3];
        forward_regs_n<T>(s_w, v, box, y);
#pragma unroll
        for(int t = 0; t < T; ++t)
            if(half == 0 && base + t * TILE + col < n)
                emit_answer(qi[t], y[t], film, film64, acc);
    }
}

// raw network outputs for test hooks: out[3 i + c], the inference kernel's arithmetic
__global__ void __launch_bounds__(256) k_nrc_eval(NrcWeights net, NrcBox box, const NrcVertex *v, int32_t n,
                                                  float *out)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_w[NRC_ROWS * WS];
    load_weights_padded<256>(net.w_bf16, s_w);
    __syncthreads();
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, col = lane % 32, half = lane / 32;
    const int i = (blockIdx.x * 4 + wave) * TILE + col;
    const bool valid = i < n;
    NrcVertex vi{};
    if(valid)
        vi = v[i];
    float y[3];
    forward_regs(s_w, vi, box, y);
    if(half == 0 && valid)
    {
        out[3 * (size_t)i + 0] = y[0];
        out[3 * (size_t)i + 1] = y[1];
        out[3 * (size_t)i + 2] = y[2];
    }
}

// ---------------------------------------------------------------- training

// Fused forward and backward of one batch: samples [0, batch) of `rec`
// (invalid records contribute nothing). Writes every layer's input as
// [unit][sample] (xt) and every layer's output gradient as [unit][sample]
// (dzt), both BF16, for k_nrc_wgrad; and one FP64 loss partial per
// wavefront. The output gradient is d loss / d y times loss_scale, where the
// loss is the mean relative L2 over the valid samples' three channels.
// Activations and gradients stay in registers (the swap23 row order makes
// each product's result the next product's B operand, forward and
// backward); the weights sit in LDS as [out][in] and [in][out].
__global__ void __launch_bounds__(128) k_nrc_train_fb(NrcWeights net, NrcBox box, NrcBatch b, NrcScratch sc)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_w[NRC_ROWS * WS];
    __shared__ __attribute__((aligned(16))) uint16_t s_wt[NRC_LAYERS * NRC_IN * WS];
    int64_t first, count;
    nrc_batch_range(b, &first, &count);
    if((int64_t)blockIdx.x * 2 * TILE >= count)
    {
        // the whole workgroup lies beyond the step's samples: no weights to stage
        if(threadIdx.x % WAVE == 0)
        {
            sc.loss_partial[blockIdx.x * 2 + threadIdx.x / WAVE] = 0.0;
            sc.count_partial[blockIdx.x * 2 + threadIdx.x / WAVE] = 0;
        }
        return;
    }
    load_weights_padded<128>(net.w_bf16, s_w);
    // [in][out] copies, per layer 64 rows of `rows` values (64, or the
    // output layer's NRC_OUT_ROWS), padded to WS; the source is contiguous
    stage_lds<WEIGHT_CHUNKS, 128, 11>(reinterpret_cast<const uint4 *>(net.wt_bf16), s_wt, [](int i) {
        constexpr int HIDDEN = (NRC_LAYERS - 1) * NRC_IN * (NRC_WIDTH / 8); // chunks of the hidden layers
        constexpr int OUT_CHUNKS = NRC_OUT_ROWS / 8;
        if(i < HIDDEN)
            return (i / (NRC_WIDTH / 8)) * WS + 8 * (i % (NRC_WIDTH / 8));
        const int j = i - HIDDEN;
        return ((NRC_LAYERS - 1) * NRC_IN + j / OUT_CHUNKS) * WS + 8 * (j % OUT_CHUNKS);
    });
    __syncthreads();

    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, col = lane % 32, h = lane / 32;
    const int64_t S = b.batch;
    const int64_t sample0 = (int64_t)(blockIdx.x * 2 + wave) * TILE;
    if(sample0 >= count)
    {
        // beyond the step's samples (k_nrc_wgrad reads none of them)
        if(lane == 0)
        {
            sc.loss_partial[blockIdx.x * 2 + wave] = 0.0;
            sc.count_partial[blockIdx.x * 2 + wave] = 0;
        }
        return;
    }
    const int64_t i = sample0 + col;
    bool valid = false;
    float target[3] = { 0.0f, 0.0f, 0.0f };
    NrcVertex vtx{};
    if(i < count)
    {
        const int64_t slot = b.index ? (int64_t)b.index[first + i]
                                     : (b.permute ? nrc_record_slot(first + i) : first + i);
        const NrcRecord &r = b.rec[slot];
        valid = r.valid != 0;
        if(valid)
        {
            vtx = r.v;
            // radiance leaving the vertex: what the path gathered after it,
            // divided by the throughput the vertex was reached with
            const float *acc = b.acc + 4 * (size_t)r.path;
            for(int c = 0; c < 3; ++c)
                target[c] = r.beta[c] > 0.0f ? (acc[c] - r.gathered[c]) / r.beta[c] : 0.0f;
            if(!(isfinite(target[0]) && isfinite(target[1]) && isfinite(target[2])))
                valid = false;
        }
    }

    // lane (col, h) holds units 16 s + 8 h + 0..7 of sample col in slice s
    auto store_units = [&](uint16_t *dst, const Frag &f, int slices) {
#pragma unroll
        for(int s = 0; s < 4; ++s)
        {
            if(s >= slices)
                break;
#pragma unroll
            for(int j = 0; j < 8; ++j)
                dst[(int64_t)(16 * s + 8 * h + j) * S + i] = f.s[s][j];
        }
    };

    // forward: X[l] is the input of layer l
    Frag X[NRC_LAYERS];
    encode_frag(vtx, box, h, X[0]);
    store_units(sc.xt, X[0], 4);
#pragma unroll
    for(int l = 0; l < NRC_LAYERS - 1; ++l)
    {
        X[l + 1] = X[l];
        hidden_layer(s_w, l, col, h, X[l + 1]);
        store_units(sc.xt + (int64_t)(l + 1) * NRC_IN * S, X[l + 1], 4);
    }
    f32x16 out = {};
    {
        const uint16_t *r = s_w + ((NRC_LAYERS - 1) * NRC_WIDTH + col) * WS + 8 * h;
#pragma unroll
        for(int s = 0; s < 4; ++s)
            out = mfma(lds8(r + 16 * s), X[NRC_LAYERS - 1].s[s], out);
    }

    // loss and output gradient (lanes of half 0 hold rows 0..2 of their
    // sample): the sum over the batch, times loss_scale (the 1 / (3 n) of the
    // mean is applied with the update, once n is known)
    double loss = 0.0;
    float g[3] = { 0.0f, 0.0f, 0.0f };
    if(h == 0 && valid)
        for(int r = 0; r < NRC_OUT; ++r)
        {
            g[r] = nrc_rel_l2_grad(out[r], target[r]) * b.loss_scale;
            loss += (double)nrc_rel_l2(out[r], target[r]);
        }
    for(int off = 32; off > 0; off >>= 1)
        loss += __shfl_xor(loss, off);
    const int32_t nvalid = (int32_t)__popcll(__ballot(valid && h == 0));
    if(lane == 0)
    {
        sc.loss_partial[blockIdx.x * 2 + wave] = loss;
        sc.count_partial[blockIdx.x * 2 + wave] = nvalid;
    }
    // dz of the output layer as a B operand: rows 0..2 in slice 0, half 0
    Frag dz;
    dz.s[0] = words8(h == 0 ? pk_bf16(g[0], g[1]) : 0u, h == 0 ? pk_bf16(g[2], 0.0f) : 0u, 0u, 0u);
    dz.s[1] = dz.s[2] = dz.s[3] = words8(0u, 0u, 0u, 0u);
    {
        // the output layer's 32 gradient rows: 0..2 from half 0, zeros beyond
        uint16_t *d5 = sc.dzt + (int64_t)(NRC_LAYERS - 1) * NRC_IN * S;
#pragma unroll
        for(int j = 0; j < 16; ++j)
        {
            const int r = 16 * h + j;
            d5[(int64_t)r * S + i] = r < NRC_OUT ? dz.s[0][r] : (uint16_t)0;
        }
    }

    // backward: dX_l = W_l^T dz_l, then dz_(l-1) = dX_l where X_l > 0
#pragma unroll
    for(int l = NRC_LAYERS - 1; l >= 1; --l)
    {
        // the output layer's gradient is zero beyond row 2: one k slice
        const int slices = l == NRC_LAYERS - 1 ? 1 : 4;
        const uint16_t *r0 = s_wt + (l * NRC_IN + swap23(col)) * WS + 8 * h;
        const uint16_t *r1 = r0 + 32 * WS;
        f32x16 acc0 = {}, acc1 = {};
#pragma unroll
        for(int s = 0; s < 4; ++s)
        {
            if(s >= slices)
                break;
            acc0 = mfma(lds8(r0 + 16 * s), dz.s[s], acc0);
            acc1 = mfma(lds8(r1 + 16 * s), dz.s[s], acc1);
        }
#pragma unroll
        for(int q = 0; q < 4; ++q)
        {
            const f32x16 &a = q < 2 ? acc0 : acc1;
            const int o = 8 * (q & 1);
            const u16x8 x = X[l].s[q];
            u16x8 d;
#pragma unroll
            for(int j = 0; j < 8; j += 2)
            {
                const uint32_t v = pk_bf16(a[o + j], a[o + j + 1]);
                const uint32_t m = (x[j] ? 0xffffu : 0u) | (x[j + 1] ? 0xffff0000u : 0u);
                d[j] = (uint16_t)(v & m);
                d[j + 1] = (uint16_t)((v & m) >> 16);
            }
            dz.s[q] = d;
        }
        store_units(sc.dzt + (int64_t)(l - 1) * NRC_IN * S, dz, 4);
    }
}

// dW_l = sum over samples of dz_l (x) x_l. One wavefront per 32 x 32 tile
// and slice of the samples: tiles 0..19 cover the five 64 x 64 hidden
// layers, tiles 20..21 the output layer's 32 x 64; slice kc of NRC_KSPLIT
// writes its partial tile to partial[kc]. k_nrc_reduce adds the slices in
// a fixed order (deterministic).
__global__ void __launch_bounds__(64) k_nrc_wgrad(NrcScratch sc, NrcBatch b, float *partial)
{
    // operand rows have stride b.batch; the step's samples, rounded up to a
    // tile of 32 (the tail samples hold zeros), are the K extent
    const int64_t stride = b.batch;
    int64_t first, count;
    nrc_batch_range(b, &first, &count);
    const int64_t S = (count + 31) / 32 * 32;
    const int t = blockIdx.x % 22, kc = blockIdx.x / 22;
    const int l = t < 20 ? t / 4 : 5;
    const int mt = t < 20 ? (t % 4) / 2 : 0;
    const int nt = t < 20 ? t % 2 : t - 20;
    const int lane = threadIdx.x, col = lane % 32, half = lane / 32;
    const int64_t chunk = (S + NRC_KSPLIT * 16 - 1) / (NRC_KSPLIT * 16) * 16;
    const int64_t k0 = kc * chunk, k1 = k0 + chunk < S ? k0 + chunk : S;
    const uint16_t *dzt = sc.dzt + (int64_t)l * NRC_IN * stride;
    const uint16_t *xt = sc.xt + (int64_t)l * NRC_IN * stride;
    const uint16_t *arow = dzt + (int64_t)(mt * 32 + col) * stride + 8 * half;
    const uint16_t *brow = xt + (int64_t)(nt * 32 + col) * stride + 8 * half;
    f32x16 acc0 = {}, acc1 = {};
    int64_t k = k0;
    for(; k + 32 <= k1; k += 32)
    {
        const u16x8 a0 = *reinterpret_cast<const u16x8 *>(arow + k);
        const u16x8 b0 = *reinterpret_cast<const u16x8 *>(brow + k);
        const u16x8 a1 = *reinterpret_cast<const u16x8 *>(arow + k + 16);
        const u16x8 b1 = *reinterpret_cast<const u16x8 *>(brow + k + 16);
        acc0 = mfma(a0, b0, acc0);
        acc1 = mfma(a1, b1, acc1);
    }
    for(; k < k1; k += 16)
        acc0 = mfma(*reinterpret_cast<const u16x8 *>(arow + k), *reinterpret_cast<const u16x8 *>(brow + k), acc0);
    float *g = partial + (size_t)kc * NRC_PARAMS + nrc_layer_offset(l);
#pragma unroll
    for(int e = 0; e < 16; ++e)
    {
        const int row = mt * 32 + 8 * (e / 4) + 4 * half + (e % 4);
        g[row * NRC_IN + nt * 32 + col] = acc0[e] + acc1[e];
    }
}

// grad = sum over the NRC_KSPLIT slices, in slice order
// Also the workgroup's share of the squared gradient norm (FP64, fixed-order
// tree), so the step's reductions read one value per workgroup.
__global__ void __launch_bounds__(256) k_nrc_reduce(const float *partial, float *grad, double *g2_partial)
{
    __shared__ double s_g2[256];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double g2 = 0.0;
    if(k < NRC_PARAMS)
    {
        float sum = 0.0f;
        for(int c = 0; c < NRC_KSPLIT; ++c)
            sum += partial[(size_t)c * NRC_PARAMS + k];
        grad[k] = sum;
        g2 = (double)sum * (double)sum;
    }
    s_g2[threadIdx.x] = g2;
    __syncthreads();
    for(int off = 128; off > 0; off >>= 1)
    {
        if((int)threadIdx.x < off)
            s_g2[threadIdx.x] += s_g2[threadIdx.x + off];
        __syncthreads();
    }
    if(threadIdx.x == 0)
        g2_partial[blockIdx.x] = s_g2[0];
}

// Adam (Kingma and Ba 2015) on the FP32 master weights, one weight per
// thread, then its BF16 copies ([out][in] and per layer [in][out]). Every
// workgroup first forms the step's reductions itself (fixed order, FP64):
// valid samples n, loss = sum / (3 n), and the norm of the gradient of the
// mean loss, which give the scale that turns the summed, loss-scaled
// gradient into the mean's gradient, or 0 (no update) for a non-finite
// gradient or an empty batch; workgroup 0 records them.
__global__ void __launch_bounds__(256) k_nrc_update(NrcWeights net, NrcAdam opt, const float *grad, NrcScratch sc,
                                                    int32_t partials, int32_t *valid_count, NrcStepStats *stats)
{
    __shared__ double s_a[256], s_b[256];
    __shared__ long long s_c[256];
    const int tid = threadIdx.x;
    double loss = 0.0;
    long long cnt = 0;
    for(int k = tid; k < partials; k += 256)
    {
        loss += sc.loss_partial[k];
        cnt += sc.count_partial[k];
    }
    double g2 = 0.0;
    for(int k = tid; k < NRC_REDUCE_BLOCKS; k += 256)
        g2 += sc.g2_partial[k];
    s_a[tid] = loss;
    s_b[tid] = g2;
    s_c[tid] = cnt;
    __syncthreads();
    for(int off = 128; off > 0; off >>= 1)
    {
        if(tid < off)
        {
            s_a[tid] += s_a[tid + off];
            s_b[tid] += s_b[tid + off];
            s_c[tid] += s_c[tid + off];
        }
        __syncthreads();
    }
    const long long n = s_c[0];
    const double dscale = n > 0 ? 1.0 / (3.0 * (double)n * (double)opt.loss_scale) : 0.0;
    const double gnorm = sqrt(s_b[0]) * dscale;
    const bool step = n > 0 && isfinite(gnorm);
    const float scale = step ? (float)dscale : 0.0f;
    if(blockIdx.x == 0 && tid == 0)
    {
        *sc.step_scale = scale;
        *valid_count = (int32_t)n;
        stats->loss = n > 0 ? s_a[0] / (3.0 * (double)n) : 0.0;
        stats->grad_norm = gnorm;
        stats->samples = (int32_t)n;
        stats->skipped = step ? 0 : 1;
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= NRC_PARAMS || scale == 0.0f)
        return;
    const float g = grad[k] * scale;
    const float m = opt.beta1 * net.m[k] + (1.0f - opt.beta1) * g;
    const float v = opt.beta2 * net.v[k] + (1.0f - opt.beta2) * g * g;
    net.m[k] = m;
    net.v[k] = v;
    const float mhat = m * opt.inv_bias1;
    const float vhat = v * opt.inv_bias2;
    const float w = net.w[k] - opt.lr * mhat / (sqrtf(vhat) + opt.eps);
    net.w[k] = w;
    const int l = k / (NRC_WIDTH * NRC_IN) < NRC_LAYERS ? k / (NRC_WIDTH * NRC_IN) : NRC_LAYERS - 1;
    const int off = nrc_layer_offset(l);
    const int rows = nrc_layer_rows(l);
    const int r = (k - off) / NRC_IN, c = (k - off) % NRC_IN;
    const uint16_t h = to_bf16(w);
    net.w_bf16[k] = h;
    net.wt_bf16[off + c * rows + r] = h;
    if(opt.ema_decay > 0.0f)
    {
        const float e = opt.ema_decay * net.ema[k] + (1.0f - opt.ema_decay) * w;
        net.ema[k] = e;
        net.ema_bf16[k] = to_bf16(e);
    }
}

// valid record slots in the order of the slot bijection, in two passes:
// k_nrc_valid_mask gathers the valid flags of bijection indices 64 w..64 w+63
// as one 64-bit ballot per wavefront; k_nrc_compact (one workgroup of 1024
// threads, one mask word each) counts, scans and writes the slots in order
__global__ void __launch_bounds__(64) k_nrc_valid_mask(const NrcRecord *rec, uint64_t *mask)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const uint64_t b = __ballot(rec[nrc_record_slot(i)].valid != 0);
    if(threadIdx.x == 0)
        mask[blockIdx.x] = b;
}

__global__ void __launch_bounds__(1024) k_nrc_compact(const uint64_t *mask, int32_t *index, int32_t *total)
{
    __shared__ int32_t s_sum[1024];
    const int t = threadIdx.x;
    const uint64_t m = mask[t];
    const int32_t n = (int32_t)__popcll(m);
    s_sum[t] = n;
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for(int off = 1; off < 1024; off <<= 1)
    {
        const int32_t v = t >= off ? s_sum[t - off] : 0;
        __syncthreads();
        s_sum[t] += v;
        __syncthreads();
    }
    int32_t o = s_sum[t] - n;
    for(uint64_t r = m; r != 0; r &= r - 1)
        index[o++] = (int32_t)nrc_record_slot(64 * t + __ffsll((unsigned long long)r) - 1);
    if(t == 1023)
        *total = s_sum[1023];
}

// clears the record flags and the training paths' radiance sums of a frame
__global__ void k_nrc_clear(NrcRecord *rec, int64_t nrec, float *acc, int64_t npaths)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nrec; k += stride)
        rec[k].valid = 0;
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < 4 * npaths; k += stride)
        acc[k] = 0.0f;
}

// FP32 master weights -> BF16 copies (after initialisation or a host write)
__global__ void k_nrc_publish(NrcWeights net)
{
    for(int k = blockIdx.x * blockDim.x + threadIdx.x; k < NRC_PARAMS; k += gridDim.x * blockDim.x)
    {
        const int l = k / (NRC_WIDTH * NRC_IN) < NRC_LAYERS ? k / (NRC_WIDTH * NRC_IN) : NRC_LAYERS - 1;
        const int off = nrc_layer_offset(l);
        const int rows = nrc_layer_rows(l);
        const int r = (k - off) / NRC_IN, c = (k - off) % NRC_IN;
        const uint16_t h = to_bf16(net.w[k]);
        net.w_bf16[k] = h;
        net.wt_bf16[off + c * rows + r] = h;
        // the moving average restarts at the published weights
        net.ema[k] = net.w[k];
        net.ema_bf16[k] = h;
    }
}

// ---------------------------------------------------------------- launchers

void launch_nrc_infer(hipStream_t s, int cu_count, int64_t max_queries, const NrcWeights &net, const NrcBox &box,
                      const NrcQuery *q, const int32_t *count, float *film, double *film64, float *acc)
{
    // persistent grid: as many workgroups as the LDS copies of the weights
    // (and the registers) let stay resident, fewer for a small query bound
    const int64_t span = (int64_t)INFER_THREADS / WAVE * INFER_TILES * TILE;
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)INFER_BLOCKS_PER_CU * cu_count, (max_queries + span - 1) / span));
    hipLaunchKernelGGL(k_nrc_infer, dim3(grid), dim3(INFER_THREADS), 0, s, net, box, q, count, film, film64,
                       acc);
}

void launch_nrc_eval(hipStream_t s, const NrcWeights &net, const NrcBox &box, const NrcVertex *v, int32_t n,
                     float *out)
{
    hipLaunchKernelGGL(k_nrc_eval, dim3((n + 127) / 128), dim3(256), 0, s, net, box, v, n, out);
}

void launch_nrc_train_step(hipStream_t s, const NrcWeights &net, const NrcBox &box, const NrcBatch &b,
                           const NrcScratch &sc, const NrcAdam &opt, float *grad, NrcStepStats *stats)
{
    launch_nrc_gradients(s, net, box, b, sc, grad);
    const int blocks = (int)(b.batch / (2 * TILE));
    hipLaunchKernelGGL(k_nrc_update, dim3((NRC_PARAMS + 255) / 256), dim3(256), 0, s, net, opt, grad, sc, blocks * 2,
                       b.valid_count, stats);
}

void launch_nrc_gradients(hipStream_t s, const NrcWeights &net, const NrcBox &box, const NrcBatch &b,
                          const NrcScratch &sc, float *grad)
{
    const int blocks = (int)(b.batch / (2 * TILE));
    hipLaunchKernelGGL(k_nrc_train_fb, dim3(blocks), dim3(128), 0, s, net, box, b, sc);
    hipLaunchKernelGGL(k_nrc_wgrad, dim3(22 * NRC_KSPLIT), dim3(64), 0, s, sc, b, sc.grad_partial);
    hipLaunchKernelGGL(k_nrc_reduce, dim3(NRC_REDUCE_BLOCKS), dim3(256), 0, s, sc.grad_partial, grad, sc.g2_partial);
}

void launch_nrc_clear(hipStream_t s, NrcRecord *rec, int64_t nrec, float *acc, int64_t npaths)
{
    hipLaunchKernelGGL(k_nrc_clear, dim3(256), dim3(256), 0, s, rec, nrec, acc, npaths);
}

void launch_nrc_compact(hipStream_t s, const NrcRecord *rec, uint64_t *mask, int32_t *index, int32_t *total)
{
    hipLaunchKernelGGL(k_nrc_valid_mask, dim3((1 << 16) / 64), dim3(64), 0, s, rec, mask);
    hipLaunchKernelGGL(k_nrc_compact, dim3(1), dim3(1024), 0, s, mask, index, total);
}

void launch_nrc_publish(hipStream_t s, const NrcWeights &net)
{
    hipLaunchKernelGGL(k_nrc_publish, dim3((NRC_PARAMS + 255) / 256), dim3(256), 0, s, net);
}

} // namespace btrc
