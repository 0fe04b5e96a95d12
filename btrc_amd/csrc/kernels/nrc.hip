// Neural radiance cache kernels for gfx950: inference, fused forward +
// backward of a training batch, the weight-gradient products and Adam.
//
// All matrix products run on MFMA (v_mfma_f32_32x32x16_bf16): BF16 operands,
// FP32 accumulation. One wavefront owns 32 samples; a layer is
// Y^T = W X^T with M = output units, N = the 32 samples, K = input units, so
// the 16 results a lane holds belong to one sample (lane % 32) and land in
// LDS as that sample's next input row. Weights sit in LDS as [out][in] (A
// operand) and, for the backward pass, as [in][out].
//
// Lane layout of the 32x32x16 MFMA (checked on the device before use):
//   A (32 x 16): lane l holds row l % 32, k = 8 (l / 32) + 0..7
//   B (16 x 32): lane l holds column l % 32, k = 8 (l / 32) + 0..7
//   D (32 x 32): element i of lane l is row 8 (i / 4) + 4 (l / 32) + i % 4,
//                column l % 32
#include <hip/hip_runtime.h>

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

// weights of all layers into LDS as [out][in] and, optionally, [in][out]
__device__ void load_weights(const uint16_t *w, const uint16_t *wt, uint16_t *s_w, uint16_t *s_wt)
{
    for(int i = threadIdx.x; i < NRC_PARAMS / 8; i += blockDim.x)
    {
        reinterpret_cast<uint4 *>(s_w)[i] = reinterpret_cast<const uint4 *>(w)[i];
        if(s_wt)
            reinterpret_cast<uint4 *>(s_wt)[i] = reinterpret_cast<const uint4 *>(wt)[i];
    }
}

// the encoded inputs of sample `col` of this wavefront into x (row-major
// [sample][64]); lanes 32..63 only wait
__device__ void encode_tile(const NrcVertex *v, bool valid, const NrcBox &box, uint16_t *x, int col, int half)
{
    if(half == 0)
    {
        float in[NRC_IN];
        if(valid)
            nrc_encode(*v, box, in);
        else
            for(int k = 0; k < NRC_IN; ++k)
                in[k] = 0.0f;
        for(int k = 0; k < NRC_IN; ++k)
            x[col * NRC_IN + k] = to_bf16(in[k]);
    }
}

// One layer forward for the wavefront's 32 samples: out = W_l x (ReLU on
// hidden layers, BF16 into `next`); the output layer returns y for lanes
// 0..31 in y[0..2]. With xt, hidden outputs also go to global memory as
// [unit][sample] (the next layer's input, for the weight gradients).
__device__ void layer_forward(int l, const uint16_t *s_w, const uint16_t *x, uint16_t *next, float *y,
                              uint16_t *xt, int64_t stride, int64_t sample0)
{
    const int lane = threadIdx.x % WAVE, col = lane % 32, half = lane / 32;
    const uint16_t *w = s_w + nrc_layer_offset(l);
    const int tiles = nrc_layer_rows(l) / 32;
    for(int mt = 0; mt < tiles; ++mt)
    {
        f32x16 acc = {};
#pragma unroll
        for(int s = 0; s < NRC_IN / 16; ++s)
        {
            const u16x8 a = lds8(w + (mt * 32 + col) * NRC_IN + 16 * s + 8 * half);
            const u16x8 b = lds8(x + col * NRC_IN + 16 * s + 8 * half);
            acc = mfma(a, b, acc);
        }
        if(l == NRC_LAYERS - 1)
        {
            if(half == 0)
            {
                y[0] = acc[0];
                y[1] = acc[1];
                y[2] = acc[2];
            }
        }
        else
        {
#pragma unroll
            for(int i = 0; i < 16; ++i)
            {
                const int row = mt * 32 + 8 * (i / 4) + 4 * half + (i % 4);
                const uint16_t h = to_bf16(fmaxf(acc[i], 0.0f));
                next[col * NRC_IN + row] = h;
                if(xt)
                    xt[(int64_t)row * stride + sample0 + col] = h;
            }
        }
    }
}

} // namespace

// ---------------------------------------------------------------- inference

// queries [0, n): y = max(net(x), 0); the film gets beta * y at the pixel,
// and a training path's radiance sum too when the query is its tail (one
// query per training path, after every other addition to its sum)
__global__ void __launch_bounds__(256) k_nrc_infer(NrcWeights net, NrcBox box, const NrcQuery *q, const int32_t *count,
                                                   float *film, double *film64, float *acc)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_w[NRC_PARAMS];
    __shared__ __attribute__((aligned(16))) uint16_t s_x[4][2][TILE * NRC_IN];
    load_weights(net.w_bf16, nullptr, s_w, nullptr);
    __syncthreads();
    const int n = *count;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, col = lane % 32, half = lane / 32;
    const int waves_total = gridDim.x * 4;
    for(int base = (blockIdx.x * 4 + wave) * TILE; base < n; base += waves_total * TILE)
    {
        const int i = base + col;
        const bool valid = i < n;
        encode_tile(valid ? &q[i].v : nullptr, valid, box, s_x[wave][0], col, half);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float y[3] = { 0.0f, 0.0f, 0.0f };
        for(int l = 0; l < NRC_LAYERS; ++l)
        {
            layer_forward(l, s_w, s_x[wave][l & 1], s_x[wave][(l + 1) & 1], y, nullptr, 0, 0);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if(half == 0 && valid)
        {
            const NrcQuery &qi = q[i];
            const float v[3] = { qi.beta[0] * fmaxf(y[0], 0.0f), qi.beta[1] * fmaxf(y[1], 0.0f),
                                 qi.beta[2] * fmaxf(y[2], 0.0f) };
            if(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]))
            {
                float *p = film + 4 * (size_t)qi.pixel;
                atomicAdd(p + 0, v[0]);
                atomicAdd(p + 1, v[1]);
                atomicAdd(p + 2, v[2]);
                if(film64)
                {
                    double *d = film64 + 3 * (size_t)qi.pixel;
                    atomicAdd(d + 0, (double)v[0]);
                    atomicAdd(d + 1, (double)v[1]);
                    atomicAdd(d + 2, (double)v[2]);
                }
                if(qi.train != NO_TRAIN)
                {
                    float *a = acc + 4 * (size_t)qi.train;
                    a[0] = a[0] + v[0];
                    a[1] = a[1] + v[1];
                    a[2] = a[2] + v[2];
                }
            }
        }
    }
}

// raw network outputs for test hooks: out[3 i + c]
__global__ void __launch_bounds__(256) k_nrc_eval(NrcWeights net, NrcBox box, const NrcVertex *v, int32_t n,
                                                  float *out)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_w[NRC_PARAMS];
    __shared__ __attribute__((aligned(16))) uint16_t s_x[4][2][TILE * NRC_IN];
    load_weights(net.w_bf16, nullptr, s_w, nullptr);
    __syncthreads();
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, col = lane % 32, half = lane / 32;
    const int base = (blockIdx.x * 4 + wave) * TILE;
    const int i = base + col;
    const bool valid = i < n;
    encode_tile(valid ? &v[i] : nullptr, valid, box, s_x[wave][0], col, half);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float y[3] = { 0.0f, 0.0f, 0.0f };
    for(int l = 0; l < NRC_LAYERS; ++l)
    {
        layer_forward(l, s_w, s_x[wave][l & 1], s_x[wave][(l + 1) & 1], y, nullptr, 0, 0);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
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
__global__ void __launch_bounds__(128) k_nrc_train_fb(NrcWeights net, NrcBox box, NrcBatch b, NrcScratch sc)
{
    __shared__ __attribute__((aligned(16))) uint16_t s_w[NRC_PARAMS];
    __shared__ __attribute__((aligned(16))) uint16_t s_wt[NRC_PARAMS];
    __shared__ __attribute__((aligned(16))) uint16_t s_x[2][NRC_LAYERS][TILE * NRC_IN];
    __shared__ __attribute__((aligned(16))) uint16_t s_dz[2][2][TILE * NRC_IN];
    load_weights(net.w_bf16, net.wt_bf16, s_w, s_wt);
    __syncthreads();

    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, col = lane % 32, half = lane / 32;
    const int64_t S = b.batch;
    const int64_t sample0 = (int64_t)(blockIdx.x * 2 + wave) * TILE;
    const int64_t i = sample0 + col;
    bool valid = false;
    float target[3] = { 0.0f, 0.0f, 0.0f };
    const NrcVertex *vtx = nullptr;
    if(i < S)
    {
        const NrcRecord &r = b.rec[b.permute ? nrc_record_slot(b.first + i) : b.first + i];
        valid = r.valid != 0;
        if(valid)
        {
            vtx = &r.v;
            // radiance leaving the vertex: what the path gathered after it,
            // divided by the throughput the vertex was reached with
            const float *acc = b.acc + 4 * (size_t)r.path;
            for(int c = 0; c < 3; ++c)
                target[c] = r.beta[c] > 0.0f ? (acc[c] - r.gathered[c]) / r.beta[c] : 0.0f;
            if(!(isfinite(target[0]) && isfinite(target[1]) && isfinite(target[2])))
                valid = false;
        }
    }

    uint16_t(*x)[TILE * NRC_IN] = s_x[wave];
    encode_tile(vtx, valid, box, x[0], col, half);
    __syncthreads();
    // the encoded inputs also go to global memory for the first layer's gradient
    if(half == 0)
        for(int k = 0; k < NRC_IN; ++k)
            sc.xt[(int64_t)k * S + i] = x[0][col * NRC_IN + k];

    // forward
    float y[3] = { 0.0f, 0.0f, 0.0f };
    for(int l = 0; l < NRC_LAYERS; ++l)
    {
        uint16_t *xt_next = l + 1 < NRC_LAYERS ? sc.xt + (int64_t)(l + 1) * NRC_IN * S : nullptr;
        layer_forward(l, s_w, x[l], l + 1 < NRC_LAYERS ? x[l + 1] : nullptr, y, xt_next, S, sample0);
        __syncthreads();
    }

    // loss and output gradient: the sum over the batch, times loss_scale (the
    // 1 / (3 n) of the mean is applied with the update, once n is known)
    double loss = 0.0;
    uint16_t *dz = s_dz[wave][0];
    if(half == 0)
    {
        for(int r = 0; r < NRC_OUT_ROWS; ++r)
        {
            float g = 0.0f;
            if(r < NRC_OUT && valid)
            {
                g = nrc_rel_l2_grad(y[r], target[r]) * b.loss_scale;
                loss += (double)nrc_rel_l2(y[r], target[r]);
            }
            dz[col * NRC_IN + r] = to_bf16(g);
        }
    }
    for(int off = 32; off > 0; off >>= 1)
        loss += __shfl_xor(loss, off);
    const int32_t nvalid = (int32_t)__popcll(__ballot(valid && half == 0));
    if(lane == 0)
    {
        sc.loss_partial[blockIdx.x * 2 + wave] = loss;
        sc.count_partial[blockIdx.x * 2 + wave] = nvalid;
    }
    __syncthreads();

    // backward: dz holds d/d(out) of layer l as [sample][unit]
    for(int l = NRC_LAYERS - 1; l >= 0; --l)
    {
        const int rows = nrc_layer_rows(l);
        uint16_t *dzt = sc.dzt + (int64_t)l * NRC_IN * S;
        if(half == 0)
            for(int r = 0; r < rows; ++r)
                dzt[(int64_t)r * S + i] = dz[col * NRC_IN + r];
        if(l == 0)
            break;
        // d/d(input of l) = W_l^T dz, masked by the ReLU of layer l - 1
        uint16_t *dz_next = s_dz[wave][(NRC_LAYERS - l) & 1];
        const uint16_t *wt = s_wt + nrc_layer_offset(l); // [in][rows]
        for(int mt = 0; mt < NRC_IN / 32; ++mt)
        {
            f32x16 acc = {};
            for(int s = 0; s < rows / 16; ++s)
            {
                const u16x8 a = lds8(wt + (mt * 32 + col) * rows + 16 * s + 8 * half);
                const u16x8 bb = lds8(dz + col * NRC_IN + 16 * s + 8 * half);
                acc = mfma(a, bb, acc);
            }
#pragma unroll
            for(int e = 0; e < 16; ++e)
            {
                const int unit = mt * 32 + 8 * (e / 4) + 4 * half + (e % 4);
                const bool on = (x[l][col * NRC_IN + unit] & 0x7fffu) != 0; // ReLU output > 0
                dz_next[col * NRC_IN + unit] = on ? to_bf16(acc[e]) : (uint16_t)0;
            }
        }
        __syncthreads();
        dz = dz_next;
    }
}

// dW_l = sum over samples of dz_l (x) x_l. One wavefront per 32 x 32 tile
// and slice of the samples: tiles 0..19 cover the five 64 x 64 hidden
// layers, tiles 20..21 the output layer's 32 x 64; slice kc of NRC_KSPLIT
// writes its partial tile to partial[kc]. k_nrc_reduce adds the slices in
// a fixed order (deterministic).
__global__ void __launch_bounds__(64) k_nrc_wgrad(NrcScratch sc, int64_t S, float *partial)
{
    const int t = blockIdx.x % 22, kc = blockIdx.x / 22;
    const int l = t < 20 ? t / 4 : 5;
    const int mt = t < 20 ? (t % 4) / 2 : 0;
    const int nt = t < 20 ? t % 2 : t - 20;
    const int lane = threadIdx.x, col = lane % 32, half = lane / 32;
    const int64_t chunk = (S + NRC_KSPLIT * 16 - 1) / (NRC_KSPLIT * 16) * 16;
    const int64_t k0 = kc * chunk, k1 = k0 + chunk < S ? k0 + chunk : S;
    const uint16_t *dzt = sc.dzt + (int64_t)l * NRC_IN * S;
    const uint16_t *xt = sc.xt + (int64_t)l * NRC_IN * S;
    const uint16_t *arow = dzt + (int64_t)(mt * 32 + col) * S + 8 * half;
    const uint16_t *brow = xt + (int64_t)(nt * 32 + col) * S + 8 * half;
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

// Reductions of a step (one workgroup, fixed order, FP64): valid samples n,
// loss = sum / (3 n), and the norm of the gradient of the mean loss. Sets
// the scale that turns the summed, loss-scaled gradient into the mean's
// gradient, or 0 (no update) for a non-finite gradient or an empty batch.
__global__ void __launch_bounds__(256) k_nrc_finalize(const float *grad, NrcScratch sc, int32_t partials,
                                                      float loss_scale, int32_t *valid_count, NrcStepStats *stats)
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
    if(tid == 0)
    {
        const long long n = s_c[0];
        const double scale = n > 0 ? 1.0 / (3.0 * (double)n * (double)loss_scale) : 0.0;
        const double gnorm = sqrt(s_b[0]) * scale;
        const bool step = n > 0 && isfinite(gnorm);
        *sc.step_scale = step ? (float)scale : 0.0f;
        *valid_count = (int32_t)n;
        stats->loss = n > 0 ? s_a[0] / (3.0 * (double)n) : 0.0;
        stats->grad_norm = gnorm;
        stats->samples = (int32_t)n;
        stats->skipped = step ? 0 : 1;
    }
}

// Adam (Kingma and Ba 2015) on the FP32 master weights, one weight per
// thread, then its BF16 copies ([out][in] and per layer [in][out])
__global__ void __launch_bounds__(256) k_nrc_update(NrcWeights net, NrcAdam opt, const float *grad, NrcScratch sc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const float scale = *sc.step_scale;
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
    }
}

// ---------------------------------------------------------------- launchers

void launch_nrc_infer(hipStream_t s, int grid, const NrcWeights &net, const NrcBox &box, const NrcQuery *q,
                      const int32_t *count, float *film, double *film64, float *acc)
{
    hipLaunchKernelGGL(k_nrc_infer, dim3(grid), dim3(256), 0, s, net, box, q, count, film, film64, acc);
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
    hipLaunchKernelGGL(k_nrc_finalize, dim3(1), dim3(256), 0, s, grad, sc, blocks * 2, b.loss_scale, b.valid_count,
                       stats);
    hipLaunchKernelGGL(k_nrc_update, dim3((NRC_PARAMS + 255) / 256), dim3(256), 0, s, net, opt, grad, sc);
}

void launch_nrc_gradients(hipStream_t s, const NrcWeights &net, const NrcBox &box, const NrcBatch &b,
                          const NrcScratch &sc, float *grad)
{
    const int blocks = (int)(b.batch / (2 * TILE));
    hipLaunchKernelGGL(k_nrc_train_fb, dim3(blocks), dim3(128), 0, s, net, box, b, sc);
    hipLaunchKernelGGL(k_nrc_wgrad, dim3(22 * NRC_KSPLIT), dim3(64), 0, s, sc, b.batch, sc.grad_partial);
    hipLaunchKernelGGL(k_nrc_reduce, dim3(NRC_REDUCE_BLOCKS), dim3(256), 0, s, sc.grad_partial, grad, sc.g2_partial);
}

void launch_nrc_clear(hipStream_t s, NrcRecord *rec, int64_t nrec, float *acc, int64_t npaths)
{
    hipLaunchKernelGGL(k_nrc_clear, dim3(256), dim3(256), 0, s, rec, nrec, acc, npaths);
}

void launch_nrc_publish(hipStream_t s, const NrcWeights &net)
{
    hipLaunchKernelGGL(k_nrc_publish, dim3((NRC_PARAMS + 255) / 256), dim3(256), 0, s, net);
}

} // namespace btrc
