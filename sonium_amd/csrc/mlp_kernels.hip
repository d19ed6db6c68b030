// Learned warm start: a 4-layer MLP (64 features -> W -> W -> W -> 2) with
// BF16 weights and activations, FP32 accumulation on
// v_mfma_f32_32x32x16_bf16, trained with hand-written forward and backward
// kernels and Adam on FP32 master weights.
//
// GEMM.  One kernel computes C = A B^T for row-major BF16 A (M x K) and B
// (N x K): both operands are K-contiguous, so every MFMA operand is one
// 16-byte load per lane.  The three products of a layer use it as
//   forward      H   = A_prev W^T          (A_prev: M x K_in,  W:  N_out x K_in)
//   input grad   dA  = dZ (W^T)^T          (dZ:     M x N_out, W^T: K_in x N_out)
//   weight grad  dW  = dZ^T (A_prev^T)^T   (dZ^T:   N_out x M, A_prev^T: K_in x M)
// so the elementwise kernels write activations and their gradients in both
// layouts, and the weight copy is kept as W and W^T.  Weight gradients reduce
// over the M samples with split K (FP32 partials, summed in a fixed order).
//
// Features of a sample (c centroid, n normal, d incident direction, k):
//   c, n, d, k/12, d.n, d.c, then sin/cos(2^l pi q) for l = 0..5 and
//   q in {d.c, d.n, k/12, (k/12)(d.c)}: 12 + 48 = 60, zero padded to 64.
// Activation: SiLU.  Output: 2 values (Re, Im of the incident-phase-free
// density), padded to 64 rows in the last layer.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

#include "internal.hpp"

namespace sbh {
namespace {

#include "lowp_common.hpp"

constexpr int FEAT = 64;
constexpr int FEAT_USED = 60;
constexpr int OUT_PAD = 64;
constexpr int RAW = 10;  // c[3], n[3], d[3], k

__device__ __forceinline__ u16 to_bf16(float f) { return Half<1>::from_float(f); }
__device__ __forceinline__ float from_bf16(u16 v) { return Half<1>::to_float(v); }

// C[z][M][N] = A[M, kz] B[N, kz]^T over the K slice z (tile 64 x 64, 4 waves of 32 x 32).
__global__ __launch_bounds__(BLOCK) void mlp_gemm_kernel(const u16* __restrict__ A, int lda, const u16* __restrict__ B,
                                                         int ldb, int M, int N, int kslice, float* __restrict__ C) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int m0 = blockIdx.x * 64 + 32 * (wave >> 1);
    const int n0 = blockIdx.y * 64 + 32 * (wave & 1);
    const int64_t k0 = (int64_t)blockIdx.z * kslice;
    const int kq = (lane >> 5) * 8;
    const u16* ap = A + (int64_t)(m0 + (lane & 31)) * lda + k0 + kq;
    const u16* bp = B + (int64_t)(n0 + (lane & 31)) * ldb + k0 + kq;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    for (int k = 0; k < kslice; k += 16) {
        const u32x4 a = *reinterpret_cast<const u32x4*>(ap + k);
        const u32x4 b = *reinterpret_cast<const u32x4*>(bp + k);
        acc = mfma32<1>(a, b, acc);
    }
    float* out = C + (int64_t)blockIdx.z * M * N;
    const int col = n0 + (lane & 31);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        out[(int64_t)row * N + col] = acc[v];
    }
}

// The 64 features of one raw sample {c, n, d, k} (FEAT_USED used, zero padded).
__device__ __forceinline__ void sample_features(const float* __restrict__ r, float* f) {
    const float c0 = r[0], c1 = r[1], c2 = r[2], n0 = r[3], n1 = r[4], n2 = r[5], d0 = r[6], d1 = r[7], d2 = r[8];
    const float kk = r[9] * (1.0f / 12.0f);
    const float dc = d0 * c0 + d1 * c1 + d2 * c2;
    const float dn = d0 * n0 + d1 * n1 + d2 * n2;
    f[0] = c0; f[1] = c1; f[2] = c2;
    f[3] = n0; f[4] = n1; f[5] = n2;
    f[6] = d0; f[7] = d1; f[8] = d2;
    f[9] = kk; f[10] = dn; f[11] = dc;
    const float q[4] = {dc, dn, kk, kk * dc};
    int o = 12;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        float w = 3.14159265358979f;
#pragma unroll
        for (int l = 0; l < 6; ++l) {
            f[o++] = sinf(w * q[a]);
            f[o++] = cosf(w * q[a]);
            w *= 2.0f;
        }
    }
#pragma unroll
    for (int j = FEAT_USED; j < FEAT; ++j) f[j] = 0.f;
}

// Features of samples idx[0..M) (or 0..M when idx is null) -> X (M x 64) and
// X^T (64 x M; not written when XT is null, as in inference).
__global__ void mlp_features_kernel(const float* __restrict__ raw, const int64_t* __restrict__ idx, int M,
                                    u16* __restrict__ X, u16* __restrict__ XT) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    float f[FEAT];
    sample_features(raw + (idx ? idx[s] : s) * RAW, f);
#pragma unroll
    for (int j = 0; j < FEAT; ++j) {
        const u16 h = to_bf16(f[j]);
        X[(int64_t)s * FEAT + j] = h;
        if (XT) XT[(int64_t)j * M + s] = h;
    }
}

__device__ __forceinline__ float silu(float x) { return x / (1.0f + expf(-x)); }
__device__ __forceinline__ float silu_grad(float x) {
    const float sg = 1.0f / (1.0f + expf(-x));
    return sg * (1.0f + x * (1.0f - sg));
}

// H = sum of nsplit partials + bias (FP32, kept for the backward pass);
// A = silu(H) in BF16, row-major and transposed (H and AT are null in
// inference: only training reads them).
__global__ void mlp_bias_act_kernel(const float* __restrict__ part, int nsplit, const float* __restrict__ bias, int M,
                                    int N, float* __restrict__ H, u16* __restrict__ A, u16* __restrict__ AT) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * N) return;
    const int m = (int)(t / N), j = (int)(t % N);
    float h = bias[j];
    for (int z = 0; z < nsplit; ++z) h += part[(int64_t)z * M * N + t];
    if (H) H[t] = h;
    const u16 a = to_bf16(silu(h));
    A[t] = a;
    if (AT) AT[(int64_t)j * M + m] = a;
}

// Output layer: Y = partial sum + bias (first 2 of 64 padded columns);
// loss = sum (Y - T)^2 / (2 M); dZ = (Y - T) / M in BF16 (both layouts, zero
// in the padded columns); loss partial per block.
__global__ void mlp_output_kernel(const float* __restrict__ part, int nsplit, const float* __restrict__ bias, int M,
                                  const float* __restrict__ targets, const int64_t* __restrict__ idx,
                                  float* __restrict__ Y, u16* __restrict__ dZ, u16* __restrict__ dZT,
                                  double* __restrict__ loss_part) {
    __shared__ double sh[BLOCK];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (t < (int64_t)M * OUT_PAD) {
        const int m = (int)(t / OUT_PAD), j = (int)(t % OUT_PAD);
        u16 g = 0;
        if (j < 2) {
            float y = bias[j];
            for (int z = 0; z < nsplit; ++z) y += part[(int64_t)z * M * OUT_PAD + t];
            Y[(int64_t)m * 2 + j] = y;
            if (targets) {
                const float e = y - targets[(idx ? idx[m] : m) * 2 + j];
                l = 0.5 * (double)e * (double)e;
                g = to_bf16(e / (float)M);
            }
        }
        if (dZ) {
            dZ[t] = g;
            dZT[(int64_t)j * M + m] = g;
        }
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sh[0];
}

// dZ = dA * silu'(H) from the summed partials of dA; BF16 in both layouts.
__global__ void mlp_act_grad_kernel(const float* __restrict__ part, int nsplit, const float* __restrict__ H, int M,
                                    int N, u16* __restrict__ dZ, u16* __restrict__ dZT) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * N) return;
    const int m = (int)(t / N), j = (int)(t % N);
    float da = 0.f;
    for (int z = 0; z < nsplit; ++z) da += part[(int64_t)z * M * N + t];
    const u16 g = to_bf16(da * silu_grad(H[t]));
    dZ[t] = g;
    dZT[(int64_t)j * M + m] = g;
}

// out[t] = sum over the nsplit partial slabs (slab stride `stride`) for t < count;
// count may be a prefix of the slab (the real rows of a padded layer).
__global__ void mlp_sum_kernel(const float* __restrict__ part, int nsplit, int64_t stride, int64_t count,
                               float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    float s = 0.f;
    for (int z = 0; z < nsplit; ++z) s += part[(int64_t)z * stride + t];
    out[t] = s;
}

// Bias gradient: row sums of dZ^T (N x M), one block per row, fixed order.
__global__ __launch_bounds__(BLOCK) void mlp_bias_grad_kernel(const u16* __restrict__ dZT, int M, float* __restrict__ out) {
    __shared__ float sh[BLOCK];
    const u16* row = dZT + (int64_t)blockIdx.x * M;
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += BLOCK) s += from_bf16(row[m]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = BLOCK / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Adam on the FP32 master parameters (t = step number from 1).
__global__ void mlp_adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                                float* __restrict__ m2, int64_t count, float lr, float b1, float b2, float eps,
                                float c1, float c2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float gr = g[t];
    const float a = b1 * m1[t] + (1.f - b1) * gr;
    const float b = b2 * m2[t] + (1.f - b2) * gr * gr;
    m1[t] = a;
    m2[t] = b;
    p[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 copies W (rows x cols, padded to prow rows) and W^T (cols x prow).
__global__ void mlp_cast_kernel(const float* __restrict__ w, int rows, int cols, int prow, u16* __restrict__ W,
                                u16* __restrict__ WT) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)prow * cols) return;
    const int r = (int)(t / cols), c = (int)(t % cols);
    const u16 h = r < rows ? to_bf16(w[(int64_t)r * cols + c]) : (u16)0;
    W[t] = h;
    WT[(int64_t)c * prow + r] = h;
}

// Sample indices for a step: idx[s] = hash(seed, step, s) mod count.
__global__ void mlp_sample_kernel(int64_t* __restrict__ idx, int M, uint64_t seed, int64_t count) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(s + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    idx[s] = (int64_t)(z % (uint64_t)count);
}

inline unsigned nb(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }


// Inference through all layers: a workgroup carries INFER_ROWS samples, their
// activations held in LDS (two row-major buffers of ldr halves per row, rows
// padded by 8 halves); each wave computes 32-column tiles of a layer with
// the GEMM's operand layout and accumulation order (one MFMA per 16 inputs,
// in order), then bias + SiLU in registers, so the outputs equal the
// layer-by-layer forward pass bit for bit.  Weights stream from L2.
constexpr int INFER_ROWS = 32;
constexpr int INFER_MAX_TILES = 8;  // 32-column tiles per wave (TPW <= 8): width <= 4 * 8 * 32 = 1024
struct InferArgs {
    const u16* w[4];
    const float* b[4];
    int in[4];
    int out[4];
};

template <int TPW>
__global__ __launch_bounds__(BLOCK) void mlp_infer_kernel(const float* __restrict__ raw, InferArgs a, int ldr,
                                                          float* __restrict__ Y) {
    extern __shared__ __attribute__((aligned(16))) u16 act[];
    u16* src = act;
    u16* dst = act + INFER_ROWS * ldr;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t s0 = (int64_t)blockIdx.x * INFER_ROWS;
    if (threadIdx.x < INFER_ROWS) {
        float f[FEAT];
        sample_features(raw + (s0 + threadIdx.x) * RAW, f);
#pragma unroll
        for (int j = 0; j < FEAT; ++j) src[threadIdx.x * ldr + j] = to_bf16(f[j]);
    }
    __syncthreads();
    for (int l = 0; l < 4; ++l) {
        const int in = a.in[l];
        const int ntiles = (a.out[l] + 31) / 32;
        const u16* __restrict__ w = a.w[l];
        f32x16 acc[TPW];
#pragma unroll
        for (int t = 0; t < TPW; ++t)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
        const int kq = (lane >> 5) * 8;
        for (int k = 0; k < in; k += 16) {
            const u32x4 av = *reinterpret_cast<const u32x4*>(&src[(lane & 31) * ldr + k + kq]);
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int tile = wave + 4 * t;
                if (tile < ntiles) {
                    const u32x4 bv =
                        *reinterpret_cast<const u32x4*>(w + (int64_t)(32 * tile + (lane & 31)) * in + k + kq);
                    acc[t] = mfma32<1>(av, bv, acc[t]);
                }
            }
        }
        const float* __restrict__ bias = a.b[l];
        if (l < 3) {
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int tile = wave + 4 * t;
                if (tile < ntiles) {
                    const int col = 32 * tile + (lane & 31);
                    const float bc = bias[col];
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int row = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                        dst[row * ldr + col] = to_bf16(silu(bc + acc[t][v]));
                    }
                }
            }
            __syncthreads();
            u16* tmp = src;
            src = dst;
            dst = tmp;
        } else if (wave == 0 && (lane & 31) < 2) {
            const int col = lane & 31;
            const float bc = bias[col];
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                Y[(s0 + row) * 2 + col] = bc + acc[0][v];
            }
        }
    }
}

}  // namespace

// ---- host side ----

namespace mlp {

// Layer l: rows x cols weights (rows padded to prow for the last layer).
struct Layer {
    int rows, cols, prow;
    int64_t w_off, b_off;  // offsets in the FP32 parameter vector
};

}  // namespace mlp
}  // namespace sbh

struct sbh_mlp {
    sbh_context* ctx = nullptr;
    int width = 512;
    int layers = 4;
    std::vector<sbh::mlp::Layer> L;
    int64_t nparams = 0;
    sbh::DeviceArray<float> params, grads, m1, m2;
    sbh::DeviceArray<uint16_t> Wb, WTb;  // BF16 weight copies, per layer at wb_off
    std::vector<int64_t> wb_off;
    int64_t step = 0;
    // batch workspace (sized for max_batch samples)
    int max_batch = 0;
    sbh::DeviceArray<uint16_t> X, XT;
    std::vector<sbh::DeviceArray<float>> H;     // pre-activations per hidden layer
    std::vector<sbh::DeviceArray<uint16_t>> A, AT;  // activations per hidden layer
    sbh::DeviceArray<uint16_t> dZ, dZT, dZ2, dZ2T;
    sbh::DeviceArray<float> part, Y;
    sbh::DeviceArray<double> loss_part;
    sbh::DeviceArray<int64_t> idx;
    sbh::DeviceArray<float> raw_ws;  // raw inputs of device-side predictions
    // dataset
    sbh::DeviceArray<float> data_in, data_out;
    int64_t data_count = 0;
};

namespace sbh {
inline sbh_context* ctx_of(const sbh_mlp* n) { return n ? n->ctx : nullptr; }
}  // namespace sbh

namespace sbh {
namespace {

constexpr int kSplitW = 16;  // split K (samples) of the weight-gradient products

void gemm(sbh_context* ctx, const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N, int K, int splits,
          float* C) {
    SBH_REQUIRE(M % 64 == 0 && N % 64 == 0 && K % (16 * splits) == 0, "MLP GEMM shape is not tile aligned");
    dim3 grid(M / 64, N / 64, splits);
    hipLaunchKernelGGL(mlp_gemm_kernel, grid, dim3(BLOCK), 0, ctx->stream, reinterpret_cast<const u16*>(A), lda,
                       reinterpret_cast<const u16*>(B), ldb, M, N, K / splits, C);
    SBH_HIP(hipGetLastError());
}

void ensure_batch(sbh_mlp* net, int M) {
    if (M <= net->max_batch) return;
    const int W = net->width;
    net->X.allocate((size_t)M * FEAT);
    net->XT.allocate((size_t)M * FEAT);
    net->H.clear();
    net->A.clear();
    net->AT.clear();
    for (int l = 0; l < net->layers - 1; ++l) {
        net->H.emplace_back((size_t)M * W);
        net->A.emplace_back((size_t)M * W);
        net->AT.emplace_back((size_t)M * W);
    }
    net->dZ.allocate((size_t)M * W);
    net->dZT.allocate((size_t)M * W);
    net->dZ2.allocate((size_t)M * W);
    net->dZ2T.allocate((size_t)M * W);
    const size_t wpart = (size_t)kSplitW * W * W;
    net->part.allocate(std::max<size_t>((size_t)M * W, wpart));
    net->Y.allocate((size_t)M * 2);
    net->loss_part.allocate(nb((int64_t)M * OUT_PAD));
    net->idx.allocate((size_t)M);
    net->max_batch = M;
}

void refresh_weights(sbh_mlp* net) {
    for (int l = 0; l < net->layers; ++l) {
        const mlp::Layer& ly = net->L[l];
        hipLaunchKernelGGL(mlp_cast_kernel, dim3(nb((int64_t)ly.prow * ly.cols)), dim3(BLOCK), 0, net->ctx->stream,
                           net->params.ptr + ly.w_off, ly.rows, ly.cols, ly.prow,
                           reinterpret_cast<u16*>(net->Wb.ptr + net->wb_off[l]),
                           reinterpret_cast<u16*>(net->WTb.ptr + net->wb_off[l]));
        SBH_HIP(hipGetLastError());
    }
}

// Forward pass over M samples (inputs raw, optional gather idx); with targets
// also the loss and the backward pass into net->grads.  Returns the loss
// partials' block count (0 without targets).
int forward_backward(sbh_mlp* net, const float* raw, const int64_t* idx, int M, const float* targets, bool backward) {
    sbh_context* ctx = net->ctx;
    const int W = net->width;
    // the transposed activations, pre-activations and output gradients only
    // feed the backward pass
    const bool keep = targets != nullptr;
    hipLaunchKernelGGL(mlp_features_kernel, dim3(nb(M)), dim3(BLOCK), 0, ctx->stream, raw, idx, M,
                       reinterpret_cast<u16*>(net->X.ptr), keep ? reinterpret_cast<u16*>(net->XT.ptr) : nullptr);
    SBH_HIP(hipGetLastError());
    const uint16_t* in = net->X.ptr;
    int in_dim = FEAT;
    for (int l = 0; l < net->layers - 1; ++l) {
        const mlp::Layer& ly = net->L[l];
        gemm(ctx, in, in_dim, net->Wb.ptr + net->wb_off[l], in_dim, M, W, in_dim, 1, net->part.ptr);
        hipLaunchKernelGGL(mlp_bias_act_kernel, dim3(nb((int64_t)M * W)), dim3(BLOCK), 0, ctx->stream, net->part.ptr,
                           1, net->params.ptr + ly.b_off, M, W, keep ? net->H[l].ptr : nullptr,
                           reinterpret_cast<u16*>(net->A[l].ptr),
                           keep ? reinterpret_cast<u16*>(net->AT[l].ptr) : nullptr);
        SBH_HIP(hipGetLastError());
        in = net->A[l].ptr;
        in_dim = W;
    }
    const mlp::Layer& lo = net->L[net->layers - 1];
    gemm(ctx, in, W, net->Wb.ptr + net->wb_off[net->layers - 1], W, M, OUT_PAD, W, 1, net->part.ptr);
    const int blocks = (int)nb((int64_t)M * OUT_PAD);
    hipLaunchKernelGGL(mlp_output_kernel, dim3(blocks), dim3(BLOCK), 0, ctx->stream, net->part.ptr, 1,
                       net->params.ptr + lo.b_off, M, targets, idx, net->Y.ptr,
                       keep ? reinterpret_cast<u16*>(net->dZ.ptr) : nullptr,
                       keep ? reinterpret_cast<u16*>(net->dZT.ptr) : nullptr, net->loss_part.ptr);
    SBH_HIP(hipGetLastError());
    if (!targets || !backward) return targets ? blocks : 0;
    SBH_REQUIRE(M % (16 * kSplitW) == 0, "training batch must be a multiple of 256 samples");
    // backward, last layer first; dZ/dZT hold the current layer's gradient, width = its output size
    uint16_t* dz = net->dZ.ptr;
    uint16_t* dzt = net->dZT.ptr;
    uint16_t* dz_next = net->dZ2.ptr;
    uint16_t* dzt_next = net->dZ2T.ptr;
    for (int l = net->layers - 1; l >= 0; --l) {
        const mlp::Layer& ly = net->L[l];
        const int out_dim = ly.prow;
        const int kin = ly.cols;
        const uint16_t* prevT = l == 0 ? net->XT.ptr : net->AT[l - 1].ptr;
        // dW (out_dim x kin) = dZ^T prev ; rows >= ly.rows are padding
        gemm(ctx, dzt, M, prevT, M, out_dim, kin, M, kSplitW, net->part.ptr);
        // the padded rows are never read: rows x kin is a prefix of each out_dim x kin slab
        hipLaunchKernelGGL(mlp_sum_kernel, dim3(nb((int64_t)ly.rows * kin)), dim3(BLOCK), 0, ctx->stream,
                           net->part.ptr, kSplitW, (int64_t)out_dim * kin, (int64_t)ly.rows * kin,
                           net->grads.ptr + ly.w_off);
        SBH_HIP(hipGetLastError());
        hipLaunchKernelGGL(mlp_bias_grad_kernel, dim3(ly.rows), dim3(BLOCK), 0, ctx->stream,
                           reinterpret_cast<const u16*>(dzt), M, net->grads.ptr + ly.b_off);
        SBH_HIP(hipGetLastError());
        if (l == 0) break;
        // dA_prev (M x kin) = dZ W  (W^T is kin x out_dim)
        gemm(ctx, dz, out_dim, net->WTb.ptr + net->wb_off[l], out_dim, M, kin, out_dim, 1, net->part.ptr);
        hipLaunchKernelGGL(mlp_act_grad_kernel, dim3(nb((int64_t)M * kin)), dim3(BLOCK), 0, ctx->stream,
                           net->part.ptr, 1, net->H[l - 1].ptr, M, kin, reinterpret_cast<u16*>(dz_next),
                           reinterpret_cast<u16*>(dzt_next));
        SBH_HIP(hipGetLastError());
        std::swap(dz, dz_next);
        std::swap(dzt, dzt_next);
    }
    return blocks;
}

// Predictions of Mp samples (a multiple of 256) into net->Y: the fused
// kernel when the width allows (every layer's tiles fit the waves' tile
// registers), else the layer-by-layer forward pass.  Same bits either way.
void infer(sbh_mlp* net, const float* raw, int Mp) {
    const int W = net->width;
    if (W > 4 * INFER_MAX_TILES * 32) {
        ensure_batch(net, Mp);
        forward_backward(net, raw, nullptr, Mp, nullptr, false);
        return;
    }
    if ((int64_t)net->Y.count < (int64_t)Mp * 2) net->Y.allocate((size_t)Mp * 2);
    InferArgs a;
    for (int l = 0; l < 4; ++l) {
        const mlp::Layer& ly = net->L[l];
        a.w[l] = reinterpret_cast<const u16*>(net->Wb.ptr + net->wb_off[l]);
        a.b[l] = net->params.ptr + ly.b_off;
        a.in[l] = ly.cols;
        a.out[l] = ly.rows;
    }
    const int ldr = W + 8;
    const size_t lds = (size_t)2 * INFER_ROWS * ldr * sizeof(uint16_t);
    const int tiles = (W / 32 + 3) / 4;  // 32-column tiles per wave in the hidden layers
    auto launch = [&](auto kernel) {
        SBH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        hipLaunchKernelGGL(kernel, dim3((unsigned)(Mp / INFER_ROWS)), dim3(BLOCK), lds, net->ctx->stream, raw, a, ldr,
                           net->Y.ptr);
    };
    if (tiles <= 1)
        launch(mlp_infer_kernel<1>);
    else if (tiles <= 2)
        launch(mlp_infer_kernel<2>);
    else if (tiles <= 4)
        launch(mlp_infer_kernel<4>);
    else
        launch(mlp_infer_kernel<8>);
    SBH_HIP(hipGetLastError());
}

double read_loss(sbh_mlp* net, int blocks) {
    std::vector<double> h(blocks);
    SBH_HIP(hipMemcpyAsync(h.data(), net->loss_part.ptr, blocks * sizeof(double), hipMemcpyDeviceToHost,
                           net->ctx->stream));
    SBH_HIP(hipStreamSynchronize(net->ctx->stream));
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}

inline void use_device(sbh_context* ctx) { SBH_HIP(hipSetDevice(ctx->device)); }

}  // namespace
}  // namespace sbh

namespace sbh {

sbh_context* mlp_context(const sbh_mlp* net) { return net ? net->ctx : nullptr; }

float* mlp_raw_workspace(sbh_mlp* net, int64_t M) {
    SBH_REQUIRE(M > 0 && M <= (1 << 24), "sample count out of range");
    const int64_t Mp = (M + 255) / 256 * 256;
    if (net->raw_ws.count < (size_t)Mp * RAW) net->raw_ws.allocate((size_t)Mp * RAW);
    if (Mp > M)
        SBH_HIP(hipMemsetAsync(net->raw_ws.ptr + M * RAW, 0, (size_t)(Mp - M) * RAW * sizeof(float),
                               net->ctx->stream));
    return net->raw_ws.ptr;
}

const float* mlp_forward_device(sbh_mlp* net, const float* raw, int64_t M) {
    SBH_REQUIRE(M > 0 && M <= (1 << 24), "sample count out of range");
    const int Mp = (int)((M + 255) / 256 * 256);
    infer(net, raw, Mp);
    return net->Y.ptr;
}

}  // namespace sbh

using namespace sbh;

extern "C" {

sbh_status sbh_mlp_create(sbh_context* ctx, int32_t width, uint64_t seed, sbh_mlp** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && out, "ctx/out is null");
        SBH_REQUIRE(width >= 64 && width <= 4096 && width % 64 == 0, "width must be a multiple of 64 in 64..4096");
        use_device(ctx);
        std::unique_ptr<sbh_mlp> net(new sbh_mlp());
        net->ctx = ctx;
        net->width = width;
        const int dims[5] = {FEAT, width, width, width, 2};
        int64_t off = 0, woff = 0;
        for (int l = 0; l < 4; ++l) {
            mlp::Layer ly;
            ly.rows = dims[l + 1];
            ly.cols = dims[l];
            ly.prow = l == 3 ? OUT_PAD : dims[l + 1];
            ly.w_off = off;
            off += (int64_t)ly.rows * ly.cols;
            ly.b_off = off;
            off += ly.rows;
            net->L.push_back(ly);
            net->wb_off.push_back(woff);
            woff += (int64_t)ly.prow * ly.cols;
        }
        net->nparams = off;
        // He-uniform weights from a seeded generator, zero biases
        std::vector<float> p((size_t)off, 0.f);
        std::mt19937_64 gen(seed);
        for (const auto& ly : net->L) {
            const double bound = std::sqrt(6.0 / ly.cols) * (ly.rows == 2 ? 0.1 : 1.0);
            std::uniform_real_distribution<double> U(-bound, bound);
            for (int64_t i = 0; i < (int64_t)ly.rows * ly.cols; ++i) p[ly.w_off + i] = (float)U(gen);
        }
        net->params.allocate((size_t)off);
        net->grads.allocate((size_t)off);
        net->m1.allocate((size_t)off);
        net->m2.allocate((size_t)off);
        copy_on_stream(net->ctx, net->params.ptr, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice);
        zero_on_stream(net->ctx, net->grads.ptr, net->grads.bytes());
        zero_on_stream(net->ctx, net->m1.ptr, net->m1.bytes());
        zero_on_stream(net->ctx, net->m2.ptr, net->m2.bytes());
        net->Wb.allocate((size_t)woff);
        net->WTb.allocate((size_t)woff);
        refresh_weights(net.get());
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        context_retain(ctx);
        *out = net.release();
    });
}

void sbh_mlp_destroy(sbh_mlp* net) {
    if (!net) return;
    sbh_context* ctx = net->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete net;
    context_release(ctx);
}

int64_t sbh_mlp_num_params(const sbh_mlp* net) { return net ? net->nparams : 0; }

sbh_status sbh_mlp_get_params(sbh_mlp* net, float* out) {
    return guard_on(ctx_of(net), [&] {
        SBH_REQUIRE(net && out, "null argument");
        use_device(net->ctx);
        SBH_HIP(hipStreamSynchronize(net->ctx->stream));
        copy_on_stream(net->ctx, out, net->params.ptr, net->params.bytes(), hipMemcpyDeviceToHost);
    });
}

sbh_status sbh_mlp_set_params(sbh_mlp* net, const float* in) {
    return guard_on(ctx_of(net), [&] {
        SBH_REQUIRE(net && in, "null argument");
        use_device(net->ctx);
        SBH_HIP(hipStreamSynchronize(net->ctx->stream));
        copy_on_stream(net->ctx, net->params.ptr, in, net->params.bytes(), hipMemcpyHostToDevice);
        refresh_weights(net);
        SBH_HIP(hipStreamSynchronize(net->ctx->stream));
    });
}

sbh_status sbh_mlp_loss_and_gradients(sbh_mlp* net, const float* inputs, const float* targets, int64_t M,
                                      double* loss, float* grads) {
    return guard_on(ctx_of(net), [&] {
        SBH_REQUIRE(net && inputs && targets, "null argument");
        SBH_REQUIRE(M > 0 && M % 256 == 0 && M <= (1 << 24), "M must be a positive multiple of 256");
        use_device(net->ctx);
        ensure_batch(net, (int)M);
        DeviceArray<float> din((size_t)M * RAW), dout((size_t)M * 2);
        copy_on_stream(net->ctx, din.ptr, inputs, din.bytes(), hipMemcpyHostToDevice);
        copy_on_stream(net->ctx, dout.ptr, targets, dout.bytes(), hipMemcpyHostToDevice);
        SBH_HIP(hipMemsetAsync(net->grads.ptr, 0, net->grads.bytes(), net->ctx->stream));
        const int blocks = forward_backward(net, din.ptr, nullptr, (int)M, dout.ptr, true);
        const double l = read_loss(net, blocks) / (double)M;
        if (loss) *loss = l;
        if (grads) copy_on_stream(net->ctx, grads, net->grads.ptr, net->grads.bytes(), hipMemcpyDeviceToHost);
    });
}

sbh_status sbh_mlp_set_dataset(sbh_mlp* net, const float* inputs, const float* targets, int64_t count) {
    return guard_on(ctx_of(net), [&] {
        SBH_REQUIRE(net && inputs && targets && count > 0, "null argument or empty dataset");
        use_device(net->ctx);
        SBH_HIP(hipStreamSynchronize(net->ctx->stream));
        net->data_in.allocate((size_t)count * RAW);
        net->data_out.allocate((size_t)count * 2);
        copy_on_stream(net->ctx, net->data_in.ptr, inputs, net->data_in.bytes(), hipMemcpyHostToDevice);
        copy_on_stream(net->ctx, net->data_out.ptr, targets, net->data_out.bytes(), hipMemcpyHostToDevice);
        net->data_count = count;
    });
}

sbh_status sbh_mlp_train(sbh_mlp* net, int64_t steps, int64_t batch, double lr_max, double lr_min, uint64_t seed,
                         double* losses) {
    return guard_on(ctx_of(net), [&] {
        SBH_REQUIRE(net, "net is null");
        SBH_REQUIRE(net->data_count > 0, "sbh_mlp_set_dataset has not been called");
        SBH_REQUIRE(batch > 0 && batch % 256 == 0 && batch <= (1 << 20), "batch must be a multiple of 256");
        SBH_REQUIRE(steps >= 0, "steps must be >= 0");
        use_device(net->ctx);
        sbh_context* ctx = net->ctx;
        ensure_batch(net, (int)batch);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int64_t s = 0; s < steps; ++s) {
            hipLaunchKernelGGL(mlp_sample_kernel, dim3(nb(batch)), dim3(BLOCK), 0, ctx->stream, net->idx.ptr,
                               (int)batch, seed * 0x100000001B3ull + (uint64_t)net->step, net->data_count);
            SBH_HIP(hipGetLastError());
            const int blocks =
                forward_backward(net, net->data_in.ptr, net->idx.ptr, (int)batch, net->data_out.ptr, true);
            if (losses) losses[s] = read_loss(net, blocks) / (double)batch;
            net->step += 1;
            const double frac = steps > 1 ? (double)s / (double)(steps - 1) : 0.0;
            const double lr = lr_min + 0.5 * (lr_max - lr_min) * (1.0 + std::cos(M_PI * frac));
            const float c1 = 1.f - std::pow(b1, (float)net->step);
            const float c2 = 1.f - std::pow(b2, (float)net->step);
            hipLaunchKernelGGL(mlp_adam_kernel, dim3(nb(net->nparams)), dim3(BLOCK), 0, ctx->stream, net->params.ptr,
                               net->grads.ptr, net->m1.ptr, net->m2.ptr, net->nparams, (float)lr, b1, b2, eps, c1, c2);
            SBH_HIP(hipGetLastError());
            refresh_weights(net);
        }
        SBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

sbh_status sbh_debug_mlp_predict_layered(sbh_mlp* net, const float* inputs, int64_t M, float* out) {
    return guard_on(ctx_of(net), [&] {
        SBH_REQUIRE(net && inputs && out, "null argument");
        SBH_REQUIRE(M > 0 && M <= (1 << 24), "M out of range");
        use_device(net->ctx);
        const int Mp = (int)((M + 255) / 256 * 256);
        ensure_batch(net, Mp);
        DeviceArray<float> din((size_t)Mp * RAW);
        SBH_HIP(hipMemsetAsync(din.ptr, 0, din.bytes(), net->ctx->stream));
        SBH_HIP(hipMemcpyAsync(din.ptr, inputs, (size_t)M * RAW * sizeof(float), hipMemcpyHostToDevice,
                               net->ctx->stream));
        forward_backward(net, din.ptr, nullptr, Mp, nullptr, false);
        SBH_HIP(hipMemcpyAsync(out, net->Y.ptr, (size_t)M * 2 * sizeof(float), hipMemcpyDeviceToHost,
                               net->ctx->stream));
        SBH_HIP(hipStreamSynchronize(net->ctx->stream));
    });
}

sbh_status sbh_mlp_predict(sbh_mlp* net, const float* inputs, int64_t M, float* out, int memory) {
    return guard_on(ctx_of(net), [&] {
        SBH_REQUIRE(net && inputs && out, "null argument");
        SBH_REQUIRE(M > 0 && M <= (1 << 24), "M out of range");
        use_device(net->ctx);
        const int Mp = (int)((M + 255) / 256 * 256);
        DeviceArray<float> din((size_t)Mp * RAW);
        SBH_HIP(hipMemsetAsync(din.ptr, 0, din.bytes(), net->ctx->stream));
        SBH_HIP(hipMemcpyAsync(din.ptr, inputs, (size_t)M * RAW * sizeof(float),
                               memory == SBH_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               net->ctx->stream));
        infer(net, din.ptr, Mp);
        SBH_HIP(hipMemcpyAsync(out, net->Y.ptr, (size_t)M * 2 * sizeof(float),
                               memory == SBH_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                               net->ctx->stream));
        SBH_HIP(hipStreamSynchronize(net->ctx->stream));
    });
}

}  // extern "C"
