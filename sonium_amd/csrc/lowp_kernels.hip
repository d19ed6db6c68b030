// Low-precision storage, block products and batched Krylov vector kernels.
//
// Complex FP16/BF16 storage with power-of-two row scaling (the reference's
// explicit-scale storage precedent: crates/sonium-compute/src/scaled_storage.rs:1-127)
// and its products with FP32 accumulation, plus the FP64 block product
// (DenseMatrixOperator::apply_block → DenseMatrix::matmul_into, bem_solver.rs:6567-6574,
// dense.rs:450-488) and the column-batched vector kernels that let k
// right-hand sides share one block product per Krylov step.
//
// Layout.  Row i of the low-precision matrix stores A_ij * 2^(T - e_i) where
// 2^e_i <= max_j max(|Re A_ij|, |Im A_ij|) < 2^(e_i+1) and T = 14 (FP16: the
// largest scaled entry stays below 2^15) or 0 (BF16 has the FP32 exponent
// range); row_inv[i] = 2^(e_i - T) undoes it exactly.  An entry is two 16-bit
// halves {re, im}, so a row is the real row [re0 im0 re1 im1 ...] of length
// 2*ld (ld >= n complex entries, a multiple of 512, zero padded).  The rows
// are stored in MFMA fragment order rather than row-major: the matrix is cut
// into tiles of 32 rows x 16 halves, and tile (r, s) is 1 KB at fragment index
// r * (ld / 8) + s, holding for lane l = 0..63 the 8 halves of row 32 r + l % 32
// starting at half 16 s + 8 (l / 32) -- exactly the A operand of
// v_mfma_f32_32x32x16_{f16,bf16}.  A wave therefore reads its operand for one
// MFMA step as a single contiguous 1 KB load; rows beyond n are zero.
//
// Complex GEMM on MFMA.  With A_r the N x 2N real view of the complex
// matrix, Y = A X is the real product A_r * X_r where X_r (2N x 2k) holds, for
// complex row j, the rows [xr_j, xi_j] and [-xi_j, xr_j]; so output column
// c < k is Re Y[:, c] and column k + c is Im Y[:, c].  Only the first half is
// stored, transposed (XT: k x 2 ld halves, row c = [xr_0, -xi_0, xr_1, ...]);
// the second half's operand [xi_j, xr_j] is the first's with the two halves of
// every 32-bit word swapped and the new low half negated, made in registers.
// Each lane's MFMA B operand (8 consecutive k for one column) is one 16-byte
// LDS read.  Each column of X is scaled by a power of two as well.

#include "internal.hpp"

namespace sbh {
namespace {

#include "lowp_common.hpp"

__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Complex vectors are double2 (FP64) or float2 (FP32); ld2/st2 move them to
// and from FP64 arithmetic (st2 rounds to nearest).
__device__ __forceinline__ double2 ld2(double2 v) { return v; }
__device__ __forceinline__ double2 ld2(float2 v) { return make_double2(v.x, v.y); }
template <typename T>
__device__ __forceinline__ T st2(double2 v);
template <>
__device__ __forceinline__ double2 st2<double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 st2<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

// Row scales: one wave per row finds max |re|, |im| and writes row_inv.  The
// matrix is read once here, so a lane keeps RS_LOADS non-temporal 16-byte
// loads in flight.
typedef double rs_v2d __attribute__((ext_vector_type(2)));
constexpr int RS_LOADS = 16;

__global__ __launch_bounds__(BLOCK) void lp_rowscale_kernel(const double2* __restrict__ A, int64_t lda, int64_t n,
                                                            int top, double* __restrict__ row_inv) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    if (i >= n) return;
    const rs_v2d* __restrict__ row = reinterpret_cast<const rs_v2d*>(A + i * lda);
    double m = 0.0;
    int64_t j = lane;
    for (; j + 64 * (RS_LOADS - 1) < n; j += 64 * RS_LOADS) {
        rs_v2d v[RS_LOADS];
#pragma unroll
        for (int u = 0; u < RS_LOADS; ++u) v[u] = __builtin_nontemporal_load(row + j + 64 * u);
#pragma unroll
        for (int u = 0; u < RS_LOADS; ++u) m = fmax(m, fmax(fabs(v[u].x), fabs(v[u].y)));
    }
    for (; j < n; j += 64) {
        const rs_v2d v = row[j];
        m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
    }
    m = wave_max(m);
    int e = 0;
    if (m > 0.0 && isfinite(m)) e = ilogb(m);
    if (lane == 0) row_inv[i] = ldexp(1.0, e - top);
}

// Row scales from the assembly's row maxima: one wave per row takes the
// maximum of its stride partials (the same value as the pass over the row:
// the maximum does not depend on the order) and writes row_inv as above.
__global__ __launch_bounds__(BLOCK) void lp_rowscale_from_max_kernel(const double* __restrict__ rowmax,
                                                                     int64_t stride, int64_t n, int top,
                                                                     double* __restrict__ row_inv) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    if (i >= n) return;
    double m = 0.0;
    for (int64_t c = lane; c < stride; c += 64) m = fmax(m, rowmax[i * stride + c]);
    m = wave_max(m);
    int e = 0;
    if (m > 0.0 && isfinite(m)) e = ilogb(m);
    if (lane == 0) row_inv[i] = ldexp(1.0, e - top);
}

// Packing: a block converts one 32-row tile x 64 complex entries (8 fragments).
// Each wave-wide load reads 64 consecutive entries (1 KB) of one row; the
// converted {re, im} words go through LDS (rows padded to 68 words so a
// fragment read hits 16 distinct bank groups per 16 lanes) and leave as 8
// contiguous 1 KB fragments.  Values are scaled by 1/row_inv (a power of two,
// exact) and rounded through FP32 to the format.
constexpr int LP_PACK_COLS = 64;               // complex entries per block (8 fragments)
constexpr int LP_PACK_ROW = LP_PACK_COLS + 4;  // padded LDS row, in 32-bit words

template <int FMT>
__global__ __launch_bounds__(BLOCK) void lp_pack_kernel(const double2* __restrict__ A, int64_t lda, int64_t n,
                                                        int64_t ld, const double* __restrict__ row_inv,
                                                        u16* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) unsigned sh[32 * LP_PACK_ROW];
    __shared__ double scale_sh[32];
    const int64_t groups = ld / LP_PACK_COLS;
    const int64_t rt = blockIdx.x / groups;
    const int64_t g = blockIdx.x % groups;
    const int64_t j_base = g * LP_PACK_COLS;
    if (threadIdx.x < 32) {
        const int64_t row = rt * 32 + threadIdx.x;
        scale_sh[threadIdx.x] = row < n ? 1.0 / row_inv[row] : 0.0;  // a power of two, exact
    }
    __syncthreads();
    const rs_v2d* __restrict__ a = reinterpret_cast<const rs_v2d*>(A);
#pragma unroll
    for (int i = 0; i < 32 * LP_PACK_COLS / BLOCK; ++i) {
        const int e = i * BLOCK + threadIdx.x;
        const int r = e / LP_PACK_COLS, col = e % LP_PACK_COLS;
        const int64_t row = rt * 32 + r, j = j_base + col;
        unsigned w = 0;
        if (row < n && j < n) {
            const rs_v2d v = __builtin_nontemporal_load(a + row * lda + j);  // read once here
            const double scale = scale_sh[r];
            w = (unsigned)Half<FMT>::from_float((float)(v.x * scale)) |
                ((unsigned)Half<FMT>::from_float((float)(v.y * scale)) << 16);
        }
        sh[r * LP_PACK_ROW + col] = w;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t steps = ld / 8;
    u32x4* __restrict__ dst = reinterpret_cast<u32x4*>(out) + (rt * steps + g * 8) * 64 + lane;
#pragma unroll
    for (int f = 0; f < 8 / (BLOCK / 64); ++f) {
        const int q = wave * (8 / (BLOCK / 64)) + f;  // fragment within the block's 8
        const u32x4 v = *reinterpret_cast<const u32x4*>(&sh[(lane & 31) * LP_PACK_ROW + q * 8 + 4 * (lane >> 5)]);
        dst[(int64_t)q * 64] = v;
    }
}

// Column max |re|,|im| of an n x k block (atomic max on the bit pattern of a
// non-negative float, which orders like the value).
// Each block reduces rows [b*R, (b+1)*R) in LDS first, then one atomic per column.
template <typename T>
__global__ __launch_bounds__(BLOCK) void colmax_kernel(const T* __restrict__ X, int64_t n, int k,
                                                       int64_t rows_per_block, unsigned* __restrict__ out) {
    __shared__ float sh[BLOCK];
    const int S = BLOCK / k;
    const int c = threadIdx.x % k;
    const int sub = threadIdx.x / k;
    const int64_t b = (int64_t)blockIdx.x * rows_per_block;
    const int64_t e = min(n, b + rows_per_block);
    float m = 0.f;
    if (sub < S) {
#pragma unroll 8
        for (int64_t r = b + sub; r < e; r += S) {
            double2 v = ld2(X[r * k + c]);
            m = fmaxf(m, (float)fmax(fabs(v.x), fabs(v.y)));
        }
    }
    sh[threadIdx.x] = m;
    // the S partials of a column by a tree (a maximum is exact in any order)
    for (int width = S; width > 1;) {
        const int half = (width + 1) / 2;
        __syncthreads();
        if (sub < width - half) sh[sub * k + c] = fmaxf(sh[sub * k + c], sh[(sub + half) * k + c]);
        width = half;
    }
    __syncthreads();
    if (threadIdx.x < k) atomicMax(out + threadIdx.x, __float_as_uint(sh[threadIdx.x]));
}

__global__ void fill_u32_kernel(unsigned* p, int n, unsigned v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// X (n x k complex) -> XT (k x 2*ld halves: [xr, -xi] pairs), scaled by 2^(TOP - e_c) per column.
template <int FMT, typename T>
__global__ void x_to_lp_kernel(const T* __restrict__ X, int64_t n, int k, int64_t ld,
                               const unsigned* __restrict__ colmax, u16* __restrict__ XT,
                               double* __restrict__ col_inv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int64_t j = t / k;
    const int c = (int)(t - j * k);
    float m = __uint_as_float(colmax[c]);
    int e = 0;
    if (m > 0.f && isfinite(m)) e = ilogb(m);
    double s = ldexp(1.0, Half<FMT>::TOP - e);
    if (j == 0) col_inv[c] = ldexp(1.0, e - Half<FMT>::TOP);
    double2 v = ld2(X[t]);
    const u16 xr = Half<FMT>::from_float((float)(v.x * s));
    const u16 nxi = Half<FMT>::from_float((float)(-v.y * s));
    u16* re_row = XT + (int64_t)c * 2 * ld;
    re_row[2 * j] = xr;
    re_row[2 * j + 1] = nxi;
}

// Split-K MFMA product with FP32 accumulation.  Block = LP_WAVES waves, each
// owning one 32-row tile, and one of LP_SPLIT K slabs; slab = blockIdx %
// LP_SPLIT, so on the 8-XCD part every XCD streams one slab of the block
// operand, which stays in that XCD's L2.  A wave loads its A fragments for a
// stage of LP_BK halves (8 x 1 KB, contiguous) straight into registers, one
// stage ahead; the block's XT columns for the stage go through LDS (rows padded
// to 136 halves), shared by all waves.  Per 16-half step a wave does NTR MFMAs
// for the real-part columns and NTR for the imaginary-part columns (operand
// derived in registers).  Partial sums go to Ypart[slab][row][col] (real part
// at col, imaginary part at NTR * 32 + col) and lp_reduce_kernel adds the
// slabs in a fixed order.
constexpr int LP_SPLIT = 8;
constexpr int LP_BK = 128;
constexpr int LP_LDS_ROW = LP_BK + 8;
constexpr int LP_WAVES = 8;
constexpr int LP_THREADS = LP_WAVES * 64;

__device__ __forceinline__ u32x4 imag_operand(u32x4 b) {  // [xr, -xi] -> [xi, xr] per 32-bit word
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = ((b[i] << 16) | (b[i] >> 16)) ^ 0x8000u;
    return r;
}

template <int FMT, int NTR>
__global__ __launch_bounds__(LP_THREADS) void lp_gemm_kernel(const u16* __restrict__ A, int64_t ld, int64_t n,
                                                             const u16* __restrict__ XT, float* __restrict__ Ypart) {
    constexpr int BN = 2 * NTR * 32;
    __shared__ __attribute__((aligned(16))) u16 Bs[NTR * 32 * LP_LDS_ROW];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int split = blockIdx.x % LP_SPLIT;
    const int64_t n_tiles = (n + 31) / 32;
    const int64_t rt = (int64_t)(blockIdx.x / LP_SPLIT) * LP_WAVES + wave;
    const bool live = rt < n_tiles;
    const int64_t slab = (2 * ld) / LP_SPLIT;  // a multiple of LP_BK (ld is a multiple of 512)
    const int64_t k_begin = split * slab;
    const int64_t k_end = k_begin + slab;
    const u32x4* __restrict__ afrag = reinterpret_cast<const u32x4*>(A) + (live ? rt : 0) * (ld / 8) * 64 + lane;
    constexpr int AS = LP_BK / 16;                                // fragments per stage
    constexpr int BR = (NTR * 32 * (LP_BK / 8)) / LP_THREADS;     // 16-byte XT pieces per thread per stage
    u32x4 ra[AS], rb[BR];
    auto load = [&](int64_t kc) {
        if (live) {
            const int64_t s0 = kc / 16;
#pragma unroll
            for (int i = 0; i < AS; ++i) ra[i] = __builtin_nontemporal_load(afrag + (s0 + i) * 64);
        }
#pragma unroll
        for (int i = 0; i < BR; ++i) {
            const int e = threadIdx.x + i * LP_THREADS;
            rb[i] = *reinterpret_cast<const u32x4*>(XT + (int64_t)(e >> 4) * 2 * ld + kc + (e & 15) * 8);
        }
    };
    f32x16 acc[2 * NTR];
#pragma unroll
    for (int t = 0; t < 2 * NTR; ++t)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
    const int kq = (lane >> 5) * 8;
    load(k_begin);
    for (int64_t kc = k_begin; kc < k_end; kc += LP_BK) {
        u32x4 ca[AS];
#pragma unroll
        for (int i = 0; i < AS; ++i) ca[i] = ra[i];
#pragma unroll
        for (int i = 0; i < BR; ++i) {
            const int e = threadIdx.x + i * LP_THREADS;
            *reinterpret_cast<u32x4*>(&Bs[(e >> 4) * LP_LDS_ROW + (e & 15) * 8]) = rb[i];
        }
        __syncthreads();
        if (kc + LP_BK < k_end) load(kc + LP_BK);
        if (live) {
#pragma unroll
            for (int s = 0; s < AS; ++s) {
#pragma unroll
                for (int t = 0; t < NTR; ++t) {
                    const u32x4 b =
                        *reinterpret_cast<const u32x4*>(&Bs[(t * 32 + (lane & 31)) * LP_LDS_ROW + s * 16 + kq]);
                    acc[t] = mfma32<FMT>(ca[s], b, acc[t]);
                    acc[NTR + t] = mfma32<FMT>(ca[s], imag_operand(b), acc[NTR + t]);
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
    // C layout of v_mfma_f32_32x32x*: column = lane % 32,
    // row = 8*(v/4) + 4*(lane/32) + (v%4) for accumulator register v.
    const int64_t row_base = rt * 32;
#pragma unroll
    for (int t = 0; t < 2 * NTR; ++t) {
        const int col = t * 32 + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t row = row_base + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (row < n) Ypart[((int64_t)split * n + row) * BN + col] = acc[t][v];
        }
    }
}

// One-column product (k = 1) on v_dot2_f32_{f16,bf16} instead of 32-column
// MFMA tiles, 31 of whose columns would be discarded.  A wave owns one 32-row
// tile and one of LP_SPLIT K slabs (slab = blockIdx % LP_SPLIT, as in
// lp_gemm_kernel).  Lane l streams the 16-byte pieces of row 32 r + l % 32 that
// the fragment layout gives it and reads the 16 bytes of XT row 0 at the same
// half offsets, so per complex entry Re += dot2(a, [xr, -xi]) and
// Im += dot2(a, [xi, xr]), in FP32 with four partial sums each.  Lanes l and
// l + 32 hold the two halves of a row's slab sum; Ypart[slab][row] = (re, im),
// which lp_reduce_kernel reads with bn = 2.
constexpr int LV_WAVES = 4;
constexpr int LV_UNROLL = 8;

__device__ __forceinline__ unsigned imag_word(unsigned b) {  // [xr, -xi] -> [xi, xr]
    return ((b << 16) | (b >> 16)) ^ 0x8000u;
}

template <int FMT>
__global__ __launch_bounds__(LV_WAVES * 64) void lp_gemv_kernel(const u16* __restrict__ A, int64_t ld, int64_t n,
                                                                const u16* __restrict__ XT,
                                                                float* __restrict__ Ypart) {
    const int lane = threadIdx.x & 63;
    const int split = blockIdx.x % LP_SPLIT;
    const int64_t rt = (int64_t)(blockIdx.x / LP_SPLIT) * LV_WAVES + (threadIdx.x >> 6);
    if (rt >= (n + 31) / 32) return;  // no barrier below
    const int64_t fslab = ld / (8 * LP_SPLIT);  // fragments per slab (ld / 8 per tile), a multiple of LV_UNROLL
    const int64_t f0 = split * fslab;
    const u32x4* __restrict__ a = reinterpret_cast<const u32x4*>(A) + (rt * (ld / 8) + f0) * 64 + lane;
    const u32x4* __restrict__ x = reinterpret_cast<const u32x4*>(XT) + 2 * f0 + (lane >> 5);
    float re[4] = {0.f, 0.f, 0.f, 0.f}, im[4] = {0.f, 0.f, 0.f, 0.f};
    for (int64_t s = 0; s < fslab; s += LV_UNROLL) {
        u32x4 av[LV_UNROLL], xv[LV_UNROLL];
#pragma unroll
        for (int u = 0; u < LV_UNROLL; ++u) {
            av[u] = __builtin_nontemporal_load(a + (s + u) * 64);
            xv[u] = x[2 * (s + u)];
        }
#pragma unroll
        for (int u = 0; u < LV_UNROLL; ++u)
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                re[w] = dot2<FMT>(av[u][w], xv[u][w], re[w]);
                im[w] = dot2<FMT>(av[u][w], imag_word(xv[u][w]), im[w]);
            }
    }
    float sr = (re[0] + re[1]) + (re[2] + re[3]);
    float si = (im[0] + im[1]) + (im[2] + im[3]);
    sr += __shfl_xor(sr, 32, 64);
    si += __shfl_xor(si, 32, 64);
    const int64_t row = rt * 32 + lane;
    if (lane < 32 && row < n) {
        Ypart[((int64_t)split * n + row) * 2] = sr;
        Ypart[((int64_t)split * n + row) * 2 + 1] = si;
    }
}

// Y[row][c] = row_inv[row] * col_inv[c] * sum_slab (Ypart[..][row][c] + i Ypart[..][row][bn/2 + c])
template <typename T>
__global__ void lp_reduce_kernel(const float* __restrict__ Ypart, int64_t n, int k, int bn,
                                 const double* __restrict__ row_inv, const double* __restrict__ col_inv,
                                 T* __restrict__ Y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int64_t row = t / k;
    const int c = (int)(t - row * k);
    double re = 0.0, im = 0.0;
    for (int s = 0; s < LP_SPLIT; ++s) {
        const float* p = Ypart + ((int64_t)s * n + row) * bn;
        re += (double)p[c];
        im += (double)p[bn / 2 + c];
    }
    const double sc = row_inv[row] * col_inv[c];
    Y[t] = st2<T>(make_double2(re * sc, im * sc));
}

// FP64 complex block product Y (n x k) = A (n x n, lda) X (n x k), k <= 64 per launch (column
// offset handled by the caller).  Tile 128 rows x 64 columns; each thread 8 rows x 4 columns.
// FP64 complex block product Y = A X (k <= 64 columns per launch) on
// v_mfma_f64_16x16x4_f64, in the real formulation used by the low-precision
// GEMM: the complex row-major A (n x lda) is a real n x 2n matrix, and the
// real operand Xt (2n x 2kc) has rows
//   Xt[2j]   = (Re X[j,0], Im X[j,0], Re X[j,1], ...)     (X's row j as stored)
//   Xt[2j+1] = (-Im X[j,0], Re X[j,0], -Im X[j,1], ...)
// so that Y[r] = A~[r] Xt comes out in complex interleaved order. Xt is never
// stored: its rows are formed while X is copied into LDS.
//
// Block: 4 waves over BM = 128 rows x 16*NT real columns; wave w owns rows
// 32w..32w+31 (2 x NT tiles of 16 x 16). Stages of 8 complex (16 real) columns
// of A go through LDS, transposed so that MFMA operand reads are contiguous.
// The K range is split in ZG_SPLIT slices (slice = blockIdx % ZG_SPLIT, one per
// XCD); partials go to Ypart[slice][row][16*NT] and zgemm_reduce_kernel sums
// them in slice order.
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int ZG_SPLIT = 8;
constexpr int ZG_BM = 128;
constexpr int ZG_BKC = 16;           // complex columns of A per stage
constexpr int ZG_AROW = ZG_BM + 2;   // LDS row of the transposed A stage (doubles)

template <int NT>
__global__ __launch_bounds__(BLOCK) void zgemm_f64_kernel(const double2* __restrict__ A, int64_t lda, int64_t n,
                                                          const double2* __restrict__ X, int64_t ldx, int kc,
                                                          int64_t kslice, double* __restrict__ Ypart) {
    constexpr int NR = 16 * NT;  // real output columns
    __shared__ double As[2 * ZG_BKC][ZG_AROW];
    __shared__ __attribute__((aligned(16))) double Xs[2 * ZG_BKC][NR];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int split = blockIdx.x % ZG_SPLIT;
    const int64_t row0 = (int64_t)(blockIdx.x / ZG_SPLIT) * ZG_BM;
    const int64_t kbeg = (int64_t)split * kslice;
    const int64_t kend = min(n, kbeg + kslice);
    f64x4 acc[2][NT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
    // Stage s+1 is loaded into registers while stage s is multiplied.
    constexpr int AQ = (ZG_BM * ZG_BKC) / BLOCK;                 // A elements per thread per stage
    constexpr int XQ = (ZG_BKC * 8 * NT + BLOCK - 1) / BLOCK;    // X elements per thread per stage
    double2 pa[AQ], px[XQ];
    auto gload = [&](int64_t k0) {
        // A stage: 128 rows x 8 complex; 8 consecutive threads read one row's 128 B
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = t + q * BLOCK;
            const int r = e / ZG_BKC, cc = e % ZG_BKC;
            const int64_t gr = row0 + r, gk = k0 + cc;
            pa[q] = (gr < n && gk < kend) ? A[gr * lda + gk] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < XQ; ++q) {
            const int e = t + q * BLOCK;
            const int j = e / (8 * NT), c = e % (8 * NT);
            const int64_t gk = k0 + j;
            px[q] = (e < ZG_BKC * 8 * NT && gk < kend && c < kc) ? X[gk * ldx + c] : make_double2(0.0, 0.0);
        }
    };
    gload(kbeg);
    for (int64_t k0 = kbeg; k0 < kend; k0 += ZG_BKC) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = t + q * BLOCK;
            const int r = e / ZG_BKC, cc = e % ZG_BKC;
            As[2 * cc][r] = pa[q].x;
            As[2 * cc + 1][r] = pa[q].y;
        }
        // Xt stage: 8 complex rows of X -> 16 real rows of Xt
#pragma unroll
        for (int q = 0; q < XQ; ++q) {
            const int e = t + q * BLOCK;
            if (e < ZG_BKC * 8 * NT) {
                const int j = e / (8 * NT), c = e % (8 * NT);
                *reinterpret_cast<double2*>(&Xs[2 * j][2 * c]) = px[q];
                *reinterpret_cast<double2*>(&Xs[2 * j + 1][2 * c]) = make_double2(-px[q].y, px[q].x);
            }
        }
        __syncthreads();
        if (k0 + ZG_BKC < kend) gload(k0 + ZG_BKC);
#pragma unroll
        for (int ks = 0; ks < (2 * ZG_BKC) / 4; ++ks) {
            const int kr = 4 * ks + (lane >> 4);
            double av[2], bv[NT];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = As[kr][32 * w + 16 * a + (lane & 15)];
#pragma unroll
            for (int b = 0; b < NT; ++b) bv[b] = Xs[kr][16 * b + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout of the 16x16 tile: lane l, register v -> row 4*v + l/16, column l % 16
    double* out = Ypart + (size_t)split * n * NR;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t gr = row0 + 32 * w + 16 * a + 4 * v + (lane >> 4);
            if (gr < n) {
#pragma unroll
                for (int b = 0; b < NT; ++b) out[gr * NR + 16 * b + (lane & 15)] = acc[a][b][v];
            }
        }
}

// Y[r, c] = sum over slices of the partials (complex columns c < kc), slice order fixed
__global__ void zgemm_reduce_kernel(const double* __restrict__ Ypart, int64_t n, int nr, int kc,
                                    double2* __restrict__ Y, int64_t ldy) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * kc) return;
    const int64_t r = e / kc;
    const int c = (int)(e % kc);
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int s = 0; s < ZG_SPLIT; ++s) {
        const double2 p = *reinterpret_cast<const double2*>(Ypart + ((size_t)s * n + r) * nr + 2 * c);
        sr += p.x;
        si += p.y;
    }
    Y[r * ldy + c] = make_double2(sr, si);
}

// ---- column-batched Krylov vector kernels (n x k row-major blocks) ----
// A block reduces over rows [chunk*len, ...) for every column: thread t owns
// column t % k and rows sub, sub + S, ... (S = 256 / k); the per-column
// partials are summed in a fixed order.  The element type T is double2 (FP64
// vectors) or float2 (FP32 vectors of the inner low-precision solve); either
// way the arithmetic and the partial sums are FP64.

__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Rows per thread per pass of the reductions below: a thread keeps BT loads in
// flight (one per row it owns) instead of one dependent load at a time.
constexpr int BT = 8;


// out[(chunk * (nvec + 1) + vec) * k + c] = sum conj(V_vec[:,c]) W[:,c]; vec == nvec -> ||W[:,c]||^2.
// Block (chunk, group) takes vectors group*BV .. group*BV + BV - 1 for its rows
// together: per row pass it issues the loads of all of them, keeps one
// partial per vector in registers, and reduces the BV partials in one tree.
// Its W values stay in registers when the chunk is one pass (S * BT rows),
// which is the usual case.  Per vector a thread adds its rows in row order.
constexpr int BV = 8;
template <typename T>
__global__ __launch_bounds__(BLOCK) void bdots_kernel(const T* __restrict__ V, int64_t ldv, int nvec,
                                                      const T* __restrict__ W, int64_t n, int k,
                                                      int64_t chunk_len, double2* __restrict__ out) {
    __shared__ double2 sh[BV * BLOCK];
    const int S = BLOCK / k;
    const int c = threadIdx.x % k;
    const int sub = threadIdx.x / k;
    const int64_t b = (int64_t)blockIdx.x * chunk_len;
    const int64_t e = min(n, b + chunk_len);
    const bool one_pass = e - b <= (int64_t)S * BT;
    const int vec0 = (int)blockIdx.y * BV;
    const int nv = min(nvec + 1, vec0 + BV) - vec0;
    double2 w[BT];
    if (one_pass && sub < S) {
#pragma unroll
        for (int t = 0; t < BT; ++t) {
            const int64_t r = b + sub + (int64_t)S * t;
            w[t] = r < e ? ld2(W[r * k + c]) : make_double2(0.0, 0.0);
        }
    }
    double2 s[BV];
#pragma unroll
    for (int v = 0; v < BV; ++v) s[v] = make_double2(0.0, 0.0);
    if (sub < S) {
        for (int64_t g = b + sub; g < e; g += (int64_t)S * BT) {
            if (!one_pass) {
#pragma unroll
                for (int t = 0; t < BT; ++t) {
                    const int64_t r = g + (int64_t)S * t;
                    w[t] = r < e ? ld2(W[r * k + c]) : make_double2(0.0, 0.0);
                }
            }
#pragma unroll
            for (int v = 0; v < BV; ++v) {
                if (v >= nv) break;
                const int vec = vec0 + v;
                if (vec < nvec) {
                    const T* __restrict__ vp = V + (int64_t)vec * ldv;
                    double2 x[BT];
#pragma unroll
                    for (int t = 0; t < BT; ++t) {
                        const int64_t r = g + (int64_t)S * t;
                        x[t] = r < e ? ld2(vp[r * k + c]) : make_double2(0.0, 0.0);
                    }
#pragma unroll
                    for (int t = 0; t < BT; ++t) {
                        if (g + (int64_t)S * t < e) {
                            const double2 p = cmulc(x[t], w[t]);
                            s[v].x += p.x;
                            s[v].y += p.y;
                        }
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < BT; ++t)
                        if (g + (int64_t)S * t < e) s[v].x += w[t].x * w[t].x + w[t].y * w[t].y;
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < BV; ++v)
        if (v < nv) sh[v * BLOCK + threadIdx.x] = s[v];
    column_tree_multi(sh, S, k, sub, c, nv);
    if (threadIdx.x < k)
        for (int v = 0; v < nv; ++v)
            out[((int64_t)blockIdx.x * (nvec + 1) + vec0 + v) * k + threadIdx.x] = sh[v * BLOCK + threadIdx.x];
}

// W[:,c] -= sum_i C[i*k+c] V_i[:,c]; norm partials per column into out[chunk*k + c].
// Per element the sum over i runs in order before the subtraction.  A thread
// owns R rows per pass and unrolls the vector loop by U: R = 8, U = 1 for wide
// blocks; R = 2, U = 4 for short chunks (at most 2 rows per thread), where the
// loads of four vectors are in flight together.  Both give the same sums.
template <typename T, int R, int U>
__global__ __launch_bounds__(BLOCK) void bproject_kernel(const T* __restrict__ V, int64_t ldv, int nvec,
                                                         const double2* __restrict__ C, T* __restrict__ W,
                                                         int64_t n, int k, int64_t chunk_len,
                                                         double* __restrict__ out) {
    __shared__ double sh[BLOCK];
    const int S = BLOCK / k;
    const int c = threadIdx.x % k;
    const int sub = threadIdx.x / k;
    const int64_t b = (int64_t)blockIdx.x * chunk_len;
    const int64_t e = min(n, b + chunk_len);
    double nr = 0.0;
    if (sub < S) {
        for (int64_t g = b + sub; g < e; g += (int64_t)S * R) {
            double sr[R], si[R];
#pragma unroll
            for (int t = 0; t < R; ++t) sr[t] = si[t] = 0.0;
#pragma unroll U
            for (int i = 0; i < nvec; ++i) {
                const double2 coef = C[i * k + c];
                const T* __restrict__ v = V + (int64_t)i * ldv;
                double2 x[R];
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const int64_t r = g + (int64_t)S * t;
                    x[t] = r < e ? ld2(v[r * k + c]) : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const double2 p = cmul(coef, x[t]);
                    sr[t] += p.x;
                    si[t] += p.y;
                }
            }
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const int64_t r = g + (int64_t)S * t;
                if (r < e) {
                    double2 w = ld2(W[r * k + c]);
                    w.x -= sr[t];
                    w.y -= si[t];
                    const T ws = st2<T>(w);
                    W[r * k + c] = ws;
                    w = ld2(ws);  // the norm of the stored vector
                    nr += w.x * w.x + w.y * w.y;
                }
            }
        }
    }
    sh[threadIdx.x] = nr;
    column_tree(sh, S, k, sub, c);
    if (threadIdx.x < k) out[(int64_t)blockIdx.x * k + threadIdx.x] = sh[threadIdx.x];
}

// R = B - R (or only the norm when B == nullptr); norm partials per column
template <typename T>
__global__ __launch_bounds__(BLOCK) void bresidual_kernel(const T* __restrict__ B, T* __restrict__ R,
                                                          int64_t n, int k, int64_t chunk_len,
                                                          double* __restrict__ out) {
    __shared__ double sh[BLOCK];
    const int S = BLOCK / k;
    const int c = threadIdx.x % k;
    const int sub = threadIdx.x / k;
    const int64_t b = (int64_t)blockIdx.x * chunk_len;
    const int64_t e = min(n, b + chunk_len);
    double nr = 0.0;
    if (sub < S) {
        for (int64_t r = b + sub; r < e; r += S) {
            double2 v = ld2(R[r * k + c]);
            if (B) {
                const double2 bb = ld2(B[r * k + c]);
                const T vs = st2<T>(make_double2(bb.x - v.x, bb.y - v.y));
                R[r * k + c] = vs;
                v = ld2(vs);
            }
            nr += v.x * v.x + v.y * v.y;
        }
    }
    sh[threadIdx.x] = nr;
    column_tree(sh, S, k, sub, c);
    if (threadIdx.x < k) out[(int64_t)blockIdx.x * k + threadIdx.x] = sh[threadIdx.x];
}

// Column sums of chunk partials in two coalesced passes with a fixed order:
// group g of G adds part[q * nout + o] over q = g, g + G, ...; then
// out[o] = sum over g of the group sums.
constexpr int BSUM_GROUPS = 64;

__global__ void bsum_groups_kernel(const double* __restrict__ part, int chunks, int64_t nout,
                                   double* __restrict__ tmp) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    double s = 0.0;
#pragma unroll 8
    for (int q = blockIdx.y; q < chunks; q += BSUM_GROUPS) s += part[(int64_t)q * nout + o];
    tmp[(int64_t)blockIdx.y * nout + o] = s;
}

// out[o] = sum over g < groups (at most BSUM_GROUPS) of tmp[g * nout + o], in
// g order; all loads are issued before the additions.
__global__ __launch_bounds__(BLOCK) void bsum_small_kernel(const double* __restrict__ tmp, int64_t nout, int groups,
                                                          double* __restrict__ out) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    double v[BSUM_GROUPS];
#pragma unroll
    for (int g = 0; g < BSUM_GROUPS; ++g) v[g] = g < groups ? tmp[(int64_t)g * nout + o] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int g = 0; g < BSUM_GROUPS; ++g)
        if (g < groups) s += v[g];
    out[o] = s;
}

// X[:,c] += sum_i Yc[i*k+c] V_i[:,c]
template <typename T>
__global__ void bupdate_kernel(const T* __restrict__ V, int64_t ldv, int nvec, const double2* __restrict__ Yc,
                               T* __restrict__ X, int64_t n, int k) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int c = (int)(t % k);
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < nvec; ++i) {
        double2 p = cmul(Yc[i * k + c], ld2(V[(int64_t)i * ldv + t]));
        sr += p.x;
        si += p.y;
    }
    const double2 x = ld2(X[t]);
    X[t] = st2<T>(make_double2(x.x + sr, x.y + si));
}

// Y[:,c] = alpha[c] * X[:,c]
template <typename T>
__global__ void bscale_kernel(const T* __restrict__ X, const double* __restrict__ alpha,
                              T* __restrict__ Y, int64_t n, int k) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const double a = alpha[t % k];
    const double2 v = ld2(X[t]);
    Y[t] = st2<T>(make_double2(v.x * a, v.y * a));
}

template <typename T>
__global__ void zero_kernel(T* __restrict__ X, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) X[t] = st2<T>(make_double2(0.0, 0.0));
}

// Element-type conversion of a block (FP64 <-> FP32, rounding to nearest).
template <typename S, typename D>
__global__ void convert_block_kernel(const S* __restrict__ X, D* __restrict__ Y, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) Y[t] = st2<D>(ld2(X[t]));
}

inline unsigned nblocks(int64_t n, int per = BLOCK) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int64_t lowp_padded_ld(int64_t n) { return (n + 511) / 512 * 512; }

int64_t lowp_split_count() { return LP_SPLIT; }

// Rows of the fragment-ordered copy: n rounded up to whole 32-row tiles.
int64_t lowp_padded_rows(int64_t n) { return (n + 31) / 32 * 32; }

void lowp_convert(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, int64_t ld, int format,
                  uint16_t* out, double* row_inv, const double* rowmax, int64_t stride) {
    const int top = format == 0 ? Half<0>::TOP : Half<1>::TOP;
    if (rowmax)
        hipLaunchKernelGGL(lp_rowscale_from_max_kernel, dim3(nblocks(n, BLOCK / 64)), dim3(BLOCK), 0, ctx->stream,
                           rowmax, stride, n, top, row_inv);
    else
        hipLaunchKernelGGL(lp_rowscale_kernel, dim3(nblocks(n, BLOCK / 64)), dim3(BLOCK), 0, ctx->stream, A, lda, n,
                           top, row_inv);
    SBH_HIP(hipGetLastError());
    const int64_t pack_blocks = (n + 31) / 32 * (ld / LP_PACK_COLS);
    SBH_REQUIRE(pack_blocks < ((int64_t)1 << 31), "low-precision copy too large");
    const dim3 grid((unsigned)pack_blocks);
    if (format == 0)
        hipLaunchKernelGGL(lp_pack_kernel<0>, grid, dim3(BLOCK), 0, ctx->stream, A, lda, n, ld, row_inv,
                           reinterpret_cast<u16*>(out));
    else
        hipLaunchKernelGGL(lp_pack_kernel<1>, grid, dim3(BLOCK), 0, ctx->stream, A, lda, n, ld, row_inv,
                           reinterpret_cast<u16*>(out));
    SBH_HIP(hipGetLastError());
}

// Y = A X for k in 1..64 columns (k = 1 on lp_gemv_kernel).  Rows k..32*ntr-1 of XT (left from a wider
// product or zero) feed only output columns that lp_reduce_kernel discards.
template <typename T>
static void lowp_gemm_t(sbh_context* ctx, const uint16_t* A, int64_t ld, int64_t n, int format, const double* row_inv,
                        const T* X, int k, uint16_t* XT, unsigned* colmax, double* col_inv, float* Ypart, T* Y) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, ctx->stream, colmax, k, 0u);
    {
        const int64_t rpb = std::max<int64_t>(1, 4096 / k);
        hipLaunchKernelGGL(colmax_kernel<T>, dim3((unsigned)((n + rpb - 1) / rpb)), dim3(BLOCK), 0, ctx->stream, X, n,
                           k, rpb, colmax);
    }
    if (format == 0)
        hipLaunchKernelGGL((x_to_lp_kernel<0, T>), dim3(nblocks(n * k)), dim3(BLOCK), 0, ctx->stream, X, n, k, ld,
                           colmax, XT, col_inv);
    else
        hipLaunchKernelGGL((x_to_lp_kernel<1, T>), dim3(nblocks(n * k)), dim3(BLOCK), 0, ctx->stream, X, n, k, ld,
                           colmax, XT, col_inv);
    SBH_HIP(hipGetLastError());
    const int ntr = (k + 31) / 32;  // 32-column tiles of the real part (the imaginary part has as many)
    const int64_t n_tiles = (n + 31) / 32;
    if (k == 1) {
        const int64_t blocks = (n_tiles + LV_WAVES - 1) / LV_WAVES * LP_SPLIT;
        SBH_REQUIRE(blocks < ((int64_t)1 << 31), "low-precision product too large");
        hipEvent_t ev;
        ctx->begin_timed(SBH_KERNEL_LOWP, &ev);
        if (format == 0)
            hipLaunchKernelGGL(lp_gemv_kernel<0>, dim3((unsigned)blocks), dim3(LV_WAVES * 64), 0, ctx->stream,
                               reinterpret_cast<const u16*>(A), ld, n, reinterpret_cast<const u16*>(XT), Ypart);
        else
            hipLaunchKernelGGL(lp_gemv_kernel<1>, dim3((unsigned)blocks), dim3(LV_WAVES * 64), 0, ctx->stream,
                               reinterpret_cast<const u16*>(A), ld, n, reinterpret_cast<const u16*>(XT), Ypart);
        SBH_HIP(hipGetLastError());
        hipLaunchKernelGGL(lp_reduce_kernel<T>, dim3(nblocks(n)), dim3(BLOCK), 0, ctx->stream, Ypart, n, 1, 2,
                           row_inv, col_inv, Y);
        SBH_HIP(hipGetLastError());
        ctx->end_timed(SBH_KERNEL_LOWP, ev);
        return;
    }
    const int bn = 2 * ntr * 32;
    const int64_t row_blocks = (n_tiles + LP_WAVES - 1) / LP_WAVES;
    SBH_REQUIRE(row_blocks * LP_SPLIT < ((int64_t)1 << 31), "low-precision product too large");
    const unsigned grid = (unsigned)(row_blocks * LP_SPLIT);
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_LOWP, &ev);
#define SBH_LP_GEMM(F, NTV)                                                                                          \
    hipLaunchKernelGGL((lp_gemm_kernel<F, NTV>), dim3(grid), dim3(LP_THREADS), 0, ctx->stream,                       \
                       reinterpret_cast<const u16*>(A), ld, n, reinterpret_cast<const u16*>(XT), Ypart)
    if (format == 0) {
        if (ntr == 1) SBH_LP_GEMM(0, 1);
        else SBH_LP_GEMM(0, 2);
    } else {
        if (ntr == 1) SBH_LP_GEMM(1, 1);
        else SBH_LP_GEMM(1, 2);
    }
#undef SBH_LP_GEMM
    SBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(lp_reduce_kernel<T>, dim3(nblocks(n * k)), dim3(BLOCK), 0, ctx->stream, Ypart, n, k, bn,
                       row_inv, col_inv, Y);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_LOWP, ev);
}

#define SBH_LOWP_PRODUCTS(T)                                                                                         \
    void lowp_gemm(sbh_context* ctx, const uint16_t* A, int64_t ld, int64_t n, int format, const double* row_inv,    \
                   const T* X, int k, uint16_t* XT, unsigned* colmax, double* col_inv, float* Ypart, T* Y) {         \
        lowp_gemm_t(ctx, A, ld, n, format, row_inv, X, k, XT, colmax, col_inv, Ypart, Y);                            \
    }
SBH_LOWP_PRODUCTS(double2)
SBH_LOWP_PRODUCTS(float2)
#undef SBH_LOWP_PRODUCTS

// Split-K partials of one 64-column chunk: ZG_SPLIT x n x nr doubles.
static size_t zgemm_partial_doubles(int64_t n, int64_t k) {
    const int kc = (int)std::min<int64_t>(64, k);
    const int nt = kc <= 8 ? 1 : kc <= 16 ? 2 : kc <= 32 ? 4 : 8;
    return (size_t)ZG_SPLIT * (size_t)n * (size_t)(16 * nt);
}

// Y = A X in 64-column chunks; part (null: the context's scratch, grown on
// demand) holds the split-K partials of one chunk and is fully rewritten by
// every chunk.
static void zgemm_f64_into(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X, int64_t k,
                           double2* Y, double* part) {
    const int64_t kslice = ((n + ZG_SPLIT - 1) / ZG_SPLIT + ZG_BKC - 1) / ZG_BKC * ZG_BKC;
    const int64_t row_tiles = (n + ZG_BM - 1) / ZG_BM;
    SBH_REQUIRE(row_tiles * ZG_SPLIT < ((int64_t)1 << 31), "block product too large");
    for (int64_t c0 = 0; c0 < k; c0 += 64) {
        const int kc = (int)std::min<int64_t>(64, k - c0);
        const int nt = kc <= 8 ? 1 : kc <= 16 ? 2 : kc <= 32 ? 4 : 8;
        const int nr = 16 * nt;
        double* chunk_part = part ? part : reinterpret_cast<double*>(ctx->device_scratch((size_t)ZG_SPLIT * n * nr / 2));
        const dim3 grid((unsigned)(row_tiles * ZG_SPLIT));
#define SBH_ZG(NT_)                                                                                           \
    hipLaunchKernelGGL(zgemm_f64_kernel<NT_>, grid, dim3(BLOCK), 0, ctx->stream, A, lda, n, X + c0, k, kc, kslice, \
                       chunk_part)
        if (nt == 1) SBH_ZG(1);
        else if (nt == 2) SBH_ZG(2);
        else if (nt == 4) SBH_ZG(4);
        else SBH_ZG(8);
#undef SBH_ZG
        SBH_HIP(hipGetLastError());
        hipLaunchKernelGGL(zgemm_reduce_kernel, dim3(nblocks(n * kc)), dim3(BLOCK), 0, ctx->stream, chunk_part, n, nr,
                           kc, Y + c0, k);
        SBH_HIP(hipGetLastError());
    }
}

void zgemm_f64(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X, int64_t k, double2* Y) {
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_BLOCK, &ev);
    zgemm_f64_into(ctx, A, lda, n, X, k, Y, nullptr);
    ctx->end_timed(SBH_KERNEL_BLOCK, ev);
}

// ---- allocation-free block product (sbh_operator_apply_block_noalloc) ----
// The split-K partials live in a fixed device array instead of the context's
// scratch.  Host code never passes the array's name to a kernel: its device
// address on each GPU comes from hipGetSymbolAddress, taken once per device
// at context set-up with that device current.  Every call on one GPU shares
// the array; calls that overlap on one GPU race on it.  Each chunk's product
// rewrites every partial it reads, so nothing needs clearing between calls.
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

void noalloc_scratch_init(int device) {
    SBH_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, "device index beyond the allocation-free scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    SBH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SBH_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address for the block-product scratch");
    g_scratch_ptr[device] = static_cast<double*>(ptr);
}

size_t noalloc_block_scratch_need(int64_t n, int64_t k) { return zgemm_partial_doubles(n, k); }

void zgemm_f64_noalloc(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X, int64_t k,
                       double2* Y) {
    double* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ptr = (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[ctx->device] : nullptr;
    }
    if (!ptr) throw Error(SBH_ERR_BACKEND, "allocation-free scratch was not set up on this device");
    zgemm_f64_into(ctx, A, lda, n, X, k, Y, ptr);
}

// Rows per block of the column-batched reductions: one pass of BT rows per
// thread (S * BT rows, S = BLOCK / k threads per column), at most 4096 blocks.
static int batched_chunks(int64_t n, int k, int64_t* len) {
    const int64_t rpb = (int64_t)(BLOCK / k) * BT;
    int64_t chunks = (n + rpb - 1) / rpb;
    // few columns: shorter chunks (down to one row per thread) for more
    // workgroups, up to the count whose sums still take one pass
    if (chunks < BSUM_GROUPS) chunks = std::min<int64_t>(BSUM_GROUPS, (n + BLOCK / k - 1) / (BLOCK / k));
    chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, 4096));
    *len = (n + chunks - 1) / chunks;
    return (int)chunks;
}

// out[o] = sum over the chunks of part[q * nout + o]; the group sums go right
// after the partials in the same scratch buffer.
static void column_sums(sbh_context* ctx, double* part, int chunks, int64_t nout, double* out) {
    if (chunks <= BSUM_GROUPS) {
        // one pass in chunk order: the same sums as the two passes, whose groups then hold one chunk each
        hipLaunchKernelGGL(bsum_small_kernel, dim3(nblocks(nout)), dim3(BLOCK), 0, ctx->stream, part, nout, chunks,
                           out);
        SBH_HIP(hipGetLastError());
        return;
    }
    double* tmp = part + (int64_t)chunks * nout;
    hipLaunchKernelGGL(bsum_groups_kernel, dim3(nblocks(nout), BSUM_GROUPS), dim3(BLOCK), 0, ctx->stream, part,
                       chunks, nout, tmp);
    SBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(bsum_small_kernel, dim3(nblocks(nout)), dim3(BLOCK), 0, ctx->stream, tmp, nout, BSUM_GROUPS,
                       out);
    SBH_HIP(hipGetLastError());
}

// The wrappers below exist for double2 (FP64 vectors) and float2 (FP32
// vectors); the partial sums and reduced values are FP64 either way.

// Partials land in `scratch` (chunks x (nvec+1) x k complex); the reduced
// (nvec+1) x k complex values land in `out`.
template <typename T>
static void bdots_t(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const T* W, int64_t n, int k,
                    double2* scratch, double2* out) {
    int64_t len;
    const int chunks = batched_chunks(n, k, &len);
    hipLaunchKernelGGL(bdots_kernel<T>, dim3((unsigned)chunks, (unsigned)((nvec + BV) / BV)), dim3(BLOCK), 0,
                       ctx->stream, V, ldv, nvec, W, n, k, len, scratch);
    SBH_HIP(hipGetLastError());
    column_sums(ctx, reinterpret_cast<double*>(scratch), chunks, (int64_t)(nvec + 1) * k * 2,
                reinterpret_cast<double*>(out));
}

// W -= sum C_i V_i per column; the reduced k squared norms land in `out`
// (out == nullptr: no norms wanted, and their sums are not launched).
template <typename T>
static void bproject_t(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const double2* C, T* W, int64_t n, int k,
                       double* scratch, double* out) {
    int64_t len;
    const int chunks = batched_chunks(n, k, &len);
    if (len <= 2 * (int64_t)(BLOCK / k))
        hipLaunchKernelGGL((bproject_kernel<T, 2, 4>), dim3((unsigned)chunks), dim3(BLOCK), 0, ctx->stream, V, ldv,
                           nvec, C, W, n, k, len, scratch);
    else
        hipLaunchKernelGGL((bproject_kernel<T, BT, 1>), dim3((unsigned)chunks), dim3(BLOCK), 0, ctx->stream, V, ldv,
                           nvec, C, W, n, k, len, scratch);
    SBH_HIP(hipGetLastError());
    if (out) column_sums(ctx, scratch, chunks, (int64_t)k, out);
}

// R = B - R (B may be null: norms only); the reduced k squared norms land in `out`.
template <typename T>
static void bresidual_t(sbh_context* ctx, const T* B, T* R, int64_t n, int k, double* scratch, double* out) {
    int64_t len;
    const int chunks = batched_chunks(n, k, &len);
    hipLaunchKernelGGL(bresidual_kernel<T>, dim3((unsigned)chunks), dim3(BLOCK), 0, ctx->stream, B, R, n, k, len,
                       scratch);
    SBH_HIP(hipGetLastError());
    column_sums(ctx, scratch, chunks, (int64_t)k, out);
}

int64_t batched_scratch_elems(int64_t n, int k, int max_vec) {
    int64_t len;
    const int chunks = batched_chunks(n, k, &len);
    return (int64_t)(chunks + BSUM_GROUPS) * (max_vec + 1) * k;  // complex elements: partials + group sums
}

template <typename T>
static void bupdate_t(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const double2* Yc, T* X, int64_t n, int k) {
    hipLaunchKernelGGL(bupdate_kernel<T>, dim3(nblocks(n * k)), dim3(BLOCK), 0, ctx->stream, V, ldv, nvec, Yc, X, n,
                       k);
    SBH_HIP(hipGetLastError());
}

template <typename T>
static void bscale_t(sbh_context* ctx, const T* X, const double* alpha, T* Y, int64_t n, int k) {
    hipLaunchKernelGGL(bscale_kernel<T>, dim3(nblocks(n * k)), dim3(BLOCK), 0, ctx->stream, X, alpha, Y, n, k);
    SBH_HIP(hipGetLastError());
}

template <typename T>
static void zero_fill_t(sbh_context* ctx, T* X, int64_t count) {
    if (count <= 0) return;
    hipLaunchKernelGGL(zero_kernel<T>, dim3(nblocks(count)), dim3(BLOCK), 0, ctx->stream, X, count);
    SBH_HIP(hipGetLastError());
}

#define SBH_VECTOR_OPS(T)                                                                                       \
    void bdots(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const T* W, int64_t n, int k,               \
               double2* scratch, double2* out) {                                                                \
        bdots_t(ctx, V, ldv, nvec, W, n, k, scratch, out);                                                      \
    }                                                                                                           \
    void bproject(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const double2* C, T* W, int64_t n, int k, \
                  double* scratch, double* out) {                                                               \
        bproject_t(ctx, V, ldv, nvec, C, W, n, k, scratch, out);                                                \
    }                                                                                                           \
    void bresidual(sbh_context* ctx, const T* B, T* R, int64_t n, int k, double* scratch, double* out) {        \
        bresidual_t(ctx, B, R, n, k, scratch, out);                                                             \
    }                                                                                                           \
    void bupdate(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const double2* Yc, T* X, int64_t n,       \
                 int k) {                                                                                       \
        bupdate_t(ctx, V, ldv, nvec, Yc, X, n, k);                                                              \
    }                                                                                                           \
    void bscale(sbh_context* ctx, const T* X, const double* alpha, T* Y, int64_t n, int k) {                    \
        bscale_t(ctx, X, alpha, Y, n, k);                                                                       \
    }                                                                                                           \
    void zero_fill(sbh_context* ctx, T* X, int64_t count) { zero_fill_t(ctx, X, count); }
SBH_VECTOR_OPS(double2)
SBH_VECTOR_OPS(float2)
#undef SBH_VECTOR_OPS

void convert_block(sbh_context* ctx, const double2* X, float2* Y, int64_t count) {
    if (count <= 0) return;
    hipLaunchKernelGGL((convert_block_kernel<double2, float2>), dim3(nblocks(count)), dim3(BLOCK), 0, ctx->stream, X,
                       Y, count);
    SBH_HIP(hipGetLastError());
}

void convert_block(sbh_context* ctx, const float2* X, double2* Y, int64_t count) {
    if (count <= 0) return;
    hipLaunchKernelGGL((convert_block_kernel<float2, double2>), dim3(nblocks(count)), dim3(BLOCK), 0, ctx->stream, X,
                       Y, count);
    SBH_HIP(hipGetLastError());
}

}  // namespace sbh
