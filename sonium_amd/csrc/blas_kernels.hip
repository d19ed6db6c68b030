// Dense complex FP64 kernels behind the LinearOperator and GMRES entry points.
//
//  zgemv_n   y = A x, A row-major (DenseMatrix::matvec_into → gemv_c64,
//            crates/sonium-solvers/crates/sonium-solvers-core/src/dense.rs:416-420,
//            traits.rs:1430-1453).  HBM-bound: 16 N^2 bytes per call.  Each
//            wave owns 4 rows and streams them with 16-byte loads, reusing
//            every x element it loads for the 4 rows; the row sums are
//            reduced across the wave with cross-lane shuffles, so the result
//            does not depend on scheduling.
//  zgemv_t   y = A^T x or A^H x: column-per-lane partial sums over row chunks
//            plus a fixed-order reduction over chunks.
//  Krylov vector kernels for classical Gram–Schmidt with DGKS
//  reorthogonalisation (cgs_orth.rs:180-301): batched conjugated dots
//  <v_i, w> = sum conj(v_i) w, the fused projection w -= sum c_i v_i with the
//  new norm, the basis update x += sum y_i v_i, and residual/scale helpers.
//  Every reduction writes per-block partials that the host sums in index
//  order, so results are deterministic.

#include <algorithm>

#include "internal.hpp"

namespace sbh {
namespace {

constexpr int BLOCK = 256;

typedef double v2d __attribute__((ext_vector_type(2)));

// Streaming load of one complex element that is read exactly once.
__device__ __forceinline__ double2 load_stream(const double2* p) {
    v2d v = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
    return make_double2(v.x, v.y);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of one double in a fixed order: wave sums, then waves 0..3.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < BLOCK / 64; ++w) t += sh[w];
    }
    return t;
}

template <int ROWS>
__global__ __launch_bounds__(BLOCK) void zgemv_n_kernel(const double2* __restrict__ A, int64_t lda,
                                                        int64_t rows, int64_t cols,
                                                        const double2* __restrict__ x, int64_t incx,
                                                        double2* __restrict__ y, int64_t incy) {
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)) * ROWS;
    if (r0 >= rows) return;
    const double2* a[ROWS];
    double sr[ROWS], si[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        int64_t rr = min(r0 + r, rows - 1);
        a[r] = A + rr * lda;
        sr[r] = 0.0;
        si[r] = 0.0;
    }
    constexpr int U = 4;
    int64_t c = lane;
    if (incx == 1) {
        for (; c + 64 * (U - 1) < cols; c += 64 * U) {
            double2 xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[c + 64 * u];
            double2 av[ROWS][U];
#pragma unroll
            for (int r = 0; r < ROWS; ++r)
#pragma unroll
                for (int u = 0; u < U; ++u) av[r][u] = load_stream(a[r] + c + 64 * u);
#pragma unroll
            for (int r = 0; r < ROWS; ++r)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    sr[r] = __builtin_fma(av[r][u].x, xv[u].x, sr[r]);
                    sr[r] = __builtin_fma(-av[r][u].y, xv[u].y, sr[r]);
                    si[r] = __builtin_fma(av[r][u].x, xv[u].y, si[r]);
                    si[r] = __builtin_fma(av[r][u].y, xv[u].x, si[r]);
                }
        }
    }
    for (; c < cols; c += 64) {
        double2 xv = x[c * incx];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            double2 av = a[r][c];
            sr[r] = __builtin_fma(av.x, xv.x, sr[r]);
            sr[r] = __builtin_fma(-av.y, xv.y, sr[r]);
            si[r] = __builtin_fma(av.x, xv.y, si[r]);
            si[r] = __builtin_fma(av.y, xv.x, si[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        sr[r] = wave_sum(sr[r]);
        si[r] = wave_sum(si[r]);
    }
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
            if (r0 + r < rows) y[(r0 + r) * incy] = make_double2(sr[r], si[r]);
    }
}

// Partial column sums of A^T x (or A^H x) over rows [chunk*len, ...).
__global__ __launch_bounds__(BLOCK) void zgemv_t_partial_kernel(const double2* __restrict__ A,
                                                                int64_t lda, int64_t rows,
                                                                int64_t cols, const double2* __restrict__ x,
                                                                int64_t rows_per_chunk, int conj,
                                                                double2* __restrict__ part) {
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.y * rows_per_chunk;
    const int64_t i1 = min(rows, i0 + rows_per_chunk);
    if (j >= cols) return;
    double sr = 0.0, si = 0.0;
    for (int64_t i = i0; i < i1; ++i) {
        double2 a = A[i * lda + j];
        double2 xv = x[i];
        if (conj) a.y = -a.y;
        sr = __builtin_fma(a.x, xv.x, sr);
        sr = __builtin_fma(-a.y, xv.y, sr);
        si = __builtin_fma(a.x, xv.y, si);
        si = __builtin_fma(a.y, xv.x, si);
    }
    part[(int64_t)blockIdx.y * cols + j] = make_double2(sr, si);
}

__global__ void sum_chunks_kernel(const double2* __restrict__ part, int64_t cols, int chunks,
                                  double2* __restrict__ y) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < chunks; ++c) {
        double2 p = part[(int64_t)c * cols + j];
        s.x += p.x;
        s.y += p.y;
    }
    y[j] = s;
}

// grid (chunks, nvec+1): vec < nvec → <v_vec, w>; vec == nvec → ||w||^2.
__global__ __launch_bounds__(BLOCK) void dots_norm_kernel(const double2* __restrict__ V, int64_t ldv,
                                                          int nvec, const double2* __restrict__ w,
                                                          int64_t n, int64_t chunk_len,
                                                          double2* __restrict__ out) {
    __shared__ double sh[BLOCK / 64];
    const int vec = blockIdx.y;
    const int64_t b = (int64_t)blockIdx.x * chunk_len;
    const int64_t e = min(n, b + chunk_len);
    double sr = 0.0, si = 0.0;
    if (vec < nvec) {
        const double2* v = V + (int64_t)vec * ldv;
        for (int64_t k = b + threadIdx.x; k < e; k += BLOCK) {
            double2 p = cmulc(v[k], w[k]);
            sr += p.x;
            si += p.y;
        }
    } else {
        for (int64_t k = b + threadIdx.x; k < e; k += BLOCK) {
            double2 q = w[k];
            sr += q.x * q.x + q.y * q.y;
        }
    }
    double tr = block_sum(sr, sh);
    double ti = block_sum(si, sh);
    if (threadIdx.x == 0) out[(int64_t)vec * gridDim.x + blockIdx.x] = make_double2(tr, ti);
}

__global__ __launch_bounds__(BLOCK) void project_norm_kernel(const double2* __restrict__ V, int64_t ldv,
                                                             int nvec, const double2* __restrict__ c,
                                                             double2* __restrict__ w, int64_t n,
                                                             double* __restrict__ norm_out) {
    __shared__ double sh[BLOCK / 64];
    const int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double nr = 0.0;
    if (k < n) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < nvec; ++i) {
            double2 p = cmul(c[i], V[(int64_t)i * ldv + k]);
            sr += p.x;
            si += p.y;
        }
        double2 wk = w[k];
        wk.x -= sr;
        wk.y -= si;
        w[k] = wk;
        nr = wk.x * wk.x + wk.y * wk.y;
    }
    double t = block_sum(nr, sh);
    if (threadIdx.x == 0) norm_out[blockIdx.x] = t;
}

__global__ void basis_update_kernel(const double2* __restrict__ V, int64_t ldv, int nvec,
                                    const double2* __restrict__ c, double2* __restrict__ x, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < nvec; ++i) {
        double2 p = cmul(c[i], V[(int64_t)i * ldv + k]);
        sr += p.x;
        si += p.y;
    }
    double2 xk = x[k];
    x[k] = make_double2(xk.x + sr, xk.y + si);
}

__global__ void scale_copy_kernel(const double2* __restrict__ x, double alpha, double2* __restrict__ y,
                                  int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double2 v = x[k];
    y[k] = make_double2(v.x * alpha, v.y * alpha);
}

__global__ __launch_bounds__(BLOCK) void residual_norm_kernel(const double2* __restrict__ b,
                                                              double2* __restrict__ r, int64_t n,
                                                              double* __restrict__ norm_out) {
    __shared__ double sh[BLOCK / 64];
    const int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double nr = 0.0;
    if (k < n) {
        double2 bv = b[k], rv = r[k];
        double2 d = make_double2(bv.x - rv.x, bv.y - rv.y);
        r[k] = d;
        nr = d.x * d.x + d.y * d.y;
    }
    double t = block_sum(nr, sh);
    if (threadIdx.x == 0) norm_out[blockIdx.x] = t;
}

__global__ __launch_bounds__(BLOCK) void norm_kernel(const double2* __restrict__ x, int64_t n,
                                                     double* __restrict__ norm_out) {
    __shared__ double sh[BLOCK / 64];
    const int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double nr = 0.0;
    if (k < n) {
        double2 v = x[k];
        nr = v.x * v.x + v.y * v.y;
    }
    double t = block_sum(nr, sh);
    if (threadIdx.x == 0) norm_out[blockIdx.x] = t;
}

__global__ void diag_scale_kernel(const double2* __restrict__ d, const double2* __restrict__ x,
                                  double2* __restrict__ y, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    y[k] = cmul(d[k], x[k]);
}

// num-complex inv(): (re/|z|^2, -im/|z|^2)
__global__ void inv_diag_kernel(const double2* __restrict__ A, int64_t lda, int64_t n,
                                double2* __restrict__ d) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double2 a = A[k * lda + k];
    double m = a.x * a.x + a.y * a.y;
    d[k] = make_double2(a.x / m, -a.y / m);
}

// Plain copy of n complex values (used to calibrate HBM byte counters).
__global__ void copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

inline unsigned blocks_for(int64_t n, int per = BLOCK) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int reduction_chunks(int64_t n) {
    int64_t c = (n + 4095) / 4096;
    return (int)std::max<int64_t>(1, std::min<int64_t>(c, 64));
}

void zgemv_n(sbh_context* ctx, const double2* A, int64_t lda, int64_t rows, int64_t cols,
             const double2* x, double2* y) {
    if (rows == 0) return;
    constexpr int ROWS = 4;
    const int64_t waves = (rows + ROWS - 1) / ROWS;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_MATVEC, &ev);
    hipLaunchKernelGGL(zgemv_n_kernel<ROWS>, dim3(blocks_for(waves, BLOCK / 64)), dim3(BLOCK), 0,
                       ctx->stream, A, lda, rows, cols, x, (int64_t)1, y, (int64_t)1);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_MATVEC, ev);
}

static void zgemv_n_strided(sbh_context* ctx, const double2* A, int64_t lda, int64_t rows,
                            int64_t cols, const double2* x, int64_t incx, double2* y, int64_t incy) {
    constexpr int ROWS = 4;
    const int64_t waves = (rows + ROWS - 1) / ROWS;
    hipLaunchKernelGGL(zgemv_n_kernel<ROWS>, dim3(blocks_for(waves, BLOCK / 64)), dim3(BLOCK), 0,
                       ctx->stream, A, lda, rows, cols, x, incx, y, incy);
    SBH_HIP(hipGetLastError());
}

void zgemv_n_untimed(sbh_context* ctx, const double2* A, int64_t lda, int64_t rows, int64_t cols,
                     const double2* x, double2* y) {
    if (rows == 0) return;
    zgemv_n_strided(ctx, A, lda, rows, cols, x, 1, y, 1);
}

void zgemv_t(sbh_context* ctx, const double2* A, int64_t lda, int64_t rows, int64_t cols,
             const double2* x, double2* y, bool conjugate) {
    if (cols == 0) return;
    int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(64, (rows + 255) / 256));
    int64_t rows_per_chunk = (rows + chunks - 1) / chunks;
    double2* part = ctx->device_scratch((size_t)chunks * (size_t)cols);
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_MATVEC, &ev);
    hipLaunchKernelGGL(zgemv_t_partial_kernel, dim3(blocks_for(cols), (unsigned)chunks), dim3(BLOCK), 0,
                       ctx->stream, A, lda, rows, cols, x, rows_per_chunk, conjugate ? 1 : 0, part);
    SBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_chunks_kernel, dim3(blocks_for(cols)), dim3(BLOCK), 0, ctx->stream,
                       part, cols, chunks, y);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_MATVEC, ev);
}

void zgemm_block(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X,
                 double2* Y, int64_t k) {
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_BLOCK, &ev);
    for (int64_t c = 0; c < k; ++c) zgemv_n_strided(ctx, A, lda, n, n, X + c, k, Y + c, k);
    ctx->end_timed(SBH_KERNEL_BLOCK, ev);
}

int dots_and_norm_partials(sbh_context* ctx, const double2* V, int64_t ldv, int nvec,
                           const double2* w, int64_t n, double2* out) {
    int chunks = reduction_chunks(n);
    int64_t len = (n + chunks - 1) / chunks;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_KRYLOV_VECTOR, &ev);
    hipLaunchKernelGGL(dots_norm_kernel, dim3((unsigned)chunks, (unsigned)(nvec + 1)), dim3(BLOCK), 0,
                       ctx->stream, V, ldv, nvec, w, n, len, out);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_KRYLOV_VECTOR, ev);
    return chunks;
}

// s[v] = sum over c < chunks of part[v * chunks + c], added in c order (the
// order in which the host adds the same partials).
__global__ void chunk_sums_kernel(const double2* __restrict__ part, int nv, int chunks, double2* __restrict__ s) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < chunks; ++c) {
        const double2 p = part[(int64_t)v * chunks + c];
        sr += p.x;
        si += p.y;
    }
    s[v] = make_double2(sr, si);
}

void sum_chunk_partials(sbh_context* ctx, const double2* part, int nv, int chunks, double2* s) {
    hipLaunchKernelGGL(chunk_sums_kernel, dim3((unsigned)((nv + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, ctx->stream,
                       part, nv, chunks, s);
    SBH_HIP(hipGetLastError());
}

int project_and_norm(sbh_context* ctx, const double2* V, int64_t ldv, int nvec, const double2* c,
                     double2* w, int64_t n, double* norm_out) {
    unsigned nb = blocks_for(n);
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_KRYLOV_VECTOR, &ev);
    hipLaunchKernelGGL(project_norm_kernel, dim3(nb), dim3(BLOCK), 0, ctx->stream, V, ldv, nvec, c, w,
                       n, norm_out);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_KRYLOV_VECTOR, ev);
    return (int)nb;
}

void basis_update(sbh_context* ctx, const double2* V, int64_t ldv, int nvec, const double2* c,
                  double2* x, int64_t n) {
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_KRYLOV_VECTOR, &ev);
    hipLaunchKernelGGL(basis_update_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, ctx->stream, V, ldv,
                       nvec, c, x, n);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_KRYLOV_VECTOR, ev);
}

void scale_copy(sbh_context* ctx, const double2* x, double alpha, double2* y, int64_t n) {
    hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, ctx->stream, x, alpha,
                       y, n);
    SBH_HIP(hipGetLastError());
}

int residual_and_norm(sbh_context* ctx, const double2* b, double2* r, int64_t n, double* norm_out) {
    unsigned nb = blocks_for(n);
    hipLaunchKernelGGL(residual_norm_kernel, dim3(nb), dim3(BLOCK), 0, ctx->stream, b, r, n, norm_out);
    SBH_HIP(hipGetLastError());
    return (int)nb;
}

int norm_partials(sbh_context* ctx, const double2* x, int64_t n, double* out) {
    unsigned nb = blocks_for(n);
    hipLaunchKernelGGL(norm_kernel, dim3(nb), dim3(BLOCK), 0, ctx->stream, x, n, out);
    SBH_HIP(hipGetLastError());
    return (int)nb;
}

void diag_scale(sbh_context* ctx, const double2* d, const double2* x, double2* y, int64_t n) {
    hipLaunchKernelGGL(diag_scale_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, ctx->stream, d, x, y, n);
    SBH_HIP(hipGetLastError());
}

// Ceiling probes for the rooflines.  stream_read_kernel reads `count` 16-byte
// words: workgroup b owns the contiguous range of tiles [b*full/grid,
// (b+1)*full/grid), a tile being STREAM_LOADS x 256 words (32 KB) read with
// STREAM_LOADS non-temporal loads in flight per thread; the words past the last
// full tile are read grid-stride.  The XOR of the data is stored only if it
// equals `magic`, so the loads stay.
typedef unsigned probe_u32x4 __attribute__((ext_vector_type(4)));
constexpr int STREAM_LOADS = 8;

__global__ __launch_bounds__(256) void stream_read_kernel(const probe_u32x4* __restrict__ x, int64_t count,
                                                         unsigned magic, unsigned* __restrict__ sink) {
    constexpr int64_t TILE = (int64_t)STREAM_LOADS * 256;
    const int64_t full = count / TILE;
    const int64_t t0 = full * blockIdx.x / gridDim.x, t1 = full * (blockIdx.x + 1) / gridDim.x;
    probe_u32x4 acc = {0u, 0u, 0u, 0u};
    for (int64_t t = t0; t < t1; ++t) {
        const probe_u32x4* p = x + t * TILE + threadIdx.x;
        probe_u32x4 v[STREAM_LOADS];
#pragma unroll
        for (int j = 0; j < STREAM_LOADS; ++j) v[j] = __builtin_nontemporal_load(p + j * 256);
#pragma unroll
        for (int j = 0; j < STREAM_LOADS; ++j) acc ^= v[j];
    }
    for (int64_t i = full * TILE + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        acc ^= x[i];
    const unsigned t = acc.x ^ acc.y ^ acc.z ^ acc.w;
    if (t == magic) sink[0] = t;
}

// FMA_CHAINS independent FP64 FMA chains of `iters` steps per thread; the sum
// is stored only if it equals `magic`.
constexpr int FMA_CHAINS = 16;
__global__ __launch_bounds__(256) void fma_rate_kernel(int64_t iters, double b, double c, double magic,
                                                      double* __restrict__ sink) {
    double a[FMA_CHAINS];
#pragma unroll
    for (int u = 0; u < FMA_CHAINS; ++u) a[u] = (double)(threadIdx.x + u);
    for (int64_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < FMA_CHAINS; ++u) a[u] = __builtin_fma(a[u], b, c);
    }
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < FMA_CHAINS; ++u) t += a[u];
    if (t == magic) sink[0] = t;
}

// Two timing events, destroyed on every exit path.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EventPair() {
        SBH_HIP(hipEventCreate(&e0));
        const hipError_t r = hipEventCreate(&e1);
        if (r != hipSuccess) {
            (void)hipEventDestroy(e0);
            e0 = nullptr;
            SBH_HIP(r);
        }
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
};

// The best rate over launch shapes of 1, 2, 4 and 8 workgroups (4 .. 32 waves)
// per CU: few long streams per CU can read faster than a full occupancy.
double probe_stream_read(sbh_context* ctx, const void* x, int64_t bytes, int reps, unsigned* sink) {
    const int64_t count = bytes / 16;
    EventPair ev;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    int cus = 0;
    SBH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const probe_u32x4* px = static_cast<const probe_u32x4*>(x);
    double best = 0.0;
    for (int per_cu = 1; per_cu <= 8; per_cu *= 2) {
        const dim3 grid((unsigned)std::max(cus, 1) * per_cu), block(256);
        hipLaunchKernelGGL(stream_read_kernel, grid, block, 0, ctx->stream, px, count, 0x9e3779b9u, sink);
        SBH_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(stream_read_kernel, grid, block, 0, ctx->stream, px, count, 0x9e3779b9u, sink);
        SBH_HIP(hipGetLastError());
        SBH_HIP(hipEventRecord(e1, ctx->stream));
        SBH_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SBH_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::max(best, (double)count * 16.0 * reps / (ms * 1e-3) / 1e9);
    }
    return best;
}

double probe_fma_rate(sbh_context* ctx, double* sink) {
    int cus = 0;
    SBH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const dim3 grid((unsigned)std::max(cus, 1) * 16), block(256);
    const int64_t iters = 1 << 14;
    EventPair ev;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    hipLaunchKernelGGL(fma_rate_kernel, grid, block, 0, ctx->stream, iters, 0.999999, 1e-7, -1.0, sink);
    SBH_HIP(hipEventRecord(e0, ctx->stream));
    const int reps = 3;
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(fma_rate_kernel, grid, block, 0, ctx->stream, iters, 0.999999, 1e-7, -1.0, sink);
    SBH_HIP(hipGetLastError());
    SBH_HIP(hipEventRecord(e1, ctx->stream));
    SBH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SBH_HIP(hipEventElapsedTime(&ms, e0, e1));
    const double flops = 2.0 * FMA_CHAINS * (double)iters * grid.x * block.x * reps;
    return flops / (ms * 1e-3) / 1e12;
}

void device_copy(sbh_context* ctx, const double2* src, double2* dst, int64_t n) {
    hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, ctx->stream, src, dst, n);
    SBH_HIP(hipGetLastError());
}

void extract_inverse_diagonal(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, double2* d) {
    hipLaunchKernelGGL(inv_diag_kernel, dim3(blocks_for(n)), dim3(BLOCK), 0, ctx->stream, A, lda, n, d);
    SBH_HIP(hipGetLastError());
}

}  // namespace sbh
