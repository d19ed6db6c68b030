// Device kernels of the shared-basis block GMRES (block_krylov.cpp), which
// restates block_gmres_inner (crates/sonium-solvers/crates/
// sonium-solvers-iterative/src/block_gmres.rs:666-1120).
//
// Layout.  The block Krylov basis is column-major as in the reference:
// vector i of length n at V + i*n (complex FP64, interleaved).  The new
// Arnoldi block W (sj vectors) uses the same layout.  Right-hand sides and
// solutions are n x s row-major (one right-hand side per column).
//
// Block Gram G = V^H W (block_gram_into, :474-499) and the combinations
// W -= V G (block_project_into, :505-530) and X += V Y (apply_block_update,
// :545-565) run on v_mfma_f64_16x16x4_f64 in the real formulation:
//   Gram:     G~ (p x 2sj) = V~ (p x 2n) * Bt (2n x 2sj), V~ the interleaved
//             rows of V, and per complex entry w of W the 2x2 real block
//             [[Re w, Im w], [Im w, -Re w]] (columns Re g, Im g);
//   combine:  O~ (sj x 2n) = C~ (sj x 2p) * Bv (2p x 2n), C~ the rows
//             [Re c, Im c] of the coefficients, and per basis entry v the
//             block [[Re v, Im v], [-Im v, Re v]].
// Both operands of every MFMA come from LDS tiles that the global loads fill
// row-contiguously.  MFMA operand/result layout of the 16x16x4 f64 form:
// A[m][k] from lane m + 16k, B[k][n] from lane n + 16k, result register v of
// lane l holds row 4v + l/16, column l%16.
//
// The in-block QR with rank-drop deflation (CGS2 per column, :992-1025) is a
// chain of small kernels over row chunks; the accept/deflate decision is taken
// on the device from reductions summed in a fixed order, so every workgroup
// takes the same decision and the host reads the outcome once per block step.

#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.hpp"

namespace sbh {

namespace {

constexpr int BK_THREADS = 256;
constexpr int BK_XCD = 8;          // XCDs of the MI355X; K slices are grouped per XCD
constexpr int BK_BM = 64;          // basis vectors per Gram tile
constexpr int BK_STAGE = 16;       // complex K entries per LDS stage
constexpr int BK_CK = 64;          // complex output entries per combine tile
constexpr int QR_CHUNK = 256;      // rows per workgroup in the in-block QR chain (one per thread)
constexpr int QR_MAXC = 64;        // block width limit
constexpr int QR_SLOTS = QR_MAXC + 1;  // dots + squared norm per chunk

typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// ---------------------------------------------------------------- Gram

// part[(slice * p + i) * 2sj + 2c + {0,1}] = sum over the slice's rows k of
// conj(V[i][k]) W[c][k] (re, im).  Blocks: (basis tile, K slice); the K
// slices of one XCD (blockIdx % 8) are shared by consecutive basis tiles so
// the W slice stays in that XCD's L2.
template <int NT>
__global__ __launch_bounds__(BK_THREADS) void bk_gram_kernel(const double2* __restrict__ V, int64_t ldv, int p,
                                                             const double2* __restrict__ W, int64_t ldw, int sj,
                                                             int64_t n, int64_t kslice, int itiles,
                                                             double* __restrict__ part, const int* __restrict__ flag) {
    if (flag && *flag == 0) return;
    constexpr int NR = 16 * NT;  // real output columns (>= 2 sj)
    __shared__ double As[2 * BK_STAGE][BK_BM + 2];
    __shared__ __attribute__((aligned(16))) double Bs[2 * BK_STAGE][NR + 2];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int xcd = blockIdx.x % BK_XCD;
    const int local = blockIdx.x / BK_XCD;
    const int itile = local % itiles;
    const int slice = (local / itiles) * BK_XCD + xcd;
    const int i0 = itile * BK_BM;
    const int64_t kbeg = (int64_t)slice * kslice;
    const int64_t kend = min(n, kbeg + kslice);
    f64x4 acc[NT];
#pragma unroll
    for (int b = 0; b < NT; ++b) acc[b] = f64x4{0.0, 0.0, 0.0, 0.0};
    constexpr int AQ = BK_BM * BK_STAGE / BK_THREADS;                  // 4
    constexpr int BQ = (8 * NT * BK_STAGE + BK_THREADS - 1) / BK_THREADS;
    double2 pa[AQ], pb[BQ];
    auto gload = [&](int64_t k0) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = t + q * BK_THREADS;
            const int r = e >> 4, kk = e & 15;
            pa[q] = (i0 + r < p && k0 + kk < kend) ? V[(int64_t)(i0 + r) * ldv + k0 + kk] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int e = t + q * BK_THREADS;
            const int c = e >> 4, kk = e & 15;
            pb[q] = (c < sj && k0 + kk < kend) ? W[(int64_t)c * ldw + k0 + kk] : make_double2(0.0, 0.0);
        }
    };
    if (kbeg < kend) gload(kbeg);
    for (int64_t k0 = kbeg; k0 < kend; k0 += BK_STAGE) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = t + q * BK_THREADS;
            const int r = e >> 4, kk = e & 15;
            As[2 * kk][r] = pa[q].x;
            As[2 * kk + 1][r] = pa[q].y;
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int e = t + q * BK_THREADS;
            if (e < 8 * NT * BK_STAGE) {
                const int c = e >> 4, kk = e & 15;
                *reinterpret_cast<double2*>(&Bs[2 * kk][2 * c]) = pb[q];
                *reinterpret_cast<double2*>(&Bs[2 * kk + 1][2 * c]) = make_double2(pb[q].y, -pb[q].x);
            }
        }
        __syncthreads();
        if (k0 + BK_STAGE < kend) gload(k0 + BK_STAGE);
#pragma unroll
        for (int ks = 0; ks < (2 * BK_STAGE) / 4; ++ks) {
            const int kr = 4 * ks + (lane >> 4);
            const double av = As[kr][16 * w + (lane & 15)];
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                const double bv = Bs[kr][16 * b + (lane & 15)];
                acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int ncol = 2 * sj;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = i0 + 16 * w + 4 * v + (lane >> 4);
        if (i < p) {
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                const int col = 16 * b + (lane & 15);
                if (col < ncol) part[((int64_t)slice * p + i) * ncol + col] = acc[b][v];
            }
        }
    }
}

// gram[i*sj + c] = sum of the slice partials in slice order; hmode 1 sets and
// hmode 2 adds the result into hacc (the Hessenberg block h[0:p, p_start:+sj]).
__global__ void bk_gram_reduce_kernel(const double* __restrict__ part, int nslices, int p, int sj,
                                      double2* __restrict__ gram, double2* __restrict__ hacc, int hmode,
                                      const int* __restrict__ flag) {
    if (flag && *flag == 0) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)p * sj) return;
    const int64_t i = e / sj;
    const int c = (int)(e % sj);
    const int ncol = 2 * sj;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < nslices; ++s) {
        const double2 v = *reinterpret_cast<const double2*>(part + ((int64_t)s * p + i) * ncol + 2 * c);
        re += v.x;
        im += v.y;
    }
    const double2 g = make_double2(re, im);
    gram[e] = g;
    if (hmode == 1) {
        hacc[e] = g;
    } else if (hmode == 2) {
        const double2 h = hacc[e];
        hacc[e] = make_double2(h.x + g.x, h.y + g.y);
    }
}

// ---------------------------------------------------------------- combine

// O[c*osc + k*osk] += alpha * sum_{i<p} C[i*ldc + c] V[i*ldv + k] for c < sj,
// k < n.  A block owns 64 complex k (128 real output columns, 32 per wave)
// and every c; the K loop runs over the basis vectors in stages of 16.
template <int MT>
__global__ __launch_bounds__(BK_THREADS) void bk_combine_kernel(const double2* __restrict__ V, int64_t ldv, int p,
                                                                const double2* __restrict__ C, int ldc, int sj,
                                                                int64_t n, double2* __restrict__ O, int64_t osc,
                                                                int64_t osk, double alpha,
                                                                const int* __restrict__ flag) {
    if (flag && *flag == 0) return;
    constexpr int MC = 16 * MT;  // coefficient columns held (>= sj)
    __shared__ double Cs[2 * BK_STAGE][MC + 2];
    __shared__ __attribute__((aligned(16))) double Vs[2 * BK_STAGE][2 * BK_CK + 2];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int64_t k0 = (int64_t)blockIdx.x * BK_CK;
    f64x4 acc[MT][2];
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[m][0] = acc[m][1] = f64x4{0.0, 0.0, 0.0, 0.0};
    constexpr int VQ = BK_STAGE * BK_CK / BK_THREADS;  // 4
    constexpr int CQ = (BK_STAGE * MC + BK_THREADS - 1) / BK_THREADS;
    double2 pv[VQ], pc[CQ];
    auto gload = [&](int i0) {
#pragma unroll
        for (int q = 0; q < VQ; ++q) {
            const int e = t + q * BK_THREADS;
            const int ii = e >> 6, kk = e & 63;
            pv[q] = (i0 + ii < p && k0 + kk < n) ? V[(int64_t)(i0 + ii) * ldv + k0 + kk] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < CQ; ++q) {
            const int e = t + q * BK_THREADS;
            const int ii = e / MC, c = e % MC;
            pc[q] = (e < BK_STAGE * MC && i0 + ii < p && c < sj) ? C[(int64_t)(i0 + ii) * ldc + c]
                                                                  : make_double2(0.0, 0.0);
        }
    };
    gload(0);
    for (int i0 = 0; i0 < p; i0 += BK_STAGE) {
#pragma unroll
        for (int q = 0; q < VQ; ++q) {
            const int e = t + q * BK_THREADS;
            const int ii = e >> 6, kk = e & 63;
            *reinterpret_cast<double2*>(&Vs[2 * ii][2 * kk]) = pv[q];
            *reinterpret_cast<double2*>(&Vs[2 * ii + 1][2 * kk]) = make_double2(-pv[q].y, pv[q].x);
        }
#pragma unroll
        for (int q = 0; q < CQ; ++q) {
            const int e = t + q * BK_THREADS;
            if (e < BK_STAGE * MC) {
                const int ii = e / MC, c = e % MC;
                Cs[2 * ii][c] = pc[q].x;
                Cs[2 * ii + 1][c] = pc[q].y;
            }
        }
        __syncthreads();
        if (i0 + BK_STAGE < p) gload(i0 + BK_STAGE);
#pragma unroll
        for (int ks = 0; ks < (2 * BK_STAGE) / 4; ++ks) {
            const int kr = 4 * ks + (lane >> 4);
            double bv[2];
#pragma unroll
            for (int nn = 0; nn < 2; ++nn) bv[nn] = Vs[kr][32 * w + 16 * nn + (lane & 15)];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const double av = Cs[kr][16 * m + (lane & 15)];
#pragma unroll
                for (int nn = 0; nn < 2; ++nn)
                    acc[m][nn] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[nn], acc[m][nn], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 16 * m + 4 * v + (lane >> 4);
            if (c >= sj) continue;
#pragma unroll
            for (int nn = 0; nn < 2; ++nn) {
                const int col = 32 * w + 16 * nn + (lane & 15);
                const int64_t k = k0 + (col >> 1);
                if (k < n) {
                    double* o = reinterpret_cast<double*>(O + c * osc + k * osk) + (col & 1);
                    *o = *o + alpha * acc[m][nn][v];
                }
            }
        }
}

// ---------------------------------------------------------------- layout and reductions

// out[k*cols + c] = in[c*ld + k] (column-major -> row-major), 16 x 64 tiles
__global__ void bk_to_rowmajor_kernel(const double2* __restrict__ in, int64_t ld, int cols, int64_t n,
                                      double2* __restrict__ out) {
    __shared__ double2 tile[16][65];
    const int64_t k0 = (int64_t)blockIdx.x * 64;
    const int c0 = blockIdx.y * 16;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 16; r += 4) {
        const int c = c0 + r;
        tile[r][tx] = (c < cols && k0 + tx < n) ? in[(int64_t)c * ld + k0 + tx] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
        const int kk = e >> 4, cc = e & 15;
        if (c0 + cc < cols && k0 + kk < n) out[(k0 + kk) * cols + c0 + cc] = tile[cc][kk];
    }
}

// out[c*ld + k] = d[k] * in[k*cols + c] (row-major -> column-major; d may be null)
__global__ void bk_to_colmajor_kernel(const double2* __restrict__ in, int cols, int64_t n,
                                      const double2* __restrict__ d, double2* __restrict__ out, int64_t ld) {
    __shared__ double2 tile[16][65];
    const int64_t k0 = (int64_t)blockIdx.x * 64;
    const int c0 = blockIdx.y * 16;
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
        const int kk = e >> 4, cc = e & 15;
        double2 v = make_double2(0.0, 0.0);
        if (c0 + cc < cols && k0 + kk < n) {
            v = in[(k0 + kk) * cols + c0 + cc];
            if (d) v = cmul(d[k0 + kk], v);
        }
        tile[cc][kk] = v;
    }
    __syncthreads();
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 16; r += 4) {
        const int c = c0 + r;
        if (c < cols && k0 + tx < n) out[(int64_t)c * ld + k0 + tx] = tile[r][tx];
    }
}

// R[c*n + k] = d[k] * (B[k*s + c] - AX[k*s + c]) (AX may be null: zero)
__global__ void bk_residual_cm_kernel(const double2* __restrict__ B, const double2* __restrict__ AX, int s,
                                      int64_t n, const double2* __restrict__ d, double2* __restrict__ R) {
    __shared__ double2 tile[16][65];
    const int64_t k0 = (int64_t)blockIdx.x * 64;
    const int c0 = blockIdx.y * 16;
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
        const int kk = e >> 4, cc = e & 15;
        double2 v = make_double2(0.0, 0.0);
        if (c0 + cc < s && k0 + kk < n) {
            const int64_t idx = (k0 + kk) * s + c0 + cc;
            v = B[idx];
            if (AX) {
                const double2 a = AX[idx];
                v = make_double2(v.x - a.x, v.y - a.y);
            }
            if (d) v = cmul(d[k0 + kk], v);
        }
        tile[cc][kk] = v;
    }
    __syncthreads();
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 16; r += 4) {
        const int c = c0 + r;
        if (c < s && k0 + tx < n) R[(int64_t)c * n + k0 + tx] = tile[r][tx];
    }
}

// Squared norms of `cols` contiguous vectors of length n: part[chunk*cols + c]
// over rows [chunk*len, ...), each thread in row order, then a fixed tree.
__global__ __launch_bounds__(BK_THREADS) void bk_colnorm_partials_kernel(const double2* __restrict__ X, int64_t ld,
                                                                         int cols, int64_t n, int64_t len,
                                                                         double* __restrict__ part) {
    __shared__ double red[BK_THREADS];
    const int c = blockIdx.y;
    const int64_t b = (int64_t)blockIdx.x * len;
    const int64_t e = min(n, b + len);
    double s = 0.0;
    for (int64_t k = b + threadIdx.x; k < e; k += BK_THREADS) {
        const double2 v = X[(int64_t)c * ld + k];
        s += v.x * v.x + v.y * v.y;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = BK_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * cols + c] = red[0];
}

// out[c] = sqrt(sum over chunks in chunk order of part[chunk*cols + c])
__global__ void bk_sum_sqrt_kernel(const double* __restrict__ part, int chunks, int cols, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    double s = 0.0;
    for (int q = 0; q < chunks; ++q) s += part[(int64_t)q * cols + c];
    out[c] = sqrt(s);
}

// DGKS trigger on the block Frobenius norm (block_gmres.rs:962-987):
// flag = ||W_after||_F < 0.707 ||W_before||_F, from squared-norm partials.
__global__ void bk_dgks_flag_kernel(const double* __restrict__ before, const double* __restrict__ after, int chunks,
                                    int* __restrict__ flag, double* __restrict__ frob) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double b = 0.0, a = 0.0;
    for (int q = 0; q < chunks; ++q) b += before[q];
    for (int q = 0; q < chunks; ++q) a += after[q];
    const double fb = sqrt(b), fa = sqrt(a);
    frob[0] = fb;
    frob[1] = fa;
    *flag = (fa >= 0.707 * fb) ? 0 : 1;
}

// Per column of n x s row-major blocks: ||B - AX||^2, ||AX||^2, ||B||^2
// partials (block_true_metrics, block_gmres.rs:567-612) into part[chunk][3s].
__global__ __launch_bounds__(BK_THREADS) void bk_true_partials_kernel(const double2* __restrict__ B,
                                                                      const double2* __restrict__ AX, int s,
                                                                      int64_t n, int64_t len,
                                                                      double* __restrict__ part) {
    __shared__ double red[3][BK_THREADS];
    const int c = blockIdx.y;
    const int64_t b = (int64_t)blockIdx.x * len;
    const int64_t e = min(n, b + len);
    double r2 = 0.0, a2 = 0.0, b2 = 0.0;
    for (int64_t k = b + threadIdx.x; k < e; k += BK_THREADS) {
        const double2 bb = B[k * s + c], aa = AX[k * s + c];
        const double dr = bb.x - aa.x, di = bb.y - aa.y;
        r2 += dr * dr + di * di;
        a2 += aa.x * aa.x + aa.y * aa.y;
        b2 += bb.x * bb.x + bb.y * bb.y;
    }
    red[0][threadIdx.x] = r2;
    red[1][threadIdx.x] = a2;
    red[2][threadIdx.x] = b2;
    __syncthreads();
    for (int h = BK_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[((int64_t)blockIdx.x * 3 + q) * s + c] = red[q][0];
}

// ---------------------------------------------------------------- in-block QR chain

// Device state of one block QR (initial block or a new Arnoldi block).
struct QrState {
    int snext[QR_MAXC + 1];   // accepted directions before column c (snext[0] = 0)
    int accepted[QR_MAXC];    // slot of column c, or -1 if deflated
    unsigned counter;         // finished workgroups of the running launch
    int pad[2];
    double orig[QR_MAXC];     // ||w_c|| before the passes
    double wnorm[QR_MAXC];    // ||w_c|| after the passes
    double2 r[QR_MAXC][QR_MAXC];  // h[p + a][p_start + c] += d1[a] + d2[a]  (a < snext[c])
    double2 coef[QR_MAXC];    // coefficients of the pass in progress
};

struct QrArgs {
    double2* W;       // sj column-major vectors of length n
    double2* Q;       // the block's accepted directions: slot a at Q + a*n
    int64_t n;
    int sj;
    int budget;       // accepted directions allowed in this block
    double defl_tol;  // sqrt(eps)
    QrState* st;
    double2* part;    // chunks x QR_SLOTS partial sums (dots in slots 0..63, ||w||^2 in slot 64)
};

// True in the workgroup that finishes last: each workgroup publishes its
// partials, then takes a ticket; the last one sums every chunk's partials
// (in chunk order) and clears the counter for the next launch.
__device__ bool qr_last_block(unsigned* counter) {
    __shared__ bool last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
    __syncthreads();
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return last;
}

// out[v] = sum over chunks of part[chunk][v] for v < cnt and, with the norm,
// v = QR_MAXC: four strided quarter sums, added in order.
__device__ void qr_sum_chunks(const double2* __restrict__ part, int chunks, int cnt, bool with_norm, double2* out,
                              double2 (*quarter)[4]) {
    const int t = threadIdx.x;
    const int items = 4 * (cnt + (with_norm ? 1 : 0));
    for (int w = t; w < items; w += BK_THREADS) {
        const int idx = w >> 2, q = w & 3;
        const int v = idx < cnt ? idx : QR_MAXC;
        double re = 0.0, im = 0.0;
        for (int ch = q; ch < chunks; ch += 4) {
            const double2 x = part[(int64_t)ch * QR_SLOTS + v];
            re += x.x;
            im += x.y;
        }
        quarter[v][q] = make_double2(re, im);
    }
    __syncthreads();
    if (t < cnt || (with_norm && t == QR_MAXC)) {
        double re = quarter[t][0].x, im = quarter[t][0].y;
        for (int q = 1; q < 4; ++q) {
            re += quarter[t][q].x;
            im += quarter[t][q].y;
        }
        out[t] = make_double2(re, im);
    }
    __syncthreads();
}

// This chunk's partials of conj(q_a) . w (a < cnt; the LDS copy ql replaces
// slot cnt - 1 when given) and ||w||^2, w in LDS.  Thread (a = t % 64,
// segment = t / 64) sums a quarter of the chunk in row order; the quarters
// are added in order.
__device__ void qr_chunk_dots(const double2* __restrict__ Q, int64_t n, int64_t r0, int len, int cnt,
                              const double2* wl, const double2* ql, double2* __restrict__ part_out,
                              double2 (*seg)[4], double* red) {
    const int t = threadIdx.x;
    const int a = t & 63, sg = t >> 6;
    const int slen = (len + 3) / 4;
    const int b = sg * slen, e = min(len, b + slen);
    if (a < cnt) {
        double re = 0.0, im = 0.0;
        const bool from_lds = ql && a == cnt - 1;
        const double2* q = Q + (int64_t)a * n + r0;
#pragma unroll 8
        for (int k = b; k < e; ++k) {
            const double2 qa = from_lds ? ql[k] : q[k];
            const double2 d = cmulc(qa, wl[k]);
            re += d.x;
            im += d.y;
        }
        seg[a][sg] = make_double2(re, im);
    }
    double s2 = 0.0;
    for (int k = t; k < len; k += BK_THREADS) s2 += wl[k].x * wl[k].x + wl[k].y * wl[k].y;
    red[t] = s2;
    __syncthreads();
    for (int h = BK_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t < cnt) {
        double re = seg[t][0].x, im = seg[t][0].y;
        for (int q = 1; q < 4; ++q) {
            re += seg[t][q].x;
            im += seg[t][q].y;
        }
        part_out[t] = make_double2(re, im);
    }
    if (t == 0) part_out[QR_MAXC] = make_double2(red[0], 0.0);
}

// Start of column c (c == sj: finish only): the previous column's direction
// q = w * (1/||w||) if it was accepted, then pass-1 partials of column c;
// the last workgroup turns them into d1 (the first h entries) and ||w_c||.
__global__ __launch_bounds__(BK_THREADS) void bk_qr_start_kernel(QrArgs a, int c) {
    __shared__ double2 wl[QR_CHUNK];
    __shared__ double2 ql[QR_CHUNK];
    __shared__ double2 seg[QR_MAXC][4];
    __shared__ double2 quarter[QR_SLOTS][4];
    __shared__ double2 sums[QR_SLOTS];
    __shared__ double red[BK_THREADS];
    const int t = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * QR_CHUNK;
    const int len = (int)min((int64_t)QR_CHUNK, a.n - r0);
    bool fresh = false;
    if (c > 0) {
        const int slot = a.st->accepted[c - 1];
        if (slot >= 0) {
            const double inv = 1.0 / a.st->wnorm[c - 1];
            const double2* wp = a.W + (int64_t)(c - 1) * a.n + r0;
            double2* q = a.Q + (int64_t)slot * a.n + r0;
            for (int k = t; k < len; k += BK_THREADS) {
                const double2 v = wp[k];
                const double2 qv = make_double2(v.x * inv, v.y * inv);
                q[k] = qv;
                ql[k] = qv;
            }
            fresh = true;
        }
    }
    if (c >= a.sj) return;
    const int cnt = c == 0 ? 0 : a.st->snext[c];
    const double2* wc = a.W + (int64_t)c * a.n + r0;
    for (int k = t; k < len; k += BK_THREADS) wl[k] = wc[k];
    __syncthreads();
    qr_chunk_dots(a.Q, a.n, r0, len, cnt, wl, fresh ? ql : nullptr, a.part + (int64_t)blockIdx.x * QR_SLOTS, seg,
                  red);
    if (!qr_last_block(&a.st->counter)) return;
    qr_sum_chunks(a.part, gridDim.x, cnt, true, sums, quarter);
    if (t < cnt) {
        a.st->r[c][t] = sums[t];
        a.st->coef[t] = sums[t];
    }
    if (t == 0) {
        if (c == 0) a.st->snext[0] = 0;
        a.st->orig[c] = sqrt(sums[QR_MAXC].x);
        a.st->counter = 0;
    }
}

// w_c -= sum_a coef[a] q_a (the reference's axpy order), then (second) the
// pass-2 partials, whose sums become coef and are added into the h entries,
// or (third) the partial ||w_c||^2, whose square root decides acceptance:
// wn > defl_tol * max(orig, tiny) and room left in the block
// (block_gmres.rs:1014-1024).
__global__ __launch_bounds__(BK_THREADS) void bk_qr_project_kernel(QrArgs a, int c, int third) {
    __shared__ double2 wl[QR_CHUNK];
    __shared__ double2 d[QR_MAXC];
    __shared__ double2 seg[QR_MAXC][4];
    __shared__ double2 quarter[QR_SLOTS][4];
    __shared__ double2 sums[QR_SLOTS];
    __shared__ double red[BK_THREADS];
    const int t = threadIdx.x;
    const int cnt = a.st->snext[c];
    const int64_t r0 = (int64_t)blockIdx.x * QR_CHUNK;
    const int len = (int)min((int64_t)QR_CHUNK, a.n - r0);
    if (t < cnt) d[t] = a.st->coef[t];
    __syncthreads();
    double2* wc = a.W + (int64_t)c * a.n + r0;
    for (int k = t; k < len; k += BK_THREADS) {
        double2 v = wc[k];
        for (int q = 0; q < cnt; ++q) {
            const double2 nd = make_double2(-d[q].x, -d[q].y);
            const double2 t2 = cmul(nd, a.Q[(int64_t)q * a.n + r0 + k]);
            v = make_double2(v.x + t2.x, v.y + t2.y);
        }
        wc[k] = v;
        wl[k] = v;
    }
    __syncthreads();
    qr_chunk_dots(a.Q, a.n, r0, len, third ? 0 : cnt, wl, nullptr, a.part + (int64_t)blockIdx.x * QR_SLOTS, seg,
                  red);
    if (!qr_last_block(&a.st->counter)) return;
    qr_sum_chunks(a.part, gridDim.x, third ? 0 : cnt, third != 0, sums, quarter);
    if (!third) {
        if (t < cnt) {
            const double2 h = a.st->r[c][t];
            a.st->r[c][t] = make_double2(h.x + sums[t].x, h.y + sums[t].y);
            a.st->coef[t] = sums[t];
        }
    } else if (t == 0) {
        const double wn = sqrt(sums[QR_MAXC].x);
        const bool accept = wn > a.defl_tol * fmax(a.st->orig[c], 1e-30) && cnt < a.budget;
        a.st->wnorm[c] = wn;
        a.st->accepted[c] = accept ? cnt : -1;
        a.st->snext[c + 1] = cnt + (accept ? 1 : 0);
    }
    if (t == 0) a.st->counter = 0;
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ---------------------------------------------------------------- launchers

int bk_gram_slices(int64_t n, int p) {
    const int itiles = (p + BK_BM - 1) / BK_BM;
    int groups = std::max(1, 256 / std::max(1, itiles));          // ~2048 blocks
    const int64_t max_groups = std::max<int64_t>(1, n / (BK_XCD * 256));  // slices of >= 256 rows
    groups = (int)std::min<int64_t>(groups, max_groups);
    return groups * BK_XCD;
}

// part: at least slices * p * 2sj doubles; gram: p*sj; hacc: p*sj (hmode 0 none, 1 set, 2 add)
void bk_gram(sbh_context* ctx, const double2* V, int64_t ldv, int p, const double2* W, int64_t ldw, int sj,
             int64_t n, double* part, double2* gram, double2* hacc, int hmode, const int* flag) {
    SBH_REQUIRE(sj >= 1 && sj <= QR_MAXC && p >= 1, "block Gram shape");
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_KRYLOV_VECTOR, &ev);
    const int slices = bk_gram_slices(n, p);
    const int itiles = (p + BK_BM - 1) / BK_BM;
    const int64_t kslice = ((n + slices - 1) / slices + BK_STAGE - 1) / BK_STAGE * BK_STAGE;
    const dim3 grid((unsigned)(slices * itiles));
    const int nt = sj <= 8 ? 1 : sj <= 16 ? 2 : sj <= 32 ? 4 : 8;
#define SBH_BG(NT_)                                                                                              \
    hipLaunchKernelGGL(bk_gram_kernel<NT_>, grid, dim3(BK_THREADS), 0, ctx->stream, V, ldv, p, W, ldw, sj, n, kslice, \
                       itiles, part, flag)
    if (nt == 1) SBH_BG(1);
    else if (nt == 2) SBH_BG(2);
    else if (nt == 4) SBH_BG(4);
    else SBH_BG(8);
#undef SBH_BG
    SBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(bk_gram_reduce_kernel, dim3(blocks_for((int64_t)p * sj, BK_THREADS)), dim3(BK_THREADS), 0,
                       ctx->stream, part, slices, p, sj, gram, hacc, hmode, flag);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_KRYLOV_VECTOR, ev);
}

void bk_combine(sbh_context* ctx, const double2* V, int64_t ldv, int p, const double2* C, int ldc, int sj, int64_t n,
                double2* O, int64_t osc, int64_t osk, double alpha, const int* flag) {
    SBH_REQUIRE(sj >= 1 && sj <= QR_MAXC, "block combination width");
    if (p <= 0) return;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_KRYLOV_VECTOR, &ev);
    const dim3 grid(blocks_for(n, BK_CK));
    const int mt = (sj + 15) / 16;
#define SBH_BC(MT_)                                                                                             \
    hipLaunchKernelGGL(bk_combine_kernel<MT_>, grid, dim3(BK_THREADS), 0, ctx->stream, V, ldv, p, C, ldc, sj, n, O, \
                       osc, osk, alpha, flag)
    if (mt == 1) SBH_BC(1);
    else if (mt == 2) SBH_BC(2);
    else if (mt == 3) SBH_BC(3);
    else SBH_BC(4);
#undef SBH_BC
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_KRYLOV_VECTOR, ev);
}

void bk_to_rowmajor(sbh_context* ctx, const double2* in, int64_t ld, int cols, int64_t n, double2* out) {
    hipLaunchKernelGGL(bk_to_rowmajor_kernel, dim3(blocks_for(n, 64), blocks_for(cols, 16)), dim3(BK_THREADS), 0,
                       ctx->stream, in, ld, cols, n, out);
    SBH_HIP(hipGetLastError());
}

void bk_to_colmajor(sbh_context* ctx, const double2* in, int cols, int64_t n, const double2* d, double2* out,
                    int64_t ld) {
    hipLaunchKernelGGL(bk_to_colmajor_kernel, dim3(blocks_for(n, 64), blocks_for(cols, 16)), dim3(BK_THREADS), 0,
                       ctx->stream, in, cols, n, d, out, ld);
    SBH_HIP(hipGetLastError());
}

void bk_residual_cm(sbh_context* ctx, const double2* B, const double2* AX, int s, int64_t n, const double2* d,
                    double2* R) {
    hipLaunchKernelGGL(bk_residual_cm_kernel, dim3(blocks_for(n, 64), blocks_for(s, 16)), dim3(BK_THREADS), 0,
                       ctx->stream, B, AX, s, n, d, R);
    SBH_HIP(hipGetLastError());
}

int bk_norm_chunks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(256, (n + 4095) / 4096)); }

// out[c] = ||X_c|| for cols contiguous vectors (stride ld); part: chunks * cols doubles
void bk_colnorms(sbh_context* ctx, const double2* X, int64_t ld, int cols, int64_t n, double* part, double* out) {
    const int chunks = bk_norm_chunks(n);
    const int64_t len = (n + chunks - 1) / chunks;
    hipLaunchKernelGGL(bk_colnorm_partials_kernel, dim3(chunks, cols), dim3(BK_THREADS), 0, ctx->stream, X, ld, cols,
                       n, len, part);
    SBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(bk_sum_sqrt_kernel, dim3(blocks_for(cols, 64)), dim3(64), 0, ctx->stream, part, chunks, cols,
                       out);
    SBH_HIP(hipGetLastError());
}

// squared Frobenius partials of one contiguous block (count complex values)
int bk_frob_partials(sbh_context* ctx, const double2* X, int64_t count, double* part) {
    const int chunks = bk_norm_chunks(count);
    const int64_t len = (count + chunks - 1) / chunks;
    hipLaunchKernelGGL(bk_colnorm_partials_kernel, dim3(chunks, 1), dim3(BK_THREADS), 0, ctx->stream, X, count, 1,
                       count, len, part);
    SBH_HIP(hipGetLastError());
    return chunks;
}

void bk_dgks_flag(sbh_context* ctx, const double* before, const double* after, int chunks, int* flag, double* frob) {
    hipLaunchKernelGGL(bk_dgks_flag_kernel, dim3(1), dim3(64), 0, ctx->stream, before, after, chunks, flag, frob);
    SBH_HIP(hipGetLastError());
}

// part: chunks * 3s doubles; returns the chunk count (partials to be summed by the caller's kernel)
void bk_true_metrics(sbh_context* ctx, const double2* B, const double2* AX, int s, int64_t n, double* part,
                     double* out) {
    const int chunks = bk_norm_chunks(n);
    const int64_t len = (n + chunks - 1) / chunks;
    hipLaunchKernelGGL(bk_true_partials_kernel, dim3(chunks, s), dim3(BK_THREADS), 0, ctx->stream, B, AX, s, n, len,
                       part);
    SBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(bk_sum_sqrt_kernel, dim3(blocks_for(3 * s, 64)), dim3(64), 0, ctx->stream, part, chunks, 3 * s,
                       out);
    SBH_HIP(hipGetLastError());
}

size_t bk_qr_state_bytes() { return sizeof(QrState); }
int bk_qr_chunks(int64_t n) { return (int)((n + QR_CHUNK - 1) / QR_CHUNK); }
size_t bk_qr_partial_elems(int64_t n) { return (size_t)bk_qr_chunks(n) * QR_SLOTS; }

// Block QR with rank-drop deflation of the sj column-major vectors W into the
// directions Q (slots 0..); state (its counter zero between launches) and
// partials on the device.  Three launches per column plus one to write the
// last direction.
void bk_block_qr(sbh_context* ctx, double2* W, double2* Q, int64_t n, int sj, int budget, void* state,
                 double2* part) {
    SBH_REQUIRE(sj >= 1 && sj <= QR_MAXC, "block QR width");
    QrArgs a;
    a.W = W;
    a.Q = Q;
    a.n = n;
    a.sj = sj;
    a.budget = budget;
    a.defl_tol = std::sqrt(2.220446049250313e-16);
    a.st = static_cast<QrState*>(state);
    a.part = part;
    const int chunks = bk_qr_chunks(n);
    const dim3 grid((unsigned)chunks);
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_KRYLOV_VECTOR, &ev);
    for (int c = 0; c <= sj; ++c) {
        hipLaunchKernelGGL(bk_qr_start_kernel, grid, dim3(BK_THREADS), 0, ctx->stream, a, c);
        if (c == sj) break;
        hipLaunchKernelGGL(bk_qr_project_kernel, grid, dim3(BK_THREADS), 0, ctx->stream, a, c, 0);
        hipLaunchKernelGGL(bk_qr_project_kernel, grid, dim3(BK_THREADS), 0, ctx->stream, a, c, 1);
    }
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_KRYLOV_VECTOR, ev);
}

}  // namespace sbh
