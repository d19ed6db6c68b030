// Kernels of the batched solver: many independent instances (each an n x n
// complex system with ncol <= 16 right-hand sides) advanced in lock step.
//
// Layout.  Instance b's vectors are n x ncol row-major blocks at vec_off of a
// packed vector of length L = sum n*ncol; global column c = col_off + q.
// Vector kernels run over chunks (instance, rows [r0, r0 + rows)) of about
// 4096 elements; a thread owns local column t % ncol.  Per-column partial sums
// go to part[(chunk * MV + slot) * 16 + q] and are added per column over the
// instance's chunks in chunk order, so every reduction has a fixed order.
//
// Products.  Grouped tiles (instance, 128 rows): the FP64 product uses
// v_mfma_f64_16x16x4_f64 and the FP16/BF16 product v_mfma_f32_32x32x16 with
// FP32 accumulation; both use the real formulation of the complex product
// (A as an n x 2n real matrix, the block operand's rows [Re, Im] and
// [-Im, Re] formed while it is copied into LDS).
//
// Arnoldi state per column: Hessenberg column j at H[(c * m + j) * (m + 1)],
// rotations cs/sn[c * m + i], g[c * (m + 1) + i].  The small per-column
// work (Givens rotations, back substitution, restart) runs on the device, so
// a whole restart cycle is a fixed sequence of launches that can be captured
// in one graph.  Kernels skip instances whose flag is clear.

#include "batch.hpp"

namespace sbh {
namespace {

#include "lowp_common.hpp"

constexpr int MAXCOL = 16;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double cabs2(double2 a) { return hypot(a.x, a.y); }
__device__ __forceinline__ double2 cinv(double2 a) {
    const double n = a.x * a.x + a.y * a.y;
    return make_double2(a.x / n, -a.y / n);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// ---- grouped products ----

constexpr int GP_BM = 128;
constexpr int GLP_ROW = 128 + 8;  // LDS row (halves) of the transposed block operand

// Y_b = diag(row_inv) (A_lp,b X_b) on the instance's 128-row tile.  X holds
// Krylov basis vectors (unit 2-norm per column), so its entries are converted
// to the format without a scale; the row-scaled A keeps |A_ij| < 2^15.  Each
// stage of 128 real k moves the tile's rows of A (16 lanes per 256-byte row
// segment) and the formed operand rows into LDS; the next stage is loaded
// into registers while the current one runs on MFMA.
template <int FMT>
__global__ __launch_bounds__(BLOCK) void glp_gemm_kernel(const int2* __restrict__ tiles,
                                                         const BatchInstance* __restrict__ inst,
                                                         const int* __restrict__ flag, const u16* __restrict__ Alp,
                                                         const double* __restrict__ row_inv,
                                                         const double2* __restrict__ X, double2* __restrict__ Y) {
    __shared__ __attribute__((aligned(16))) u16 As[GP_BM * GLP_ROW];
    __shared__ __attribute__((aligned(16))) u16 Bs[2 * MAXCOL * GLP_ROW];
    const int2 tile = tiles[blockIdx.x];
    if (!flag[tile.x]) return;  // uniform over the block
    const BatchInstance& I = inst[tile.x];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t n = I.n, ld = I.ld;
    const int kb = I.ncol;
    const int64_t row0 = tile.y;
    const int ar = threadIdx.x >> 4;
    const int aq = threadIdx.x & 15;
    const u16* arow[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) arow[r] = Alp + I.lp_off + min(row0 + ar + 16 * r, n - 1) * 2 * ld + aq * 8;
    const double2* Xb = X + I.vec_off;
    // operand stage: complex rows j0 .. j0+63 of X_b, 16 columns; thread takes 4 (row, column) pairs
    u32x4 ra[8];
    double2 rx[4];
    auto load = [&](int64_t kc) {
#pragma unroll
        for (int r = 0; r < 8; ++r) ra[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(arow[r] + kc));
        const int64_t j0 = kc / 2;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = threadIdx.x + i * BLOCK;
            const int jl = e / MAXCOL, c = e % MAXCOL;
            const int64_t j = j0 + jl;
            rx[i] = (j < n && c < kb) ? Xb[j * kb + c] : make_double2(0.0, 0.0);
        }
    };
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    const int kq = (lane >> 5) * 8;
    load(0);
    for (int64_t kc = 0; kc < 2 * ld; kc += 128) {
#pragma unroll
        for (int r = 0; r < 8; ++r) *reinterpret_cast<u32x4*>(&As[(ar + 16 * r) * GLP_ROW + aq * 8]) = ra[r];
        // operand column 2c is Re, 2c+1 is Im of output column c:
        // rows (2j, 2j+1) = (Re x, -Im x) and (Im x, Re x)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = threadIdx.x + i * BLOCK;
            const int jl = e / MAXCOL, c = e % MAXCOL;
            const unsigned xr = Half<FMT>::from_float((float)rx[i].x);
            const unsigned xi = Half<FMT>::from_float((float)rx[i].y);
            const unsigned nxi = Half<FMT>::from_float((float)-rx[i].y);
            *reinterpret_cast<unsigned*>(&Bs[(2 * c) * GLP_ROW + 2 * jl]) = xr | (nxi << 16);
            *reinterpret_cast<unsigned*>(&Bs[(2 * c + 1) * GLP_ROW + 2 * jl]) = xi | (xr << 16);
        }
        __syncthreads();
        if (kc + 128 < 2 * ld) load(kc + 128);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const u32x4 a = *reinterpret_cast<const u32x4*>(&As[(wave * 32 + (lane & 31)) * GLP_ROW + s * 16 + kq]);
            const u32x4 b = *reinterpret_cast<const u32x4*>(&Bs[(lane & 31) * GLP_ROW + s * 16 + kq]);
            acc = mfma32<FMT>(a, b, acc);
        }
        __syncthreads();
    }
    // accumulator layout: column lane % 32, row 8*(v/4) + 4*(lane/32) + v%4
    const int64_t row_base = row0 + wave * 32;
    const int col = lane & 31;
    const int c = col >> 1;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const float mine = acc[v];
        const float other = __shfl_xor(mine, 1, 64);
        const int64_t row = row_base + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if ((col & 1) == 0 && c < kb && row < n) {
            const double s = row_inv[I.panel_off + row];
            Y[I.vec_off + row * kb + c] = make_double2((double)mine * s, (double)other * s);
        }
    }
}

// FP64: Y_b = A_b X_b on the instance's 128-row tile, 8*NT complex columns
// (16*NT real) per tile; wave w owns rows 32w..32w+31 as 2 x NT tiles of 16 x 16.
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int G64_BKC = 8;           // complex columns of A per stage
constexpr int G64_AROW = GP_BM + 2;  // LDS row of the transposed A stage

template <int NT>
__global__ __launch_bounds__(BLOCK) void g64_gemm_kernel(const int2* __restrict__ tiles,
                                                         const BatchInstance* __restrict__ inst,
                                                         const int* __restrict__ flag, const double2* __restrict__ A,
                                                         const double2* __restrict__ X, double2* __restrict__ Y) {
    __shared__ double As[2 * G64_BKC][G64_AROW];
    __shared__ __attribute__((aligned(16))) double Xs[2 * G64_BKC][16 * NT];
    const int2 tile = tiles[blockIdx.x];
    if (!flag[tile.x]) return;
    const BatchInstance& I = inst[tile.x];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int64_t n = I.n;
    const int kb = I.ncol;
    const int64_t row0 = tile.y;
    const double2* Ab = A + I.a_off;
    const double2* Xb = X + I.vec_off;
    f64x4 acc[2][NT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
    // Stage s+1 is loaded into registers while stage s is multiplied.
    constexpr int AQ = (GP_BM * G64_BKC) / BLOCK;
    double2 pa[AQ], px;
    auto gload = [&](int64_t k0) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = t + q * BLOCK;
            const int r = e / G64_BKC, cc = e % G64_BKC;
            const int64_t gr = row0 + r, gk = k0 + cc;
            pa[q] = (gr < n && gk < n) ? Ab[gr * n + gk] : make_double2(0.0, 0.0);
        }
        const int j = t / (8 * NT), c = t % (8 * NT);
        const int64_t gk = k0 + j;
        px = (t < G64_BKC * 8 * NT && gk < n && c < kb) ? Xb[gk * kb + c] : make_double2(0.0, 0.0);
    };
    gload(0);
    for (int64_t k0 = 0; k0 < n; k0 += G64_BKC) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = t + q * BLOCK;
            const int r = e / G64_BKC, cc = e % G64_BKC;
            As[2 * cc][r] = pa[q].x;
            As[2 * cc + 1][r] = pa[q].y;
        }
        if (t < G64_BKC * 8 * NT) {
            const int j = t / (8 * NT), c = t % (8 * NT);
            *reinterpret_cast<double2*>(&Xs[2 * j][2 * c]) = px;
            *reinterpret_cast<double2*>(&Xs[2 * j + 1][2 * c]) = make_double2(-px.y, px.x);
        }
        __syncthreads();
        if (k0 + G64_BKC < n) gload(k0 + G64_BKC);
#pragma unroll
        for (int ks = 0; ks < (2 * G64_BKC) / 4; ++ks) {
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
    // 16 x 16 tile of v_mfma_f64_16x16x4_f64: lane l, register v -> row 4*v + l/16,
    // real column l % 16 (A operand lane l = A[l%16][l/16], B lane l = B[l/16][l%16])
    const int col = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t gr = row0 + 32 * w + 16 * a + 4 * v + (lane >> 4);
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                const double mine = acc[a][b][v];
                const double other = __shfl_xor(mine, 1, 64);
                const int c = (16 * b + col) >> 1;
                if ((col & 1) == 0 && c < kb && gr < n) Y[I.vec_off + gr * kb + c] = make_double2(mine, other);
            }
        }
}

// Row-scaled FP16/BF16 copy of every instance matrix: one block per global row.
constexpr int CONV_BLOCK = 1024;

template <int FMT>
__global__ __launch_bounds__(CONV_BLOCK) void glp_convert_kernel(const int* __restrict__ row_inst,
                                                            const BatchInstance* __restrict__ inst,
                                                            const double2* __restrict__ A, u16* __restrict__ Alp,
                                                            double* __restrict__ row_inv) {
    const int64_t g = blockIdx.x;
    const BatchInstance& I = inst[row_inst[g]];
    const int64_t r = g - I.panel_off;
    convert_row<FMT, CONV_BLOCK>(A + I.a_off + r * I.n, I.n, I.ld, Alp + I.lp_off + r * 2 * I.ld, row_inv + g);
}

// ---- ragged vector kernels ----

// slot s < nvec: sum conj(V_s) W per column; slot nvec: ||W||^2.
// Rows per thread per pass of the vector reductions (loads kept in flight together).
constexpr int RB = 8;

__global__ __launch_bounds__(BLOCK) void rdots_kernel(const VecChunk* __restrict__ chunks,
                                                      const BatchInstance* __restrict__ inst,
                                                      const int* __restrict__ flag, const double2* __restrict__ V,
                                                      int64_t L, int nvec, const double2* __restrict__ W,
                                                      double2* __restrict__ part, int MV) {
    __shared__ double2 sh[BLOCK];
    const VecChunk ch = chunks[blockIdx.x];
    if (!flag[ch.inst]) return;
    const BatchInstance& I = inst[ch.inst];
    const int kb = I.ncol;
    const int S = BLOCK / kb;
    const int c = threadIdx.x % kb;
    const int sub = threadIdx.x / kb;
    const int slot = blockIdx.y;
    const int64_t base = I.vec_off + ch.r0 * kb + c;
    double2 s = make_double2(0.0, 0.0);
    if (sub < S) {
        // RB rows per pass, their loads issued together; sums in row order
        for (int r0 = sub; r0 < ch.rows; r0 += S * RB) {
            double2 y[RB];
#pragma unroll
            for (int t = 0; t < RB; ++t) {
                const int r = r0 + S * t;
                y[t] = r < ch.rows ? W[base + (int64_t)r * kb] : make_double2(0.0, 0.0);
            }
            if (slot < nvec) {
                const double2* v = V + (int64_t)slot * L;
                double2 x[RB];
#pragma unroll
                for (int t = 0; t < RB; ++t) {
                    const int r = r0 + S * t;
                    x[t] = r < ch.rows ? v[base + (int64_t)r * kb] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int t = 0; t < RB; ++t)
                    if (r0 + S * t < ch.rows) {
                        const double2 p = cmulc(x[t], y[t]);
                        s.x += p.x;
                        s.y += p.y;
                    }
            } else {
#pragma unroll
                for (int t = 0; t < RB; ++t)
                    if (r0 + S * t < ch.rows) s.x += y[t].x * y[t].x + y[t].y * y[t].y;
            }
        }
    }
    sh[threadIdx.x] = s;
    column_tree(sh, S, kb, sub, c);
    if (threadIdx.x < kb) part[((int64_t)blockIdx.x * MV + slot) * MAXCOL + threadIdx.x] = sh[threadIdx.x];
}

// Column sums of the dot partials (thread per (column, slot)).  pass 0:
// coef = h(:, j) = the sums, wn0 = ||W|| before projection.  pass 1 (second
// Gram–Schmidt pass): only columns flagged for reorthogonalisation take the
// sums (added to h); the others get coef = 0.
__global__ void rsum_dots_kernel(int C, const int* __restrict__ col_inst, const int* __restrict__ col_q,
                                 const int2* __restrict__ chunk_range, const double2* __restrict__ part, int MV,
                                 int nvec, int j, int m, const int* __restrict__ active,
                                 const int* __restrict__ reorth, int pass, double2* __restrict__ H,
                                 double2* __restrict__ coef, double* __restrict__ wn0) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)C * (nvec + 1)) return;
    const int c = (int)(t / (nvec + 1));
    const int slot = (int)(t % (nvec + 1));
    const bool take = pass == 0 ? active[c] != 0 : reorth[c] != 0;
    double2 s = make_double2(0.0, 0.0);
    if (take) {
        const int2 r = chunk_range[col_inst[c]];
        const int q = col_q[c];
        for (int ch = r.x; ch < r.y; ++ch) {
            const double2 p = part[((int64_t)ch * MV + slot) * MAXCOL + q];
            s.x += p.x;
            s.y += p.y;
        }
    }
    if (slot < nvec) {
        coef[(int64_t)c * MV + slot] = s;
        if (take) {
            double2& h = H[((int64_t)c * m + j) * (m + 1) + slot];
            h = pass == 0 ? s : cadd(h, s);
        }
    } else if (pass == 0 && take) {
        wn0[c] = sqrt(s.x);
    }
}

// W -= sum_{s < nvec} coef[c][s] V_s per column; ||W||^2 partials into slot 0.
__global__ __launch_bounds__(BLOCK) void rproject_kernel(const VecChunk* __restrict__ chunks,
                                                         const BatchInstance* __restrict__ inst,
                                                         const int* __restrict__ flag, const double2* __restrict__ V,
                                                         int64_t L, int nvec, const double2* __restrict__ coef, int MV,
                                                         double2* __restrict__ W, double2* __restrict__ part) {
    __shared__ double sh[BLOCK];
    const VecChunk ch = chunks[blockIdx.x];
    if (!flag[ch.inst]) return;
    const BatchInstance& I = inst[ch.inst];
    const int kb = I.ncol;
    const int S = BLOCK / kb;
    const int c = threadIdx.x % kb;
    const int sub = threadIdx.x / kb;
    const int64_t base = I.vec_off + ch.r0 * kb + c;
    const double2* cf = coef + (int64_t)(I.col_off + c) * MV;
    double nr = 0.0;
    if (sub < S) {
        // RB rows per pass: per vector their RB loads are issued together; per
        // element the sum over vectors runs in order before the subtraction
        for (int r0 = sub; r0 < ch.rows; r0 += S * RB) {
            double sr[RB], si[RB];
#pragma unroll
            for (int t = 0; t < RB; ++t) sr[t] = si[t] = 0.0;
            for (int s = 0; s < nvec; ++s) {
                const double2 cs = cf[s];
                const double2* v = V + (int64_t)s * L;
                double2 x[RB];
#pragma unroll
                for (int t = 0; t < RB; ++t) {
                    const int r = r0 + S * t;
                    x[t] = r < ch.rows ? v[base + (int64_t)r * kb] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int t = 0; t < RB; ++t) {
                    const double2 p = cmul(cs, x[t]);
                    sr[t] += p.x;
                    si[t] += p.y;
                }
            }
#pragma unroll
            for (int t = 0; t < RB; ++t) {
                const int r = r0 + S * t;
                if (r < ch.rows) {
                    const int64_t idx = base + (int64_t)r * kb;
                    double2 w = W[idx];
                    w.x -= sr[t];
                    w.y -= si[t];
                    W[idx] = w;
                    nr += w.x * w.x + w.y * w.y;
                }
            }
        }
    }
    sh[threadIdx.x] = nr;
    column_tree(sh, S, kb, sub, c);
    if (threadIdx.x < kb) part[((int64_t)blockIdx.x * MV) * MAXCOL + threadIdx.x] = make_double2(sh[threadIdx.x], 0.0);
}

// ||W|| per column from slot 0 of the partials; pass 0 also flags the
// columns whose norm dropped below 0.707 of the norm before projection
// (the reorthogonalisation test of the single-system GMRES).
__global__ void rnorm_kernel(int C, const int* __restrict__ col_inst, const int* __restrict__ col_q,
                             const int2* __restrict__ chunk_range, const double2* __restrict__ part, int MV,
                             const int* __restrict__ active, const double* __restrict__ wn0, int pass,
                             double* __restrict__ wn, int* __restrict__ reorth, int64_t* __restrict__ reorth_count) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    if (!active[c]) {
        if (pass == 0) reorth[c] = 0;
        return;
    }
    const int2 r = chunk_range[col_inst[c]];
    const int q = col_q[c];
    double s = 0.0;
    for (int ch = r.x; ch < r.y; ++ch) s += part[((int64_t)ch * MV) * MAXCOL + q].x;
    const double v = sqrt(s);
    wn[c] = v;
    if (pass == 0) {
        const int f = v < 0.707 * wn0[c] ? 1 : 0;
        reorth[c] = f;
        reorth_count[c] += f;
    }
}

// Instance flags: any column active / flagged / not yet converged.
__global__ void rflags_kernel(int B, const BatchInstance* __restrict__ inst, const int* __restrict__ active,
                              const int* __restrict__ reorth, const int* __restrict__ done, int* __restrict__ inst_active,
                              int* __restrict__ inst_reorth, int* __restrict__ inst_pending) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const BatchInstance& I = inst[b];
    int a = 0, r = 0, p = 0;
    for (int q = 0; q < I.ncol; ++q) {
        a |= active[I.col_off + q];
        r |= reorth[I.col_off + q] & active[I.col_off + q];
        p |= !done[I.col_off + q];
    }
    inst_active[b] = a;
    inst_reorth[b] = r;
    inst_pending[b] = p;
}

// Step j of the per-column Arnoldi recurrence after the projection: store
// h(j+1, j) = ||w||, apply the previous rotations to column j, form and apply
// the new rotation (gmres_common.rs:91-105), update g and the residual
// estimate.  A column freezes (jlen = j + 1) when the estimate drops below
// its threshold, when ||w|| < 1e-20 (happy breakdown) or at the end of the
// cycle.  scale = 1/||w|| for the next basis vector (0 once frozen).
__global__ void rgivens_kernel(int C, int j, int m, int* __restrict__ active, const double* __restrict__ wn,
                               double2* __restrict__ H, double2* __restrict__ cs, double2* __restrict__ sn,
                               double2* __restrict__ g, const double* __restrict__ tol_abs, double* __restrict__ est,
                               int* __restrict__ jlen, int64_t* __restrict__ iters, double* __restrict__ scale) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    if (!active[c]) {
        scale[c] = 0.0;
        return;
    }
    double2* h = H + ((int64_t)c * m + j) * (m + 1);
    double2* C_ = cs + (int64_t)c * m;
    double2* S_ = sn + (int64_t)c * m;
    double2* gc = g + (int64_t)c * (m + 1);
    const double w = wn[c];
    h[j + 1] = make_double2(w, 0.0);
    const bool breakdown = w < 1e-20;
    for (int i = 0; i < j; ++i) {
        const double2 temp = cadd(cmul(conj2(C_[i]), h[i]), cmul(conj2(S_[i]), h[i + 1]));
        h[i + 1] = cadd(csub(make_double2(0.0, 0.0), cmul(S_[i], h[i])), cmul(C_[i], h[i + 1]));
        h[i] = temp;
    }
    const double2 a = h[j], bb = h[j + 1];
    double2 c_, s_;
    if (cabs2(bb) < 1e-30) {
        c_ = make_double2(1.0, 0.0);
        s_ = make_double2(0.0, 0.0);
    } else if (cabs2(a) < 1e-30) {
        c_ = make_double2(0.0, 0.0);
        s_ = make_double2(1.0, 0.0);
    } else {
        const double r = sqrt((a.x * a.x + a.y * a.y) + (bb.x * bb.x + bb.y * bb.y));
        c_ = make_double2(a.x * (1.0 / r), a.y * (1.0 / r));
        s_ = make_double2(bb.x * (1.0 / r), bb.y * (1.0 / r));
    }
    C_[j] = c_;
    S_[j] = s_;
    h[j] = cadd(cmul(conj2(c_), h[j]), cmul(conj2(s_), h[j + 1]));
    h[j + 1] = make_double2(0.0, 0.0);
    const double2 temp = cadd(cmul(conj2(c_), gc[j]), cmul(conj2(s_), gc[j + 1]));
    gc[j + 1] = cadd(csub(make_double2(0.0, 0.0), cmul(s_, gc[j])), cmul(c_, gc[j + 1]));
    gc[j] = temp;
    const double e = cabs2(gc[j + 1]);
    est[c] = e;
    iters[c] += 1;
    const bool stop = e < tol_abs[c] || (!breakdown && e < 1e-20) || breakdown || j + 1 == m;
    if (stop) {
        active[c] = 0;
        jlen[c] = j + 1;
        scale[c] = 0.0;
    } else {
        scale[c] = 1.0 / w;
    }
}

// Back substitution y = R^{-1} g over the column's jlen steps (gmres.rs
// solve_upper_triangular order); y goes to coef[c][0..jlen), zeros after.
__global__ void rsolve_kernel(int C, int m, int MV, const double2* __restrict__ H, const double2* __restrict__ g,
                              const int* __restrict__ jlen, double2* __restrict__ coef) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int kk = jlen[c];
    double2* y = coef + (int64_t)c * MV;
    for (int i = 0; i < MV; ++i) y[i] = make_double2(0.0, 0.0);
    const double2* gc = g + (int64_t)c * (m + 1);
    for (int i = kk - 1; i >= 0; --i) {
        double2 sum = gc[i];
        for (int jj = i + 1; jj < kk; ++jj) sum = csub(sum, cmul(H[((int64_t)c * m + jj) * (m + 1) + i], y[jj]));
        const double2 d = H[((int64_t)c * m + i) * (m + 1) + i];
        if (cabs2(d) > 1e-30) y[i] = cmul(sum, cinv(d));
    }
}

// X += sum_{s < jlen} coef[c][s] V_s for the columns of pending instances.
__global__ __launch_bounds__(BLOCK) void rupdate_kernel(const VecChunk* __restrict__ chunks,
                                                        const BatchInstance* __restrict__ inst,
                                                        const int* __restrict__ flag, const double2* __restrict__ V,
                                                        int64_t L, const int* __restrict__ jlen,
                                                        const double2* __restrict__ coef, int MV,
                                                        double2* __restrict__ X) {
    const VecChunk ch = chunks[blockIdx.x];
    if (!flag[ch.inst]) return;
    const BatchInstance& I = inst[ch.inst];
    const int kb = I.ncol;
    const int S = BLOCK / kb;
    const int c = threadIdx.x % kb;
    const int sub = threadIdx.x / kb;
    if (sub >= S) return;
    const int nv = jlen[I.col_off + c];
    const double2* cf = coef + (int64_t)(I.col_off + c) * MV;
    const int64_t base = I.vec_off + ch.r0 * kb + c;
    for (int r = sub; r < ch.rows; r += S) {
        const int64_t idx = base + (int64_t)r * kb;
        double2 x = X[idx];
        for (int s = 0; s < nv; ++s) {
            const double2 p = cmul(cf[s], V[(int64_t)s * L + idx]);
            x.x += p.x;
            x.y += p.y;
        }
        X[idx] = x;
    }
}

// R = B - R per column (R holds A X on entry); ||R||^2 partials into slot 0.
__global__ __launch_bounds__(BLOCK) void rresid_kernel(const VecChunk* __restrict__ chunks,
                                                       const BatchInstance* __restrict__ inst,
                                                       const int* __restrict__ flag, const double2* __restrict__ B,
                                                       double2* __restrict__ R, double2* __restrict__ part, int MV) {
    __shared__ double sh[BLOCK];
    const VecChunk ch = chunks[blockIdx.x];
    if (!flag[ch.inst]) return;
    const BatchInstance& I = inst[ch.inst];
    const int kb = I.ncol;
    const int S = BLOCK / kb;
    const int c = threadIdx.x % kb;
    const int sub = threadIdx.x / kb;
    const int64_t base = I.vec_off + ch.r0 * kb + c;
    double nr = 0.0;
    if (sub < S) {
        for (int r = sub; r < ch.rows; r += S) {
            const int64_t idx = base + (int64_t)r * kb;
            const double2 b = B[idx];
            const double2 ax = R[idx];
            const double2 v = make_double2(b.x - ax.x, b.y - ax.y);
            R[idx] = v;
            nr += v.x * v.x + v.y * v.y;
        }
    }
    sh[threadIdx.x] = nr;
    column_tree(sh, S, kb, sub, c);
    if (threadIdx.x < kb) part[((int64_t)blockIdx.x * MV) * MAXCOL + threadIdx.x] = make_double2(sh[threadIdx.x], 0.0);
}

// After a residual: ||r|| per column, convergence test ||r|| / ||b|| < tol,
// completion time stamp, and the start of the next cycle: g = (||r||, 0, ...),
// scale = 1/||r||, threshold of the Arnoldi estimate = inner_tol * ||r||,
// but not below a quarter of the outer target tol * ||b|| (refinement: a
// cycle that starts near the target stops as soon as its low-precision
// estimate passes it instead of gaining another inner_tol), or tol * ||b||
// (plain).  mode 0: after a cycle (counted);
// 1: the residual of x = 0, which also sets ||b||; 2: the residual of an
// initial guess (not counted).
__global__ void rrestart_kernel(int C, int m, const int* __restrict__ col_inst, const int* __restrict__ col_q,
                                const int2* __restrict__ chunk_range, const double2* __restrict__ part, int MV,
                                const int* __restrict__ pending_in, int mode, double tol, double inner_tol,
                                int refinement, double* __restrict__ bnorm, double* __restrict__ rnorm,
                                int* __restrict__ done, int* __restrict__ active, int* __restrict__ jlen,
                                double2* __restrict__ g, double* __restrict__ scale, double* __restrict__ tol_abs,
                                int64_t* __restrict__ cycles, unsigned long long* __restrict__ t_done) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    jlen[c] = 0;
    if (done[c] || !pending_in[col_inst[c]]) {
        active[c] = 0;
        scale[c] = 0.0;
        return;
    }
    const int2 r = chunk_range[col_inst[c]];
    const int q = col_q[c];
    double s = 0.0;
    for (int ch = r.x; ch < r.y; ++ch) s += part[((int64_t)ch * MV) * MAXCOL + q].x;
    const double rn = sqrt(s);
    if (mode == 1) bnorm[c] = rn;
    rnorm[c] = rn;
    const double bn = bnorm[c];
    const bool conv = bn == 0.0 || rn / bn < tol;
    if (conv || !(rn > 0.0) || !isfinite(rn)) {
        done[c] = 1;
        active[c] = 0;
        scale[c] = 0.0;
        t_done[c] = wall_clock64();
        return;
    }
    if (mode == 0) cycles[c] += 1;
    double2* gc = g + (int64_t)c * (m + 1);
    gc[0] = make_double2(rn, 0.0);
    for (int i = 1; i <= m; ++i) gc[i] = make_double2(0.0, 0.0);
    scale[c] = 1.0 / rn;
    tol_abs[c] = refinement ? fmax(inner_tol * rn, 0.25 * tol * bn) : tol * bn;
    active[c] = 1;
}

// dst = scale[c] * src per column (instances with the flag set).
__global__ __launch_bounds__(BLOCK) void rscale_kernel(const VecChunk* __restrict__ chunks,
                                                       const BatchInstance* __restrict__ inst,
                                                       const int* __restrict__ flag, const double* __restrict__ scale,
                                                       const double2* __restrict__ src, double2* __restrict__ dst) {
    const VecChunk ch = chunks[blockIdx.x];
    if (!flag[ch.inst]) return;
    const BatchInstance& I = inst[ch.inst];
    const int kb = I.ncol;
    const int S = BLOCK / kb;
    const int c = threadIdx.x % kb;
    const int sub = threadIdx.x / kb;
    if (sub >= S) return;
    const double sc = scale[I.col_off + c];
    const int64_t base = I.vec_off + ch.r0 * kb + c;
    for (int r = sub; r < ch.rows; r += S) {
        const int64_t idx = base + (int64_t)r * kb;
        const double2 v = src[idx];
        dst[idx] = make_double2(v.x * sc, v.y * sc);
    }
}

__global__ void rfill_kernel(double2* __restrict__ p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = make_double2(0.0, 0.0);
}

__global__ void rstamp_kernel(unsigned long long* __restrict__ out) { *out = wall_clock64(); }

// ---- initial guesses from device-resident vectors ----
// One grid row (blockIdx.y) per instance, one grid layer (blockIdx.z) per
// candidate slot; element e < n * ncol of an instance's block is at
// vec_off + e of a packed vector.

// dst[z * L + vec_off + e] = src[z * B + inst][e]; a null source gives zeros.
__global__ void bgather_kernel(const BatchInstance* __restrict__ inst, int B, int64_t L,
                               const double2* const* __restrict__ src, double2* __restrict__ dst) {
    const BatchInstance& I = inst[blockIdx.y];
    const int64_t cnt = I.n * I.ncol;
    const double2* s = src[(int64_t)blockIdx.z * B + blockIdx.y];
    double2* d = dst + (int64_t)blockIdx.z * L + I.vec_off;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * BLOCK)
        d[e] = s ? s[e] : make_double2(0.0, 0.0);
}

// dst[inst][e] = x[vec_off + e] for every instance with a destination.
__global__ void bscatter_kernel(const BatchInstance* __restrict__ inst, const double2* __restrict__ x,
                                double2* const* __restrict__ dst) {
    const BatchInstance& I = inst[blockIdx.y];
    double2* d = dst[blockIdx.y];
    if (!d) return;
    const int64_t cnt = I.n * I.ncol;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * BLOCK)
        d[e] = x[I.vec_off + e];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Per column c (one workgroup): the Gram matrix G[a][b] = sum_r conj(P_a) P_b
// of the M candidate products, h[a] = sum_r conj(P_a) rhs and |rhs|^2, over
// the column's rows.  Fixed order: strided per-thread sums, a butterfly per
// wave, then the 4 wave sums in order.
template <int M>
__global__ __launch_bounds__(BLOCK) void bgram_kernel(const BatchInstance* __restrict__ inst,
                                                      const int* __restrict__ col_inst, const int* __restrict__ col_q,
                                                      const double2* __restrict__ P, int64_t L,
                                                      const double2* __restrict__ rhs, double2* __restrict__ G,
                                                      double2* __restrict__ h, double* __restrict__ bb) {
    constexpr int NP = M * (M + 1) / 2;
    constexpr int NV = 2 * NP + 2 * M + 1;
    __shared__ double sh[BLOCK / 64][NV];
    const int c = blockIdx.x;
    const BatchInstance& I = inst[col_inst[c]];
    const int q = col_q[c];
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = 0.0;
    for (int64_t r = threadIdx.x; r < I.n; r += BLOCK) {
        const int64_t t = I.vec_off + r * I.ncol + q;
        double2 p[M];
#pragma unroll
        for (int a = 0; a < M; ++a) p[a] = P[(int64_t)a * L + t];
        const double2 bv = rhs[t];
        int o = 0;
#pragma unroll
        for (int a = 0; a < M; ++a)
#pragma unroll
            for (int b = a; b < M; ++b) {
                const double2 g = cmulc(p[a], p[b]);
                v[o++] += g.x;
                v[o++] += g.y;
            }
#pragma unroll
        for (int a = 0; a < M; ++a) {
            const double2 g = cmulc(p[a], bv);
            v[o++] += g.x;
            v[o++] += g.y;
        }
        v[o] += bv.x * bv.x + bv.y * bv.y;
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = wave_sum(v[i]);
        if ((threadIdx.x & 63) == 0) sh[w][i] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NV; i += BLOCK) {
        double s = sh[0][i];
        for (int j = 1; j < BLOCK / 64; ++j) s += sh[j][i];
        sh[0][i] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int o = 0;
        for (int a = 0; a < M; ++a)
            for (int b = a; b < M; ++b) {
                const double re = sh[0][o++], im = sh[0][o++];
                G[((int64_t)c * M + a) * M + b] = make_double2(re, im);
                G[((int64_t)c * M + b) * M + a] = make_double2(re, -im);
            }
        for (int a = 0; a < M; ++a) {
            const double re = sh[0][o++], im = sh[0][o++];
            h[(int64_t)c * M + a] = make_double2(re, im);
        }
        bb[c] = sh[0][o];
    }
}

// x[vec_off + e] = sum_a coef[col * m + a] S[a * L + vec_off + e], col = col_off + e % ncol.
__global__ void bcombine_kernel(const BatchInstance* __restrict__ inst, const double2* __restrict__ S, int64_t L,
                                int m, const double2* __restrict__ coef, double2* __restrict__ x) {
    const BatchInstance& I = inst[blockIdx.y];
    const int64_t cnt = I.n * I.ncol;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * BLOCK) {
        const double2* cf = coef + (int64_t)(I.col_off + (int)(e % I.ncol)) * m;
        double2 acc = make_double2(0.0, 0.0);
        for (int a = 0; a < m; ++a) acc = cadd(acc, cmul(cf[a], S[(int64_t)a * L + I.vec_off + e]));
        x[I.vec_off + e] = acc;
    }
}

// Raw network inputs of every (panel, column) sample, at the packed index:
// {centroid, outward normal, incident direction, k} as FP32.
__global__ void blearned_raw_kernel(const BatchInstance* __restrict__ inst, const Source* __restrict__ src,
                                    const double* __restrict__ dirs, float* __restrict__ raw) {
    const BatchInstance& I = inst[blockIdx.y];
    const int64_t cnt = I.n * I.ncol;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * BLOCK) {
        const Source s = src[I.panel_off + e / I.ncol];
        const double* d = dirs + 3 * (I.dir_off + e % I.ncol);
        float* o = raw + (I.vec_off + e) * 10;
        for (int j = 0; j < 3; ++j) {
            o[j] = (float)s.c[j];
            o[3 + j] = (float)s.n[j];
            o[6 + j] = (float)d[j];
        }
        o[9] = (float)I.k;
    }
}

// Guess x0 = u exp(i k d.c) from the network's incident-phase-free density u.
__global__ void blearned_phase_kernel(const BatchInstance* __restrict__ inst, const Source* __restrict__ src,
                                      const double* __restrict__ dirs, const float* __restrict__ Y,
                                      double2* __restrict__ x) {
    const BatchInstance& I = inst[blockIdx.y];
    const int64_t cnt = I.n * I.ncol;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * BLOCK) {
        const Source s = src[I.panel_off + e / I.ncol];
        const double* d = dirs + 3 * (I.dir_off + e % I.ncol);
        const int64_t t = I.vec_off + e;
        double sn, cs;
        sincos(I.k * (s.c[0] * d[0] + s.c[1] * d[1] + s.c[2] * d[2]), &sn, &cs);
        x[t] = cmul(make_double2((double)Y[2 * t], (double)Y[2 * t + 1]), make_double2(cs, sn));
    }
}

inline unsigned nb(int64_t n, int per = BLOCK) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ---- launchers (declared in batch.hpp) ----

void batch_lp_convert(sbh_context* ctx, int format, int64_t rows, const int* row_inst, const BatchInstance* inst,
                      const double2* A, uint16_t* Alp, double* row_inv) {
    if (rows == 0) return;
    if (format == 0)
        hipLaunchKernelGGL(glp_convert_kernel<0>, dim3((unsigned)rows), dim3(CONV_BLOCK), 0, ctx->stream, row_inst, inst, A,
                           Alp, row_inv);
    else
        hipLaunchKernelGGL(glp_convert_kernel<1>, dim3((unsigned)rows), dim3(CONV_BLOCK), 0, ctx->stream, row_inst, inst, A,
                           Alp, row_inv);
    SBH_HIP(hipGetLastError());
}

void batch_product(sbh_context* ctx, int format, int max_cols, int ntiles, const int2* tiles,
                   const BatchInstance* inst, const int* flag, const double2* A, const uint16_t* Alp,
                   const double* row_inv, const double2* X, double2* Y) {
    if (ntiles == 0) return;
    if (format < 0 && max_cols <= 8)
        hipLaunchKernelGGL(g64_gemm_kernel<1>, dim3((unsigned)ntiles), dim3(BLOCK), 0, ctx->stream, tiles, inst, flag,
                           A, X, Y);
    else if (format < 0)
        hipLaunchKernelGGL(g64_gemm_kernel<2>, dim3((unsigned)ntiles), dim3(BLOCK), 0, ctx->stream, tiles, inst, flag,
                           A, X, Y);
    else if (format == 0)
        hipLaunchKernelGGL(glp_gemm_kernel<0>, dim3((unsigned)ntiles), dim3(BLOCK), 0, ctx->stream, tiles, inst, flag,
                           Alp, row_inv, X, Y);
    else
        hipLaunchKernelGGL(glp_gemm_kernel<1>, dim3((unsigned)ntiles), dim3(BLOCK), 0, ctx->stream, tiles, inst, flag,
                           Alp, row_inv, X, Y);
    SBH_HIP(hipGetLastError());
}

void ragged_dots(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* V, int nvec,
                 const double2* W, double2* part) {
    hipLaunchKernelGGL(rdots_kernel, dim3((unsigned)lay.nchunks, (unsigned)(nvec + 1)), dim3(BLOCK), 0, ctx->stream,
                       lay.chunks, lay.inst, flag, V, lay.L, nvec, W, part, lay.MV);
    SBH_HIP(hipGetLastError());
}

void ragged_sum_dots(sbh_context* ctx, const RaggedLayout& lay, const double2* part, int nvec, int j,
                     const int* active, const int* reorth, int pass, double2* H, double2* coef, double* wn0) {
    hipLaunchKernelGGL(rsum_dots_kernel, dim3(nb((int64_t)lay.C * (nvec + 1))), dim3(BLOCK), 0, ctx->stream, lay.C,
                       lay.col_inst, lay.col_q, lay.chunk_range, part, lay.MV, nvec, j, lay.m, active, reorth, pass, H,
                       coef, wn0);
    SBH_HIP(hipGetLastError());
}

void ragged_project(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* V, int nvec,
                    const double2* coef, double2* W, double2* part) {
    hipLaunchKernelGGL(rproject_kernel, dim3((unsigned)lay.nchunks), dim3(BLOCK), 0, ctx->stream, lay.chunks, lay.inst,
                       flag, V, lay.L, nvec, coef, lay.MV, W, part);
    SBH_HIP(hipGetLastError());
}

void ragged_norm(sbh_context* ctx, const RaggedLayout& lay, const double2* part, const int* active, const double* wn0,
                 int pass, double* wn, int* reorth, int64_t* reorth_count) {
    hipLaunchKernelGGL(rnorm_kernel, dim3(nb(lay.C)), dim3(BLOCK), 0, ctx->stream, lay.C, lay.col_inst, lay.col_q,
                       lay.chunk_range, part, lay.MV, active, wn0, pass, wn, reorth, reorth_count);
    SBH_HIP(hipGetLastError());
}

void ragged_flags(sbh_context* ctx, const RaggedLayout& lay, const int* active, const int* reorth, const int* done,
                  int* inst_active, int* inst_reorth, int* inst_pending) {
    hipLaunchKernelGGL(rflags_kernel, dim3(nb(lay.B)), dim3(BLOCK), 0, ctx->stream, lay.B, lay.inst, active, reorth,
                       done, inst_active, inst_reorth, inst_pending);
    SBH_HIP(hipGetLastError());
}

void ragged_givens(sbh_context* ctx, const RaggedLayout& lay, int j, int* active, const double* wn, double2* H,
                   double2* cs, double2* sn, double2* g, const double* tol_abs, double* est, int* jlen, int64_t* iters,
                   double* scale) {
    hipLaunchKernelGGL(rgivens_kernel, dim3(nb(lay.C)), dim3(BLOCK), 0, ctx->stream, lay.C, j, lay.m, active, wn, H,
                       cs, sn, g, tol_abs, est, jlen, iters, scale);
    SBH_HIP(hipGetLastError());
}

void ragged_solve(sbh_context* ctx, const RaggedLayout& lay, const double2* H, const double2* g, const int* jlen,
                  double2* coef) {
    hipLaunchKernelGGL(rsolve_kernel, dim3(nb(lay.C, 64)), dim3(64), 0, ctx->stream, lay.C, lay.m, lay.MV, H, g, jlen,
                       coef);
    SBH_HIP(hipGetLastError());
}

void ragged_update(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* V, const int* jlen,
                   const double2* coef, double2* X) {
    hipLaunchKernelGGL(rupdate_kernel, dim3((unsigned)lay.nchunks), dim3(BLOCK), 0, ctx->stream, lay.chunks, lay.inst,
                       flag, V, lay.L, jlen, coef, lay.MV, X);
    SBH_HIP(hipGetLastError());
}

void ragged_residual(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* B, double2* R,
                     double2* part) {
    hipLaunchKernelGGL(rresid_kernel, dim3((unsigned)lay.nchunks), dim3(BLOCK), 0, ctx->stream, lay.chunks, lay.inst,
                       flag, B, R, part, lay.MV);
    SBH_HIP(hipGetLastError());
}

void ragged_restart(sbh_context* ctx, const RaggedLayout& lay, const double2* part, const int* pending_in, int mode,
                    double tol, double inner_tol, int refinement, double* bnorm, double* rnorm, int* done, int* active,
                    int* jlen, double2* g, double* scale, double* tol_abs, int64_t* cycles,
                    unsigned long long* t_done) {
    hipLaunchKernelGGL(rrestart_kernel, dim3(nb(lay.C)), dim3(BLOCK), 0, ctx->stream, lay.C, lay.m, lay.col_inst,
                       lay.col_q, lay.chunk_range, part, lay.MV, pending_in, mode, tol, inner_tol, refinement, bnorm,
                       rnorm, done, active, jlen, g, scale, tol_abs, cycles, t_done);
    SBH_HIP(hipGetLastError());
}

void ragged_scale(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double* scale, const double2* src,
                  double2* dst) {
    hipLaunchKernelGGL(rscale_kernel, dim3((unsigned)lay.nchunks), dim3(BLOCK), 0, ctx->stream, lay.chunks, lay.inst,
                       flag, scale, src, dst);
    SBH_HIP(hipGetLastError());
}

void ragged_fill(sbh_context* ctx, double2* p, int64_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(rfill_kernel, dim3(nb(n)), dim3(BLOCK), 0, ctx->stream, p, n);
    SBH_HIP(hipGetLastError());
}

void ragged_stamp(sbh_context* ctx, unsigned long long* out) {
    hipLaunchKernelGGL(rstamp_kernel, dim3(1), dim3(1), 0, ctx->stream, out);
    SBH_HIP(hipGetLastError());
}

namespace {
// Grid over the largest instance block, at most 4096 workgroups per row of
// instances (the kernels stride over the rest).
dim3 instance_grid(int64_t max_block, int B, int layers) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(nb(max_block), 4096)), (unsigned)B, (unsigned)layers);
}
}  // namespace

void batch_gather(sbh_context* ctx, int B, int64_t L, int64_t max_block, int layers, const BatchInstance* inst,
                  const double2* const* src, double2* dst) {
    hipLaunchKernelGGL(bgather_kernel, instance_grid(max_block, B, layers), dim3(BLOCK), 0, ctx->stream, inst, B, L,
                       src, dst);
    SBH_HIP(hipGetLastError());
}

void batch_scatter(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const double2* x,
                   double2* const* dst) {
    hipLaunchKernelGGL(bscatter_kernel, instance_grid(max_block, B, 1), dim3(BLOCK), 0, ctx->stream, inst, x, dst);
    SBH_HIP(hipGetLastError());
}

void batch_gram(sbh_context* ctx, int m, int C, const BatchInstance* inst, const int* col_inst, const int* col_q,
                const double2* P, int64_t L, const double2* rhs, double2* G, double2* h, double* bb) {
    switch (m) {
#define SBH_GRAM_CASE(M)                                                                                           \
    case M:                                                                                                        \
        hipLaunchKernelGGL(bgram_kernel<M>, dim3(C), dim3(BLOCK), 0, ctx->stream, inst, col_inst, col_q, P, L, rhs, \
                           G, h, bb);                                                                              \
        break;
        SBH_GRAM_CASE(1)
        SBH_GRAM_CASE(2)
        SBH_GRAM_CASE(3)
        SBH_GRAM_CASE(4)
        SBH_GRAM_CASE(5)
        SBH_GRAM_CASE(6)
#undef SBH_GRAM_CASE
        default:
            throw Error(SBH_ERR_INVALID_ARGUMENT, "the number of candidates must be in 1..6");
    }
    SBH_HIP(hipGetLastError());
}

void batch_combine(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const double2* S, int64_t L,
                   int m, const double2* coef, double2* x) {
    hipLaunchKernelGGL(bcombine_kernel, instance_grid(max_block, B, 1), dim3(BLOCK), 0, ctx->stream, inst, S, L, m,
                       coef, x);
    SBH_HIP(hipGetLastError());
}

void batch_learned_raw(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const Source* src,
                       const double* dirs, float* raw) {
    hipLaunchKernelGGL(blearned_raw_kernel, instance_grid(max_block, B, 1), dim3(BLOCK), 0, ctx->stream, inst, src,
                       dirs, raw);
    SBH_HIP(hipGetLastError());
}

void batch_learned_phase(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const Source* src,
                         const double* dirs, const float* Y, double2* x) {
    hipLaunchKernelGGL(blearned_phase_kernel, instance_grid(max_block, B, 1), dim3(BLOCK), 0, ctx->stream, inst, src,
                       dirs, Y, x);
    SBH_HIP(hipGetLastError());
}

}  // namespace sbh
