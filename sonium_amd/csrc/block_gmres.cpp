// Many right-hand sides: column-batched restarted GMRES and mixed-precision
// FGMRES iterative refinement.
//
// Batched GMRES.  Each of the k columns runs exactly the single right-hand-side
// algorithm of gmres_core (crates/sonium-solvers/crates/sonium-solvers-iterative/
// src/gmres.rs:631-1057: restart cycles, CGS with one DGKS pass at 0.707,
// complex Givens, convergence on the projected residual confirmed by the
// true residual) with its own Hessenberg matrix, but the k columns advance
// in lock step so every Arnoldi step needs one block product A*[v_j^(1) ..
// v_j^(k)] — a GEMM instead of k GEMVs.  A column that converges, breaks
// down or must restart pauses until the cycle ends; pausing changes only
// when its next cycle starts, not what it computes.  This serves the
// reference's multi-RHS entry (BemSolver::solve_many → apply_block,
// bem_solver.rs:2032, 2175-2181, 6567-6574) with per-column solutions of the
// single-RHS method instead of block-Krylov ones.
//
// FGMRES-IR.  The reference's mixed-precision precedent
// (crates/sonium-compute/src/gmres_ir.rs:1-80, 536-660; Carson & Higham 2018,
// Alg. 5.3): a flexible GMRES outer loop in FP64 (Arnoldi, Givens, residual,
// update) whose right preconditioner z = M(v) is an inner GMRES solve with the
// low-precision operator (complex FP16/BF16 storage, FP32 accumulation).  The
// outer update uses the preconditioned vectors Z.  Defaults follow
// GmresIrConfig::default (gmres_ir.rs:560-580).

#include <algorithm>
#include <cmath>
#include <complex>
#include <memory>
#include <vector>

#include "internal.hpp"

namespace sbh {

int64_t lowp_padded_ld(int64_t n);
void zgemm_f64(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X, int64_t k, double2* Y);
int64_t batched_scratch_elems(int64_t n, int k, int max_vec);
// Block vectors of complex FP64 (double2) or FP32 (float2) entries; the
// reductions return FP64 values for both.
#define SBH_DECLARE_VECTOR_OPS(T)                                                                                 \
    void lowp_gemm(sbh_context* ctx, const uint16_t* A, int64_t ld, int64_t n, int format, const double* row_inv, \
                   const T* X, int k, uint16_t* XT, unsigned* colmax, double* col_inv, float* Ypart, T* Y);       \
    void bdots(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const T* W, int64_t n, int k,                 \
               double2* scratch, double2* out);                                                                   \
    void bproject(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const double2* C, T* W, int64_t n, int k,   \
                  double* scratch, double* out);                                                                  \
    void bresidual(sbh_context* ctx, const T* B, T* R, int64_t n, int k, double* scratch, double* out);           \
    void bupdate(sbh_context* ctx, const T* V, int64_t ldv, int nvec, const double2* Yc, T* X, int64_t n, int k); \
    void bscale(sbh_context* ctx, const T* X, const double* alpha, T* Y, int64_t n, int k);                       \
    void zero_fill(sbh_context* ctx, T* X, int64_t count);
SBH_DECLARE_VECTOR_OPS(double2)
SBH_DECLARE_VECTOR_OPS(float2)
#undef SBH_DECLARE_VECTOR_OPS
void convert_block(sbh_context* ctx, const double2* X, float2* Y, int64_t count);
void convert_block(sbh_context* ctx, const float2* X, double2* Y, int64_t count);

namespace {

using cplx = std::complex<double>;

inline cplx cm(cplx a, cplx b) {
    return cplx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cplx cinv(cplx a) {
    double n = a.real() * a.real() + a.imag() * a.imag();
    return cplx(a.real() / n, -a.imag() / n);
}
inline double cnorm(cplx a) { return std::hypot(a.real(), a.imag()); }

}  // namespace

// Block operator: Y (n x k) = A X (n x k), device memory, context stream;
// T is the vector element type (double2 or float2).
template <typename T>
struct BlockOperatorT {
    sbh_context* ctx = nullptr;
    int64_t n = 0;
    int64_t applies = 0;
    virtual void apply(const T* X, T* Y, int k) = 0;
    virtual ~BlockOperatorT() = default;
};
using BlockOperator = BlockOperatorT<double2>;

struct DenseBlockOperator : BlockOperator {
    const sbh_system* sys;
    explicit DenseBlockOperator(const sbh_system* s) : sys(s) {
        ctx = s->ctx;
        n = s->n;
    }
    void apply(const double2* X, double2* Y, int k) override {
        if (k == 1)
            zgemv_n(ctx, sys->d_A.ptr, sys->lda, n, n, X, Y);
        else
            zgemm_f64(ctx, sys->d_A.ptr, sys->lda, n, X, k, Y);
        applies += 1;
    }
};

template <typename T>
struct LowpBlockOperator : BlockOperatorT<T> {
    sbh_lowp_matrix* lp;
    explicit LowpBlockOperator(sbh_lowp_matrix* m) : lp(m) {
        this->ctx = m->ctx;
        this->n = m->n;
    }
    void apply(const T* X, T* Y, int k) override {
        lowp_gemm(this->ctx, lp->d_A.ptr, lp->ld, this->n, lp->format, lp->d_row_inv.ptr, X, k, lp->d_XT.ptr,
                  lp->d_colmax.ptr, lp->d_colinv.ptr, lp->d_ypart.ptr, Y);
        this->applies += 1;
    }
};

// Right preconditioner of the flexible variant: Z (n x k) = M(V).
template <typename T>
struct BlockPreconditionerT {
    virtual void apply(const T* V, T* Z, int k) = 0;
    virtual ~BlockPreconditionerT() = default;
};
using BlockPreconditioner = BlockPreconditionerT<double2>;

struct BatchedParams {
    int64_t max_iterations = 100;  // restart cycles
    int64_t restart = 30;
    double tol = 1e-8;
    bool exact_checks = true;  // confirm convergence with the true residual (gmres_core)
    bool finalize = true;      // finalize_gmres_solution
    // X is zero on entry: the first residual is B itself and A*0 is not formed.
    // The result is the same for any finite A (a non-finite A then shows at the
    // first Krylov product instead of the first residual).
    bool x_zero = false;
    // Unconditional two-pass classical Gram-Schmidt whose coefficients go from
    // the dot products straight into the projection on the device, as the
    // reference's inner solve does (gmres_inner, gmres_ir.rs:360-394: h = d1 + d2,
    // then the norm after the second pass); one host read per Arnoldi step.
    // Otherwise: one pass plus the DGKS pass when ||w|| < 0.707 ||w0|| (gmres_core).
    bool cgs2 = false;
};

struct ColumnResult {
    int64_t iterations = 0, restarts = 0, reorth = 0;
    double residual = 0, estimated = 0, backward_error = 0;
    bool converged = false;
    int status = SBH_SOLVER_MAX_ITERATIONS;
};

// Reusable workspace for one (n, k, restart, flexible) shape; T is the
// element type of the Krylov vectors (the Hessenberg matrix, the rotations
// and every reduction stay FP64).
template <typename T>
class BatchedGmres {
   public:
    BatchedGmres(sbh_context* ctx, int64_t n, int k, int m, bool flexible)
        : ctx_(ctx), n_(n), k_(k), m_(m), flexible_(flexible) {
        const size_t blk = (size_t)n * k;
        admit_device(ctx, (int64_t)(sizeof(T) * blk * ((size_t)(m + 1) + (flexible ? (size_t)m : 0) + 3)),
                     "the block GMRES Krylov basis", false);
        V_.allocate((size_t)(m + 1) * blk);
        if (flexible) Z_.allocate((size_t)m * blk);
        W_.allocate(blk);
        R_.allocate(blk);
        scratch_.allocate((size_t)batched_scratch_elems(n, k, m + 1) + 16);
        part_.allocate((size_t)(m + 2) * k + 16);
        C_.allocate((size_t)(m + 1) * k + 1);
        alpha_.allocate((size_t)k);
        // Separate pinned regions: each is rewritten by the host only after a
        // stream synchronisation has retired the copy that last used it.
        down_.allocate(((size_t)(m + 2) * k * 2 + 16) * 2 + k);
        gs_.allocate((size_t)2 * (m + 2) * k + (k + 1) / 2 + 1);
        up_coef_.allocate((size_t)(m + 1) * k * 2 + 16);
        up_alpha_.allocate((size_t)k + 16);
        pinned_ = down_.ptr;
    }

    void run(BlockOperatorT<T>& A, BlockPreconditionerT<T>* M, const T* B, T* X, const BatchedParams& p,
             std::vector<ColumnResult>& out);

   private:
    sbh_context* ctx_;
    int64_t n_;
    int k_, m_;
    bool flexible_;
    DeviceArray<T> V_, Z_, W_, R_;
    DeviceArray<double2> part_, C_, scratch_, gs_;
    DeviceArray<double> alpha_;
    PinnedArray<double> down_, up_coef_, up_alpha_;
    double* pinned_ = nullptr;  // download region

    T* vec(int i) { return V_.ptr + (size_t)i * n_ * k_; }
    T* zvec(int i) { return Z_.ptr + (size_t)i * n_ * k_; }

    void fetch(const void* dev, size_t doubles) {
        SBH_HIP(hipMemcpyAsync(pinned_, dev, sizeof(double) * doubles, hipMemcpyDeviceToHost, ctx_->stream));
        SBH_HIP(hipStreamSynchronize(ctx_->stream));
    }
    // k squared column norms reduced on the device into part_
    void column_sums(std::vector<double>& out) {
        fetch(part_.ptr, (size_t)k_);
        out.assign(pinned_, pinned_ + k_);
    }
    std::vector<double> norms(const T* Xd) {
        bresidual(ctx_, static_cast<const T*>(nullptr), const_cast<T*>(Xd), n_, k_,
                  reinterpret_cast<double*>(scratch_.ptr),
                  reinterpret_cast<double*>(part_.ptr));
        std::vector<double> s;
        column_sums(s);
        for (auto& v : s) v = std::sqrt(v);
        return s;
    }
    std::vector<double> residual_norms(const T* B, T* R) {  // R = B - R
        bresidual(ctx_, B, R, n_, k_, reinterpret_cast<double*>(scratch_.ptr), reinterpret_cast<double*>(part_.ptr));
        std::vector<double> s;
        column_sums(s);
        for (auto& v : s) v = std::sqrt(v);
        return s;
    }
    // coefficients <v_i, w> for i < nvec (per column) and ||w||
    void dots(int nvec, const T* Wd, std::vector<std::vector<cplx>>& coef, std::vector<double>& wn) {
        bdots(ctx_, V_.ptr, n_ * k_, nvec, Wd, n_, k_, scratch_.ptr, part_.ptr);
        fetch(part_.ptr, (size_t)(nvec + 1) * k_ * 2);
        coef.assign(k_, std::vector<cplx>(nvec));
        wn.assign(k_, 0.0);
        for (int c = 0; c < k_; ++c) {
            for (int v = 0; v < nvec; ++v) {
                size_t idx = ((size_t)v * k_ + c) * 2;
                coef[c][v] = cplx(pinned_[idx], pinned_[idx + 1]);
            }
            wn[c] = std::sqrt(pinned_[((size_t)nvec * k_ + c) * 2]);
        }
    }
    void upload_coeffs(const std::vector<std::vector<cplx>>& coef, int nvec, const std::vector<char>& use) {
        double* up = up_coef_.ptr;
        for (int i = 0; i < nvec; ++i)
            for (int c = 0; c < k_; ++c) {
                cplx v = (use[c] && i < (int)coef[c].size()) ? coef[c][i] : cplx(0.0, 0.0);
                up[2 * ((size_t)i * k_ + c)] = v.real();
                up[2 * ((size_t)i * k_ + c) + 1] = v.imag();
            }
        SBH_HIP(hipMemcpyAsync(C_.ptr, up, sizeof(double2) * nvec * k_, hipMemcpyHostToDevice, ctx_->stream));
    }
    void upload_alpha(const std::vector<double>& a) {
        std::copy(a.begin(), a.end(), up_alpha_.ptr);
        SBH_HIP(hipMemcpyAsync(alpha_.ptr, up_alpha_.ptr, sizeof(double) * k_, hipMemcpyHostToDevice, ctx_->stream));
    }
};

template <typename T>
void BatchedGmres<T>::run(BlockOperatorT<T>& A, BlockPreconditionerT<T>* M, const T* B, T* X,
                          const BatchedParams& p, std::vector<ColumnResult>& out) {
    const int k = k_;
    const int m = m_;
    out.assign(k, ColumnResult{});
    std::vector<char> done(k, 0);
    auto finish = [&](int c, double res, double est, double bw, bool conv, int status) {
        out[c].residual = res;
        out[c].estimated = est;
        out[c].backward_error = bw;
        out[c].converged = conv;
        out[c].status = status;
        done[c] = 1;
    };
    // invalid_gmres_input (gmres.rs:252-275)
    std::vector<double> bn = norms(B);
    std::vector<double> xn = norms(X);
    const bool bad_cfg = p.max_iterations <= 0 || p.restart <= 0 || !std::isfinite(p.tol) || p.tol <= 0.0;
    std::vector<double> true_b(k), pb(k);
    for (int c = 0; c < k; ++c) {
        if (bad_cfg || !std::isfinite(bn[c]) || !std::isfinite(xn[c])) {
            finish(c, INFINITY, INFINITY, INFINITY, false, SBH_SOLVER_NUMERICAL_ERROR);
            continue;
        }
        true_b[c] = bn[c] < 1e-15 ? 1.0 : bn[c];
        pb[c] = true_b[c];
    }
    std::vector<int64_t> cycles(k, 0);
    std::vector<std::vector<cplx>> H(k, std::vector<cplx>((size_t)(m + 1) * m)), g(k, std::vector<cplx>(m + 1));
    std::vector<std::vector<cplx>> cs(k), sn(k);
    auto h = [&](int c, int i, int j) -> cplx& { return H[c][(size_t)i * m + j]; };
    std::vector<char> cycling(k, 0), inner_conv(k, 0);
    std::vector<double> alpha(k, 0.0);
    std::vector<std::vector<cplx>> coef, re;
    std::vector<double> wnb, wn, wn2;
    // R_ = B - A X is kept with the version of X it belongs to, so a cycle
    // start or the final check right after a true-residual check reuses it
    // instead of forming A X again (every FP64 product is a full pass over A).
    int64_t x_version = 0, r_version = -1;
    std::vector<double> r_abs(k, 0.0), ax_abs(k, 0.0);
    auto refresh_residual = [&]() {
        if (r_version == x_version) return;
        if (x_version == 0 && p.x_zero) {
            zero_fill(ctx_, R_.ptr, n_ * k_);
            std::fill(ax_abs.begin(), ax_abs.end(), 0.0);
        } else {
            A.apply(X, R_.ptr, k);
            if (p.finalize) ax_abs = norms(R_.ptr);
        }
        r_abs = residual_norms(B, R_.ptr);
        r_version = x_version;
    };

    auto solve_y = [&](int c, int kk) {
        std::vector<cplx> y(kk, cplx(0.0, 0.0));
        for (int i = kk - 1; i >= 0; --i) {
            cplx sum = g[c][i];
            for (int j = i + 1; j < kk; ++j) sum -= cm(h(c, i, j), y[j]);
            if (cnorm(h(c, i, i)) > 1e-30) y[i] = cm(sum, cinv(h(c, i, i)));
        }
        return y;
    };
    // x_c += sum_i y_c[i] basis_i[:, c] for the listed columns
    auto update_x = [&](const std::vector<char>& which, int kk) {
        std::vector<std::vector<cplx>> ys(k);
        for (int c = 0; c < k; ++c)
            if (which[c]) ys[c] = solve_y(c, kk);
        upload_coeffs(ys, kk, which);
        bupdate(ctx_, flexible_ ? Z_.ptr : V_.ptr, n_ * k_, kk, C_.ptr, X, n_, k_);
        ++x_version;
    };
    auto true_residuals = [&]() {
        refresh_residual();
        std::vector<double> r = r_abs;
        for (int c = 0; c < k; ++c) r[c] /= std::max(true_b[c], 2.2250738585072014e-308);
        return r;
    };

    for (;;) {
        bool any = false;
        for (int c = 0; c < k; ++c)
            if (!done[c] && cycles[c] < p.max_iterations) any = true;
        if (!any) break;
        // start a cycle: r = b - A x
        refresh_residual();
        const std::vector<double> rn = r_abs;
        bool anycyc = false;
        for (int c = 0; c < k; ++c) {
            cycling[c] = 0;
            alpha[c] = 0.0;
            if (done[c] || cycles[c] >= p.max_iterations) continue;
            cycles[c] += 1;
            double true_rel = rn[c] / true_b[c];
            double beta = rn[c];
            double rel = beta / pb[c];
            if (rel < p.tol && true_rel < p.tol) {
                finish(c, true_rel, rel, true_rel, true, SBH_SOLVER_CONVERGED);
                continue;
            }
            if (!std::isfinite(beta) || beta == 0.0) {
                finish(c, true_rel, rel, true_rel, false, SBH_SOLVER_BREAKDOWN);
                continue;
            }
            cycling[c] = 1;
            anycyc = true;
            alpha[c] = 1.0 / beta;
            std::fill(H[c].begin(), H[c].end(), cplx(0.0, 0.0));
            std::fill(g[c].begin(), g[c].end(), cplx(0.0, 0.0));
            g[c][0] = cplx(beta, 0.0);
            cs[c].clear();
            sn[c].clear();
            inner_conv[c] = 0;
        }
        if (!anycyc) continue;
        upload_alpha(alpha);
        bscale(ctx_, R_.ptr, alpha_.ptr, vec(0), n_, k);
        std::vector<char> end_of_cycle(k, 0);
        for (int j = 0; j < m; ++j) {
            bool live = false;
            for (int c = 0; c < k; ++c) live = live || cycling[c];
            if (!live) break;
            for (int c = 0; c < k; ++c)
                if (cycling[c]) out[c].iterations += 1;
            const T* zj = vec(j);
            if (flexible_) {
                M->apply(vec(j), zvec(j), k);
                zj = zvec(j);
            }
            A.apply(zj, W_.ptr, k);
            if (p.cgs2) {
                // pass 1 and pass 2 of CGS2 with device-side coefficients; a column that has
                // stopped has zero vectors from here on, so projecting it changes nothing
                const size_t region = (size_t)(m + 2) * k;
                double2* d1 = gs_.ptr;
                double2* d2 = gs_.ptr + region;
                double* nrm = reinterpret_cast<double*>(gs_.ptr + 2 * region);
                double* scratch = reinterpret_cast<double*>(scratch_.ptr);
                bdots(ctx_, V_.ptr, n_ * k_, j + 1, W_.ptr, n_, k, scratch_.ptr, d1);
                bproject(ctx_, V_.ptr, n_ * k_, j + 1, d1, W_.ptr, n_, k, scratch, nullptr);
                bdots(ctx_, V_.ptr, n_ * k_, j + 1, W_.ptr, n_, k, scratch_.ptr, d2);
                bproject(ctx_, V_.ptr, n_ * k_, j + 1, d2, W_.ptr, n_, k, scratch, nrm);
                fetch(gs_.ptr, 2 * region * 2 + k);
                const double* h1 = pinned_;
                const double* h2 = pinned_ + 2 * region;
                const double* hn = pinned_ + 4 * region;
                wn.assign(k, 0.0);
                for (int c = 0; c < k; ++c) {
                    if (!cycling[c]) continue;
                    for (int i = 0; i <= j; ++i) {
                        const size_t idx = ((size_t)i * k + c) * 2;
                        h(c, i, j) = cplx(h1[idx], h1[idx + 1]) + cplx(h2[idx], h2[idx + 1]);
                    }
                    wn[c] = std::sqrt(hn[c]);
                }
            } else {
            dots(j + 1, W_.ptr, coef, wnb);
            for (int c = 0; c < k; ++c)
                if (cycling[c])
                    for (int i = 0; i <= j; ++i) h(c, i, j) = coef[c][i];
            upload_coeffs(coef, j + 1, cycling);
            bproject(ctx_, V_.ptr, n_ * k_, j + 1, C_.ptr, W_.ptr, n_, k, reinterpret_cast<double*>(scratch_.ptr),
                     reinterpret_cast<double*>(part_.ptr));
            column_sums(wn);
            for (auto& v : wn) v = std::sqrt(v);
            std::vector<char> reo(k, 0);
            bool anyreo = false;
            for (int c = 0; c < k; ++c)
                if (cycling[c] && wn[c] < 0.707 * wnb[c]) {
                    reo[c] = 1;
                    anyreo = true;
                }
            if (anyreo) {
                std::vector<double> dummy;
                dots(j + 1, W_.ptr, re, dummy);
                for (int c = 0; c < k; ++c)
                    if (reo[c]) {
                        out[c].reorth += 1;
                        for (int i = 0; i <= j; ++i) h(c, i, j) += re[c][i];
                    }
                upload_coeffs(re, j + 1, reo);
                bproject(ctx_, V_.ptr, n_ * k_, j + 1, C_.ptr, W_.ptr, n_, k, reinterpret_cast<double*>(scratch_.ptr),
                         reinterpret_cast<double*>(part_.ptr));
                column_sums(wn2);
                for (int c = 0; c < k; ++c)
                    if (reo[c]) wn[c] = std::sqrt(wn2[c]);
            }
            }
            for (int c = 0; c < k; ++c) {
                alpha[c] = 0.0;
                if (!cycling[c]) continue;
                h(c, j + 1, j) = cplx(wn[c], 0.0);
                if (wn[c] < 1e-20)
                    inner_conv[c] = 1;
                else
                    alpha[c] = 1.0 / wn[c];
            }
            upload_alpha(alpha);
            bscale(ctx_, W_.ptr, alpha_.ptr, vec(j + 1), n_, k);
            std::vector<char> fin(k, 0), conv(k, 0);
            std::vector<double> rrel(k, 0.0);
            bool anyfin = false;
            for (int c = 0; c < k; ++c) {
                if (!cycling[c]) continue;
                for (int i = 0; i < j; ++i) {
                    cplx temp = cm(std::conj(cs[c][i]), h(c, i, j)) + cm(std::conj(sn[c][i]), h(c, i + 1, j));
                    h(c, i + 1, j) = (cplx(0.0, 0.0) - cm(sn[c][i], h(c, i, j))) + cm(cs[c][i], h(c, i + 1, j));
                    h(c, i, j) = temp;
                }
                cplx a = h(c, j, j), b = h(c, j + 1, j), cc, ss;
                if (cnorm(b) < 1e-30) {
                    cc = cplx(1.0, 0.0);
                    ss = cplx(0.0, 0.0);
                } else if (cnorm(a) < 1e-30) {
                    cc = cplx(0.0, 0.0);
                    ss = cplx(1.0, 0.0);
                } else {
                    double r = std::sqrt((a.real() * a.real() + a.imag() * a.imag()) +
                                         (b.real() * b.real() + b.imag() * b.imag()));
                    cc = a * (1.0 / r);
                    ss = b * (1.0 / r);
                }
                cs[c].push_back(cc);
                sn[c].push_back(ss);
                h(c, j, j) = cm(std::conj(cc), h(c, j, j)) + cm(std::conj(ss), h(c, j + 1, j));
                h(c, j + 1, j) = cplx(0.0, 0.0);
                cplx temp = cm(std::conj(cc), g[c][j]) + cm(std::conj(ss), g[c][j + 1]);
                g[c][j + 1] = (cplx(0.0, 0.0) - cm(ss, g[c][j])) + cm(cc, g[c][j + 1]);
                g[c][j] = temp;
                double abs_res = cnorm(g[c][j + 1]);
                rrel[c] = abs_res / pb[c];
                conv[c] = rrel[c] < p.tol || (!inner_conv[c] && abs_res < 1e-20);
                if (conv[c] || inner_conv[c]) {
                    fin[c] = 1;
                    anyfin = true;
                }
            }
            if (anyfin) {
                update_x(fin, j + 1);
                std::vector<double> t(k, 0.0);
                if (p.exact_checks) t = true_residuals();
                for (int c = 0; c < k; ++c) {
                    if (!fin[c]) continue;
                    cycling[c] = 0;
                    if (conv[c]) {
                        if (!p.exact_checks) {
                            finish(c, rrel[c], rrel[c], rrel[c], true, SBH_SOLVER_CONVERGED);
                        } else if (t[c] >= p.tol && t[c] >= 1e-20) {
                            out[c].restarts += 1;  // restart from the true residual
                        } else {
                            finish(c, t[c], rrel[c], t[c], true, SBH_SOLVER_CONVERGED);
                        }
                    } else {
                        double tt = p.exact_checks ? t[c] : rrel[c];
                        finish(c, tt, rrel[c], tt, tt < p.tol, tt < p.tol ? SBH_SOLVER_CONVERGED : SBH_SOLVER_BREAKDOWN);
                    }
                }
            }
            if (j == m - 1)
                for (int c = 0; c < k; ++c)
                    if (cycling[c]) end_of_cycle[c] = 1;
        }
        bool anyend = false;
        for (int c = 0; c < k; ++c) anyend = anyend || end_of_cycle[c];
        if (anyend) {
            update_x(end_of_cycle, m);
            std::vector<double> t(k, 0.0);
            if (p.exact_checks) t = true_residuals();
            for (int c = 0; c < k; ++c) {
                if (!end_of_cycle[c]) continue;
                double est = cnorm(g[c][m]) / pb[c];
                double tt = p.exact_checks ? t[c] : est;
                if (tt < p.tol)
                    finish(c, tt, tt, tt, true, SBH_SOLVER_CONVERGED);
                else
                    out[c].restarts += 1;
            }
        }
    }
    bool anyleft = false;
    for (int c = 0; c < k; ++c) anyleft = anyleft || !done[c];
    if (anyleft) {
        std::vector<double> t = true_residuals();
        for (int c = 0; c < k; ++c)
            if (!done[c]) finish(c, t[c], t[c], t[c], false, SBH_SOLVER_MAX_ITERATIONS);
    }
    if (!p.finalize) return;
    // finalize_gmres_solution (gmres.rs:159-250), per column
    refresh_residual();
    const std::vector<double> axn = ax_abs;
    const std::vector<double> rn = r_abs;
    std::vector<double> xn2 = norms(X);
    for (int c = 0; c < k; ++c) {
        ColumnResult& r = out[c];
        if (r.status == SBH_SOLVER_NUMERICAL_ERROR) continue;
        if (!std::isfinite(r.estimated) || r.estimated < 0.0 || !std::isfinite(xn2[c]) || !std::isfinite(axn[c])) {
            r.residual = r.backward_error = INFINITY;
            r.converged = false;
            r.status = SBH_SOLVER_NUMERICAL_ERROR;
            continue;
        }
        double truer = bn[c] > 0.0 ? rn[c] / bn[c] : rn[c];
        r.residual = truer;
        r.backward_error = rn[c] / std::max(axn[c] + bn[c], 2.2250738585072014e-308);
        r.converged = truer <= p.tol;
        if (r.converged)
            r.status = SBH_SOLVER_CONVERGED;
        else if (r.status == SBH_SOLVER_CONVERGED)
            r.status = SBH_SOLVER_MAX_ITERATIONS;
    }
}

// Inner low-precision GMRES as the flexible preconditioner: z ≈ A_lp^{-1} v,
// with the inner Krylov vectors and products in FP32 as in the reference's
// inner solve (gmres_ir.rs:33-48: v is rounded to f32, the f32 result is
// widened back into the FP64 Z basis).
struct InnerGmresPreconditioner : BlockPreconditioner {
    BlockOperatorT<float2>& A;
    BatchedGmres<float2> engine;
    BatchedParams params;
    std::vector<int64_t> inner_iterations;
    sbh_context* ctx;
    int64_t n;
    DeviceArray<float2> V32, Z32;
    InnerGmresPreconditioner(BlockOperatorT<float2>& op, int64_t nn, int k, int max_inner, double tol)
        : A(op), engine(op.ctx, nn, k, max_inner, false), inner_iterations(k, 0), ctx(op.ctx), n(nn) {
        V32.allocate((size_t)nn * k);
        Z32.allocate((size_t)nn * k);
        params.max_iterations = 1;
        params.restart = max_inner;
        params.tol = tol;
        params.exact_checks = false;
        params.finalize = false;
        params.x_zero = true;
        params.cgs2 = true;
    }
    void apply(const double2* V, double2* Z, int k) override {
        convert_block(ctx, V, V32.ptr, n * k);
        zero_fill(ctx, Z32.ptr, n * k);
        std::vector<ColumnResult> r;
        engine.run(A, nullptr, V32.ptr, Z32.ptr, params, r);
        convert_block(ctx, Z32.ptr, Z, n * k);
        for (int c = 0; c < k; ++c) inner_iterations[c] += r[c].iterations;
    }
};

sbh_status run_gmres_block(sbh_operator* op, const double2* d_B, double2* d_X, int k, const sbh_gmres_config& cfg,
                           bool x_zero, sbh_gmres_solution* sols) {
    DenseBlockOperator A(op->sys);
    const int m = (int)std::max<int64_t>(1, std::min<int64_t>(cfg.restart, op->sys->n));
    BatchedGmres<double2> engine(op->ctx, op->sys->n, k, m, false);
    BatchedParams p;
    p.max_iterations = cfg.max_iterations;
    p.restart = cfg.restart;
    p.tol = cfg.tolerance;
    p.x_zero = x_zero;
    std::vector<ColumnResult> res;
    engine.run(A, nullptr, d_B, d_X, p, res);
    for (int c = 0; c < k; ++c) {
        sbh_gmres_solution& s = sols[c];
        s = sbh_gmres_solution{};
        s.iterations = res[c].iterations;
        s.restarts = res[c].restarts;
        s.residual = res[c].residual;
        s.estimated_residual = res[c].estimated;
        s.backward_error = res[c].backward_error;
        s.converged = res[c].converged ? 1 : 0;
        s.status = res[c].status;
        s.spmv_count = A.applies;
        s.reorthogonalizations = res[c].reorth;
    }
    return SBH_OK;
}

sbh_status run_gmres_ir(sbh_operator* op, sbh_lowp_matrix* lp, const double2* d_B, double2* d_X, int k,
                        const sbh_gmres_ir_config& cfg, sbh_gmres_ir_status* st) {
    DenseBlockOperator A(op->sys);
    LowpBlockOperator<float2> Alp(lp);
    const int64_t n = op->sys->n;
    const int max_outer = (int)std::max<int64_t>(1, cfg.max_outer_iterations);
    const int restart = (int)std::max<int64_t>(1, cfg.restart_size > 0 ? std::min<int64_t>(cfg.restart_size, max_outer)
                                                                       : max_outer);
    const int max_inner = (int)std::max<int64_t>(1, std::min<int64_t>(cfg.max_inner_iterations, n));
    InnerGmresPreconditioner M(Alp, n, k, max_inner, cfg.inner_tolerance);
    BatchedGmres<double2> engine(op->ctx, n, k, std::min<int>(restart, (int)n), true);
    BatchedParams p;
    p.restart = restart;
    p.max_iterations = (max_outer + restart - 1) / restart;
    p.tol = cfg.outer_tolerance;
    p.x_zero = true;
    std::vector<ColumnResult> res;
    zero_fill(op->ctx, d_X, n * k);
    engine.run(A, &M, d_B, d_X, p, res);
    for (int c = 0; c < k; ++c) {
        sbh_gmres_ir_status& s = st[c];
        s = sbh_gmres_ir_status{};
        s.outer_iterations = res[c].iterations;
        s.total_inner_iterations = M.inner_iterations[c];
        s.estimated_residual = res[c].estimated;
        s.final_outer_residual = res[c].residual;
        s.backward_error = res[c].backward_error;
        s.converged = res[c].converged ? 1 : 0;
        s.status = res[c].status;
        s.restart_cycles = res[c].restarts + 1;
        s.fp64_products = A.applies;
        s.lowp_products = Alp.applies;
    }
    return SBH_OK;
}

}  // namespace sbh
