// Block GMRES with one shared block Krylov basis and rank-drop deflation:
// block_gmres_preconditioned_with_guess → block_gmres_inner
// (crates/sonium-solvers/crates/sonium-solvers-iterative/src/block_gmres.rs:
// 163-215, 666-1120), the reference's solve_many path
// (BemSolver::solve_many → solve_compressed_many → block_gmres_preconditioned,
// crates/sonium-bem/src/core/bem_solver.rs:2032, 2175-2181).
//
// Device: the basis (column-major, vector i at V + i*n), the block products,
// the block Gram matrices V^H W and the projections W -= V G and X += V Y on
// FP64 MFMA, and the per-column CGS2 QR of each new block with the
// reference's rank-drop test (block_krylov_kernels.hip).  Host: the control
// flow of block_gmres_inner and the small block-Hessenberg least-squares
// problem (solve_block_ls, :614-664).
//
// The least-squares solve is incremental.  The reference refactors the whole
// (p + s') x p Hessenberg with Givens rotations after every block step,
// column by column and row by row.  Rotations of a column only touch rows
// where that column is nonzero, and a step adds columns whose new rows are
// zero in every older column, so the older columns' rotations are the same
// at every step; keeping them and applying them to the new columns (and the
// new rotations to the kept right-hand sides) performs exactly the
// reference's operations on every entry, in its order, without redoing the
// old ones.  The minimiser Y is formed only where the reference uses it
// (an exact residual check, the end of a cycle).
//
// After a step's QR result is read, the next step's device work is queued
// before the host solves the least-squares problem, so the host work overlaps
// the next block product.  That work writes only basis slots beyond the
// current basis and scratch buffers, so a step that ends the cycle leaves
// nothing to undo.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace sbh {

// block_krylov_kernels.hip
int bk_gram_slices(int64_t n, int p);
void bk_gram(sbh_context* ctx, const double2* V, int64_t ldv, int p, const double2* W, int64_t ldw, int sj,
             int64_t n, double* part, double2* gram, double2* hacc, int hmode, const int* flag);
void bk_combine(sbh_context* ctx, const double2* V, int64_t ldv, int p, const double2* C, int ldc, int sj, int64_t n,
                double2* O, int64_t osc, int64_t osk, double alpha, const int* flag);
void bk_to_rowmajor(sbh_context* ctx, const double2* in, int64_t ld, int cols, int64_t n, double2* out);
void bk_to_colmajor(sbh_context* ctx, const double2* in, int cols, int64_t n, const double2* d, double2* out,
                    int64_t ld);
void bk_residual_cm(sbh_context* ctx, const double2* B, const double2* AX, int s, int64_t n, const double2* d,
                    double2* R);
int bk_norm_chunks(int64_t n);
void bk_colnorms(sbh_context* ctx, const double2* X, int64_t ld, int cols, int64_t n, double* part, double* out);
int bk_frob_partials(sbh_context* ctx, const double2* X, int64_t count, double* part);
void bk_dgks_flag(sbh_context* ctx, const double* before, const double* after, int chunks, int* flag, double* frob);
void bk_true_metrics(sbh_context* ctx, const double2* B, const double2* AX, int s, int64_t n, double* part,
                     double* out);
size_t bk_qr_state_bytes();
size_t bk_qr_partial_elems(int64_t n);
void bk_block_qr(sbh_context* ctx, double2* W, double2* Q, int64_t n, int sj, int budget, void* state,
                 double2* part);
void zgemm_f64(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X, int64_t k, double2* Y);

namespace {

using cplx = std::complex<double>;
constexpr double TINY = 1e-30;
constexpr int BK_MAXC = 64;

inline cplx cm(cplx a, cplx b) {  // num-complex Mul
    return cplx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cplx cinv(cplx a) {  // num-complex inv
    const double n = a.real() * a.real() + a.imag() * a.imag();
    return cplx(a.real() / n, -a.imag() / n);
}
inline double cnorm(cplx a) { return std::hypot(a.real(), a.imag()); }
inline double cnorm_sqr(cplx a) { return a.real() * a.real() + a.imag() * a.imag(); }

// givens_rotation (gmres_common.rs:91-107)
inline void givens(cplx a, cplx b, cplx* c, cplx* s) {
    if (cnorm(b) < TINY) {
        *c = cplx(1.0, 0.0);
        *s = cplx(0.0, 0.0);
        return;
    }
    if (cnorm(a) < TINY) {
        *c = cplx(0.0, 0.0);
        *s = cplx(1.0, 0.0);
        return;
    }
    const double r = std::sqrt(cnorm_sqr(a) + cnorm_sqr(b));
    const cplx inv(1.0 / r, 0.0);
    *c = cm(a, inv);
    *s = cm(b, inv);
}

// (x, y) <- (conj(c) x + conj(s) y, (0 - s x) + c y), solve_block_ls's update
inline void rotate(cplx c, cplx s, cplx& x, cplx& y) {
    const cplx a = x, b = y;
    x = cm(std::conj(c), a) + cm(std::conj(s), b);
    y = (cplx(0.0, 0.0) - cm(s, a)) + cm(c, b);
}

template <typename F>
void parallel_for(int count, int threads, F&& f) {
    threads = std::max(1, std::min(threads, count));
    if (threads == 1) {
        for (int i = 0; i < count; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    auto work = [&] {
        for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) f(i);
    };
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

int host_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc ? hc : 1u));
}

// Givens QR of the block Hessenberg matrix, kept across the block steps of a
// cycle (solve_block_ls, block_gmres.rs:614-664).
class BlockLeastSquares {
   public:
    // g = [C0; 0] with C0 the s1 x s initial block-QR coefficients
    void reset(int s, const std::vector<cplx>& c0, int s1) {
        s_ = s;
        col_.clear();
        rot_.clear();
        gw_.assign((size_t)s1 * s, cplx(0.0, 0.0));
        std::copy(c0.begin(), c0.begin() + (size_t)s1 * s, gw_.begin());
    }

    // Appends columns [p_start, p) (entries for rows 0..rows-1 each) and
    // factors them; the transformed right-hand sides grow to `rows` rows.
    void add_columns(int p_start, int p, int rows, std::vector<std::vector<cplx>>&& cols, int threads) {
        col_.resize(p);
        rot_.resize(p);
        for (int j = p_start; j < p; ++j) col_[j] = std::move(cols[j - p_start]);
        if (gw_.size() < (size_t)rows * s_) gw_.resize((size_t)rows * s_, cplx(0.0, 0.0));
        // rotations of the older columns, in their order, on each new column
        parallel_for(p - p_start, threads, [&](int q) {
            std::vector<cplx>& h = col_[p_start + q];
            for (int col = 0; col < p_start; ++col)
                for (const Rot& r : rot_[col]) rotate(r.c, r.s, h[col], h[r.row]);
        });
        // the new columns, in the reference's order
        for (int col = p_start; col < p; ++col) {
            rot_[col].clear();
            for (int row = col + 1; row < rows; ++row) {
                if (cnorm(col_[col][row]) < TINY) continue;
                cplx c, s;
                givens(col_[col][col], col_[col][row], &c, &s);
                for (int j = col; j < p; ++j) rotate(c, s, col_[j][col], col_[j][row]);
                for (int q = 0; q < s_; ++q) rotate(c, s, gw_[(size_t)col * s_ + q], gw_[(size_t)row * s_ + q]);
                rot_[col].push_back(Rot{row, c, s});
            }
        }
    }

    // Per right-hand side: norm of rows [cols, rows) of the transformed g.
    std::vector<double> tail(int cols, int rows) const {
        std::vector<double> t(s_, 0.0);
        for (int i = cols; i < rows; ++i)
            for (int q = 0; q < s_; ++q) t[q] += cnorm_sqr(gw_[(size_t)i * s_ + q]);
        for (double& v : t) v = std::sqrt(v);
        return t;
    }

    // Y (cols x s, row-major) by back substitution per right-hand side
    // (solve_upper_triangular, gmres_common.rs:110-127).  Columns the basis
    // has but the Hessenberg does not yet (all zero) give y = 0.
    std::vector<cplx> solve(int cols, int threads) const {
        std::vector<cplx> y((size_t)cols * s_, cplx(0.0, 0.0));
        parallel_for(s_, threads, [&](int q) {
            std::vector<cplx> yq(cols, cplx(0.0, 0.0));
            for (int i = cols - 1; i >= 0; --i) {
                cplx sum = i < (int)gw_.size() / s_ ? gw_[(size_t)i * s_ + q] : cplx(0.0, 0.0);
                for (int j = i + 1; j < cols; ++j) sum -= cm(entry(i, j), yq[j]);
                const cplx d = entry(i, i);
                if (cnorm(d) > TINY) yq[i] = cm(sum, cinv(d));
            }
            for (int i = 0; i < cols; ++i) y[(size_t)i * s_ + q] = yq[i];
        });
        return y;
    }

   private:
    struct Rot {
        int row;
        cplx c, s;
    };
    int s_ = 0;
    std::vector<std::vector<cplx>> col_;  // rotated Hessenberg columns
    std::vector<std::vector<Rot>> rot_;   // rotations of each column, in order
    std::vector<cplx> gw_;                // transformed right-hand sides, rows x s

    cplx entry(int i, int j) const {
        if (j >= (int)col_.size()) return cplx(0.0, 0.0);
        const std::vector<cplx>& c = col_[j];
        return i < (int)c.size() ? c[i] : cplx(0.0, 0.0);
    }
};

// Layout of the QR state written by the device chain (mirrors QrState).
struct HostQrState {
    int snext[BK_MAXC + 1];
    int accepted[BK_MAXC];
    unsigned counter;
    int pad[2];
    double orig[BK_MAXC];
    double wnorm[BK_MAXC];
    double r[BK_MAXC][BK_MAXC][2];
    double coef[BK_MAXC][2];
};

}  // namespace

struct BlockGmresParams {
    int64_t max_iterations = 100;
    int64_t restart = 30;
    double tol = 1e-6;
    bool true_residual_check = true;
    int64_t exact_interval = 0;
};

struct BlockGmresResult {
    int64_t iterations = 0, spmv = 0, exact_checks = 0, cycles = 0, max_basis = 0, deflated = 0;
    int64_t cholqr_blocks = 0, chain_blocks = 0;
    bool converged = false;
    std::vector<double> projected, true_res, backward;
};

class BlockGmresEngine {
   public:
    BlockGmresEngine(sbh_operator* op, const sbh_precond* M, int s, int64_t restart)
        : op_(op), ctx_(op->ctx), n_(op->n), s_(s), dinv_(M && M->kind == 1 ? M->d_dinv.ptr : nullptr) {
        max_steps_ = (int)std::max<int64_t>(1, restart);
        max_basis_ = (max_steps_ + 1) * s;
        const size_t blk = (size_t)n_ * s;
        admit_device(ctx_, (int64_t)(sizeof(double2) * ((size_t)(max_basis_ + s) * n_ + 6 * blk)),
                     "the shared block Krylov basis", false);
        V_.allocate((size_t)(max_basis_ + s) * n_);
        W_.allocate(blk);
        Rcm_.allocate(blk);
        Xc_.allocate(blk);
        AX_.allocate(blk);
        Vrm_.allocate(blk);
        Yrm_.allocate(blk);
        int64_t part = 0;
        for (int p = 1; p <= max_basis_; ++p)
            part = std::max<int64_t>(part, (int64_t)bk_gram_slices(n_, p) * p);
        gpart_.allocate((size_t)part * 2 * s + 16);
        gram_.allocate((size_t)max_basis_ * s + 16);
        // hacc (max_basis x s) followed by the QR state, read back together
        const size_t hbytes = sizeof(double2) * (size_t)max_basis_ * s;
        hacc_off_ = 0;
        qr_off_ = (hbytes + 255) / 256 * 256;
        dl_bytes_ = qr_off_ + bk_qr_state_bytes();
        dl_dev_.allocate((dl_bytes_ + 15) / 16);
        dl_host_.allocate((dl_bytes_ + 7) / 8);
        qpart_.allocate(bk_qr_partial_elems(n_));
        nchunks_ = bk_norm_chunks((int64_t)n_ * s);
        npart_.allocate((size_t)std::max<int64_t>(3 * s, 4) * bk_norm_chunks(n_) + 2 * nchunks_ + 64);
        small_.allocate(4 * (size_t)s + 64);
        flag_.allocate(4);
        small_host_.allocate(4 * (size_t)s + 64);
        qg_.allocate((size_t)s * s + 16);
        qg_host_.allocate((size_t)2 * s * s + 32);
        SBH_REQUIRE(sizeof(HostQrState) == bk_qr_state_bytes(), "block QR state layout mismatch");
        // the QR chain's workgroup counter starts at zero and every launch leaves it there
        zero_on_stream(ctx_, dl_dev_.ptr, dl_dev_.bytes());
    }

    void run(const double2* B, double2* X, bool x_zero, const BlockGmresParams& prm, BlockGmresResult& out);

   private:
    sbh_operator* op_;
    sbh_context* ctx_;
    int64_t n_;
    int s_;
    const double2* dinv_;
    int max_steps_ = 0, max_basis_ = 0;
    DeviceArray<double2> V_, W_, Rcm_, Xc_, AX_, Vrm_, Yrm_, gram_, qpart_, dl_dev_;
    DeviceArray<double> gpart_, npart_, small_;
    DeviceArray<int> flag_;
    PinnedArray<double> dl_host_, small_host_;
    DeviceArray<double2> qg_;     // sj x sj Gram matrix of the block being factored
    PinnedArray<double> qg_host_;  // its host copy
    HostQrState qrh_;              // outcome of the block QR, whichever way it ran
    int64_t cholqr_blocks_ = 0, chain_blocks_ = 0;
    size_t hacc_off_ = 0, qr_off_ = 0, dl_bytes_ = 0;
    int nchunks_ = 1;
    int64_t spmv_ = 0, exact_checks_ = 0;

    double2* vec(int i) { return V_.ptr + (size_t)i * n_; }
    double2* hacc_dev() { return reinterpret_cast<double2*>(reinterpret_cast<char*>(dl_dev_.ptr) + hacc_off_); }
    void* qr_dev() { return reinterpret_cast<char*>(dl_dev_.ptr) + qr_off_; }
    const HostQrState* qr_host() const { return &qrh_; }
    const double2* hacc_host() const {
        return reinterpret_cast<const double2*>(reinterpret_cast<const char*>(dl_host_.ptr) + hacc_off_);
    }

    // Y (n x k row-major) = A X (n x k row-major)
    void apply_rm(const double2* X, double2* Y, int k) {
        if (op_->dense()) {
            if (k == 1)
                zgemv_n(ctx_, op_->sys->d_A.ptr, op_->sys->lda, n_, n_, X, Y);
            else
                zgemm_f64(ctx_, op_->sys->d_A.ptr, op_->sys->lda, n_, X, k, Y);
        } else {
            // column by column through column-major scratch
            bk_to_colmajor(ctx_, X, k, n_, nullptr, Vrm_.ptr, n_);
            for (int c = 0; c < k; ++c) operator_apply_device(op_, Vrm_.ptr + (size_t)c * n_, Yrm_.ptr + (size_t)c * n_);
            bk_to_rowmajor(ctx_, Yrm_.ptr, n_, k, n_, Y);
        }
        spmv_ += k;
    }

    // W (column-major, sj vectors) = M^-1 A V[first .. first + sj)
    void apply_cm(int first, int sj) {
        if (op_->dense()) {
            bk_to_rowmajor(ctx_, vec(first), n_, sj, n_, Vrm_.ptr);
            if (sj == 1)
                zgemv_n(ctx_, op_->sys->d_A.ptr, op_->sys->lda, n_, n_, Vrm_.ptr, Yrm_.ptr);
            else
                zgemm_f64(ctx_, op_->sys->d_A.ptr, op_->sys->lda, n_, Vrm_.ptr, sj, Yrm_.ptr);
            bk_to_colmajor(ctx_, Yrm_.ptr, sj, n_, dinv_, W_.ptr, n_);
        } else {
            for (int c = 0; c < sj; ++c) {
                operator_apply_device(op_, vec(first + c), Yrm_.ptr);
                if (dinv_)
                    diag_scale(ctx_, dinv_, Yrm_.ptr, W_.ptr + (size_t)c * n_, n_);
                else
                    device_copy(ctx_, Yrm_.ptr, W_.ptr + (size_t)c * n_, n_);
            }
        }
        spmv_ += sj;
    }

    // One block Arnoldi step on the device: W = M^-1 A V_j, block CGS
    // against V[0:p] with the DGKS trigger, then the QR of W into V[p:]
    // (block_gmres.rs:918-1025).  Nothing is read back here.
    void queue_step(int p, int sj) {
        const int p_start = p - sj;
        apply_cm(p_start, sj);
        double* before = npart_.ptr;
        double* after = npart_.ptr + nchunks_;
        double* frob = small_.ptr;
        const int chunks = bk_frob_partials(ctx_, W_.ptr, (int64_t)sj * n_, before);
        bk_gram(ctx_, V_.ptr, n_, p, W_.ptr, n_, sj, n_, gpart_.ptr, gram_.ptr, hacc_dev(), 1, nullptr);
        bk_combine(ctx_, V_.ptr, n_, p, gram_.ptr, sj, sj, n_, W_.ptr, n_, 1, -1.0, nullptr);
        bk_frob_partials(ctx_, W_.ptr, (int64_t)sj * n_, after);
        bk_dgks_flag(ctx_, before, after, chunks, flag_.ptr, frob);
        bk_gram(ctx_, V_.ptr, n_, p, W_.ptr, n_, sj, n_, gpart_.ptr, gram_.ptr, hacc_dev(), 2, flag_.ptr);
        bk_combine(ctx_, V_.ptr, n_, p, gram_.ptr, sj, sj, n_, W_.ptr, n_, 1, -1.0, flag_.ptr);
        bk_gram(ctx_, W_.ptr, n_, sj, W_.ptr, n_, sj, n_, gpart_.ptr, qg_.ptr, nullptr, 0, nullptr);
    }

    // Reads the step's Hessenberg block (p x sj) and forms the QR of W into
    // the basis slots from p on.
    void read_step(int p, int sj) {
        SBH_HIP(hipMemcpyAsync(dl_host_.ptr, dl_dev_.ptr, sizeof(double2) * (size_t)p * sj, hipMemcpyDeviceToHost,
                               ctx_->stream));
        block_qr(p, sj, max_basis_ - p);
    }

    double2* qg_host() { return reinterpret_cast<double2*>(qg_host_.ptr); }

    // sj x sj Gram matrix qg_ -> host (the stream is synchronized)
    std::vector<cplx> fetch_gram(int sj) {
        SBH_HIP(hipMemcpyAsync(qg_host(), qg_.ptr, sizeof(double2) * (size_t)sj * sj, hipMemcpyDeviceToHost,
                               ctx_->stream));
        SBH_HIP(hipStreamSynchronize(ctx_->stream));
        std::vector<cplx> g((size_t)sj * sj);
        for (size_t i = 0; i < g.size(); ++i) g[i] = cplx(qg_host()[i].x, qg_host()[i].y);
        return g;
    }

    // Upper R with R^H R = G (G Hermitian, sj x sj row-major); false when a
    // pivot is not clearly positive: below `floor` times sqrt(G_cc).
    static bool cholesky(const std::vector<cplx>& G, int m, double floor, std::vector<cplx>& R,
                         std::vector<double>* diag_norm) {
        R.assign((size_t)m * m, cplx(0.0, 0.0));
        for (int j = 0; j < m; ++j) {
            double d = G[(size_t)j * m + j].real();
            const double gjj = d;
            for (int k = 0; k < j; ++k) d -= cnorm_sqr(R[(size_t)k * m + j]);
            if (!(gjj > 0.0) || !(d > floor * floor * gjj) || !std::isfinite(d)) return false;
            const double rjj = std::sqrt(d);
            R[(size_t)j * m + j] = cplx(rjj, 0.0);
            if (diag_norm) (*diag_norm)[j] = std::sqrt(gjj);
            for (int i = j + 1; i < m; ++i) {
                cplx v = G[(size_t)j * m + i];
                for (int k = 0; k < j; ++k) v -= std::conj(R[(size_t)k * m + j]) * R[(size_t)k * m + i];
                R[(size_t)j * m + i] = v / rjj;
            }
        }
        return true;
    }

    // T = R^-1 (upper triangular, positive real diagonal)
    static std::vector<cplx> triangular_inverse(const std::vector<cplx>& R, int m) {
        std::vector<cplx> T((size_t)m * m, cplx(0.0, 0.0));
        for (int c = 0; c < m; ++c) {
            T[(size_t)c * m + c] = cplx(1.0 / R[(size_t)c * m + c].real(), 0.0);
            for (int i = c - 1; i >= 0; --i) {
                cplx acc(0.0, 0.0);
                for (int k = i + 1; k <= c; ++k) acc += R[(size_t)i * m + k] * T[(size_t)k * m + c];
                T[(size_t)i * m + c] = -acc / R[(size_t)i * m + i].real();
            }
        }
        return T;
    }

    // out (column-major, sj vectors) = in * T, T sj x sj upper (coefficients through gram_)
    void apply_triangular(const double2* in, const std::vector<cplx>& T, int sj, double2* out) {
        std::vector<double2> th(T.size());
        for (size_t i = 0; i < T.size(); ++i) th[i] = make_double2(T[i].real(), T[i].imag());
        SBH_HIP(hipMemcpyAsync(gram_.ptr, th.data(), sizeof(double2) * th.size(), hipMemcpyHostToDevice,
                               ctx_->stream));
        SBH_HIP(hipMemsetAsync(out, 0, sizeof(double2) * (size_t)sj * n_, ctx_->stream));
        bk_combine(ctx_, in, n_, sj, gram_.ptr, sj, sj, n_, out, n_, 1, 1.0, nullptr);
        SBH_HIP(hipStreamSynchronize(ctx_->stream));  // th must outlive the copy
    }

    // QR of the sj column-major vectors in W_ (its Gram matrix already in
    // qg_) into the basis slots dst..: Cholesky QR twice when every pivot is
    // clearly positive (the block is far from the rank-drop threshold) and the
    // block fits the budget, which gives the per-column CGS2 factors to
    // rounding; otherwise the per-column CGS2 chain with rank-drop deflation.
    // W_ is left unchanged by the first way, so the second can always follow.
    void block_qr(int dst, int sj, int budget) {
        constexpr double kPivotFloor = 1e-4;  // four orders above sqrt(eps)
        std::vector<cplx> G1 = fetch_gram(sj), R1, R2;
        std::vector<double> orig(sj, 0.0);
        bool ok = sj <= budget && cholesky(G1, sj, kPivotFloor, R1, &orig);
        if (ok) {
            apply_triangular(W_.ptr, triangular_inverse(R1, sj), sj, Yrm_.ptr);
            bk_gram(ctx_, Yrm_.ptr, n_, sj, Yrm_.ptr, n_, sj, n_, gpart_.ptr, qg_.ptr, nullptr, 0, nullptr);
            ok = cholesky(fetch_gram(sj), sj, 0.5, R2, nullptr);
        }
        if (ok) {
            apply_triangular(Yrm_.ptr, triangular_inverse(R2, sj), sj, vec(dst));
            qrh_ = HostQrState{};
            for (int c = 0; c < sj; ++c) {
                // R = R2 R1: column c of the product
                for (int a = 0; a <= c; ++a) {
                    cplx v(0.0, 0.0);
                    for (int k = a; k <= c; ++k) v += R2[(size_t)a * sj + k] * R1[(size_t)k * sj + c];
                    if (a < c) {
                        qrh_.r[c][a][0] = v.real();
                        qrh_.r[c][a][1] = v.imag();
                    } else {
                        qrh_.wnorm[c] = v.real();
                    }
                }
                qrh_.snext[c] = c;
                qrh_.accepted[c] = c;
                qrh_.orig[c] = orig[c];
            }
            qrh_.snext[sj] = sj;
            cholqr_blocks_ += 1;
            return;
        }
        bk_block_qr(ctx_, W_.ptr, vec(dst), n_, sj, budget, qr_dev(), qpart_.ptr);
        SBH_HIP(hipMemcpyAsync(&qrh_, qr_dev(), sizeof(HostQrState), hipMemcpyDeviceToHost, ctx_->stream));
        SBH_HIP(hipStreamSynchronize(ctx_->stream));
        chain_blocks_ += 1;
    }

    std::vector<double> read_small(int count) {
        SBH_HIP(hipMemcpyAsync(small_host_.ptr, small_.ptr, sizeof(double) * count, hipMemcpyDeviceToHost,
                               ctx_->stream));
        SBH_HIP(hipStreamSynchronize(ctx_->stream));
        return std::vector<double>(small_host_.ptr, small_host_.ptr + count);
    }

    // per-column norms of a row-major n x s block: through the column-major scratch
    std::vector<double> colnorms_rm(const double2* X) {
        bk_to_colmajor(ctx_, X, s_, n_, nullptr, Rcm_.ptr, n_);
        bk_colnorms(ctx_, Rcm_.ptr, n_, s_, n_, npart_.ptr, small_.ptr);
        return read_small(s_);
    }

    // block_true_metrics (block_gmres.rs:567-612): ||b - A x||/||b|| and the
    // normwise backward error per column, with one block product.
    void true_metrics(const double2* B, const double2* Xd, const std::vector<double>& b_true,
                      std::vector<double>& tr, std::vector<double>& be) {
        apply_rm(Xd, AX_.ptr, s_);
        exact_checks_ += 1;
        bk_true_metrics(ctx_, B, AX_.ptr, s_, n_, npart_.ptr, small_.ptr);
        const std::vector<double> m = read_small(3 * s_);
        tr.assign(s_, 0.0);
        be.assign(s_, 0.0);
        for (int j = 0; j < s_; ++j) {
            const double ar = m[j], axn = m[s_ + j], bn = m[2 * s_ + j];
            tr[j] = b_true[j] <= TINY ? ar : ar / b_true[j];
            be[j] = ar / std::max(axn + bn, 2.2250738585072014e-308);
        }
    }

    // X (row-major) += V[0:p] y (y p x s row-major), apply_block_update
    void update(double2* Xd, const std::vector<cplx>& y, int p) {
        if (p <= 0) return;
        std::vector<double2> yh((size_t)p * s_);
        for (size_t i = 0; i < yh.size(); ++i) yh[i] = make_double2(y[i].real(), y[i].imag());
        SBH_HIP(hipMemcpyAsync(gram_.ptr, yh.data(), sizeof(double2) * yh.size(), hipMemcpyHostToDevice,
                               ctx_->stream));
        bk_combine(ctx_, V_.ptr, n_, p, gram_.ptr, s_, s_, n_, Xd, 1, s_, 1.0, nullptr);
        // the staging vector must outlive the copy
        SBH_HIP(hipStreamSynchronize(ctx_->stream));
    }
};

void BlockGmresEngine::run(const double2* B, double2* X, bool x_zero, const BlockGmresParams& prm,
                           BlockGmresResult& out) {
    const int s = s_;
    const int threads = host_threads();
    out = BlockGmresResult{};
    out.projected.assign(s, INFINITY);
    // ||M^-1 b_j|| for the convergence test and ||b_j|| for the true residuals
    std::vector<double> b_norms, b_true;
    bk_residual_cm(ctx_, B, nullptr, s, n_, dinv_, Rcm_.ptr);
    bk_colnorms(ctx_, Rcm_.ptr, n_, s, n_, npart_.ptr, small_.ptr);
    b_norms = read_small(s);
    b_true = colnorms_rm(B);
    std::vector<double>& projected = out.projected;
    std::vector<double> col_norms(s);
    int64_t total_iters = 0;
    BlockLeastSquares ls;
    bool x_known_zero = x_zero;

    for (int64_t cycle = 0; cycle < prm.max_iterations; ++cycle) {
        out.cycles += 1;
        // R0 = M^-1 (B - A X), column-major
        if (x_known_zero) {
            // A X = 0: the product is not formed (it is still counted)
            bk_residual_cm(ctx_, B, nullptr, s, n_, dinv_, Rcm_.ptr);
            spmv_ += s;
        } else {
            apply_rm(X, AX_.ptr, s);
            bk_residual_cm(ctx_, B, AX_.ptr, s, n_, dinv_, Rcm_.ptr);
        }
        bk_colnorms(ctx_, Rcm_.ptr, n_, s, n_, npart_.ptr, small_.ptr);
        col_norms = read_small(s);
        bool all_converged = true;
        for (int j = 0; j < s; ++j) {
            projected[j] = b_norms[j] >= TINY ? col_norms[j] / b_norms[j] : 0.0;
            if (b_norms[j] >= TINY && projected[j] > prm.tol) all_converged = false;
        }
        if (all_converged) break;
        // initial block QR with rank-drop deflation (CGS2 per column) of a copy of R0
        device_copy(ctx_, Rcm_.ptr, W_.ptr, n_ * s);
        bk_gram(ctx_, W_.ptr, n_, s, W_.ptr, n_, s, n_, gpart_.ptr, qg_.ptr, nullptr, 0, nullptr);
        block_qr(0, s, 1 << 30);
        const int s1 = qr_host()->snext[s];
        out.deflated += s - s1;
        if (s1 == 0) break;
        // C0 = V1^H R0 (s1 x s)
        bk_gram(ctx_, V_.ptr, n_, s1, Rcm_.ptr, n_, s, n_, gpart_.ptr, gram_.ptr, nullptr, 0, nullptr);
        std::vector<cplx> c0((size_t)s1 * s);
        {
            std::vector<double2> g((size_t)s1 * s);
            SBH_HIP(hipMemcpyAsync(g.data(), gram_.ptr, sizeof(double2) * g.size(), hipMemcpyDeviceToHost,
                                   ctx_->stream));
            SBH_HIP(hipStreamSynchronize(ctx_->stream));
            for (size_t i = 0; i < g.size(); ++i) c0[i] = cplx(g[i].x, g[i].y);
        }
        ls.reset(s, c0, s1);
        int p = s1;
        int sj = s1;
        bool converged_cycle = false;
        bool restart_cycle = false;
        bool queued = false;
        for (int step = 0; step < max_steps_; ++step) {
            if (p >= max_basis_) break;
            const int p_start = p - sj;
            if (!queued) queue_step(p, sj);
            read_step(p, sj);
            const HostQrState* q = qr_host();
            const int s_next = q->snext[sj];
            out.deflated += sj - s_next;
            // the new Hessenberg columns p_start .. p-1, rows 0 .. p + s_next - 1
            const int rows = p + s_next;
            std::vector<std::vector<cplx>> cols(sj, std::vector<cplx>(rows, cplx(0.0, 0.0)));
            const double2* hh = hacc_host();
            for (int c = 0; c < sj; ++c) {
                std::vector<cplx>& col = cols[c];
                for (int i = 0; i < p; ++i) col[i] = cplx(hh[(size_t)i * sj + c].x, hh[(size_t)i * sj + c].y);
                for (int a = 0; a < q->snext[c]; ++a) col[p + a] = cplx(q->r[c][a][0], q->r[c][a][1]);
                if (q->accepted[c] >= 0) col[p + q->accepted[c]] = cplx(q->wnorm[c], 0.0);
            }
            // queue the next step's device work before the host solve
            queued = s_next > 0 && step + 1 < max_steps_ && p + s_next < max_basis_;
            if (queued) queue_step(p + s_next, s_next);
            total_iters += 1;
            ls.add_columns(p_start, p, rows, std::move(cols), threads);
            const std::vector<double> tail = ls.tail(p, rows);
            for (int j = 0; j < s; ++j) projected[j] = b_norms[j] >= TINY ? tail[j] / b_norms[j] : 0.0;
            bool cycle_done = true;
            for (int j = 0; j < s; ++j) cycle_done = cycle_done && projected[j] <= prm.tol;
            const bool interval_due =
                prm.exact_interval > 0 && total_iters > 0 && total_iters % prm.exact_interval == 0;
            if (s_next > 0 && (interval_due || (cycle_done && prm.true_residual_check))) {
                const std::vector<cplx> y = ls.solve(p, threads);
                device_copy(ctx_, X, Xc_.ptr, n_ * s);
                update(Xc_.ptr, y, p);
                std::vector<double> tr, be;
                true_metrics(B, Xc_.ptr, b_true, tr, be);
                bool exact_converged = true;
                for (int j = 0; j < s; ++j) exact_converged = exact_converged && tr[j] <= prm.tol;
                if (exact_converged) {
                    device_copy(ctx_, Xc_.ptr, X, n_ * s);
                    x_known_zero = false;
                    converged_cycle = true;
                    break;
                }
                if (cycle_done && prm.true_residual_check) {
                    device_copy(ctx_, Xc_.ptr, X, n_ * s);
                    x_known_zero = false;
                    restart_cycle = true;
                    break;
                }
            }
            if (cycle_done || s_next == 0) {
                update(X, ls.solve(p, threads), p);
                x_known_zero = false;
                converged_cycle = cycle_done;
                break;
            }
            p += s_next;
            sj = s_next;
            out.max_basis = std::max<int64_t>(out.max_basis, p);
        }
        out.max_basis = std::max<int64_t>(out.max_basis, p);
        if (restart_cycle) continue;
        if (!converged_cycle && p > s1) {
            // budget exhausted: the square LS over the whole basis; its tail is
            // empty, as in the reference (rows == cols)
            for (int j = 0; j < s; ++j) projected[j] = 0.0;
            update(X, ls.solve(p, threads), p);
            x_known_zero = false;
        }
        if (converged_cycle) break;
    }
    std::vector<double> tr, be;
    if (x_known_zero) {
        // A*0 = 0: the true residual is ||b|| / ||b||
        tr.assign(s, 0.0);
        be.assign(s, 0.0);
        for (int j = 0; j < s; ++j) {
            tr[j] = b_true[j] <= TINY ? b_true[j] : 1.0;
            be[j] = b_true[j] / std::max(b_true[j], 2.2250738585072014e-308);
        }
        spmv_ += s;
        exact_checks_ += 1;
    } else {
        true_metrics(B, X, b_true, tr, be);
    }
    bool conv = true;
    for (int j = 0; j < s; ++j) conv = conv && (prm.true_residual_check ? tr[j] <= prm.tol : projected[j] <= prm.tol);
    out.iterations = total_iters;
    out.true_res = tr;
    out.backward = be;
    out.converged = conv;
    out.spmv = spmv_;
    out.exact_checks = exact_checks_;
    out.cholqr_blocks = cholqr_blocks_;
    out.chain_blocks = chain_blocks_;
}

sbh_status run_block_gmres(sbh_operator* op, const sbh_precond* M, const double2* d_B, double2* d_X, int k,
                           bool x_zero, const sbh_block_gmres_config& cfg, sbh_block_gmres_info* info,
                           double* residuals, double* true_residuals, double* backward_errors) {
    BlockGmresParams prm;
    prm.max_iterations = cfg.max_iterations;
    prm.restart = cfg.restart;
    prm.tol = cfg.tolerance;
    prm.true_residual_check = cfg.true_residual_check != 0;
    prm.exact_interval = cfg.exact_residual_check_interval;
    BlockGmresEngine engine(op, M, k, cfg.restart);
    BlockGmresResult r;
    engine.run(d_B, d_X, x_zero, prm, r);
    bool finite = true;
    for (int j = 0; j < k; ++j)
        finite = finite && std::isfinite(r.projected[j]) && std::isfinite(r.true_res[j]) &&
                 std::isfinite(r.backward[j]);
    if (info) {
        *info = sbh_block_gmres_info{};
        info->iterations = r.iterations;
        info->converged = r.converged ? 1 : 0;
        info->status = !finite ? SBH_SOLVER_NUMERICAL_ERROR
                               : (r.converged ? SBH_SOLVER_CONVERGED : SBH_SOLVER_MAX_ITERATIONS);
        info->spmv_count = r.spmv;
        info->exact_residual_checks = r.exact_checks;
        info->restart_cycles = r.cycles;
        info->max_basis = r.max_basis;
        info->deflated = r.deflated;
        info->cholesky_qr_blocks = r.cholqr_blocks;
        info->chain_qr_blocks = r.chain_blocks;
    }
    for (int j = 0; j < k; ++j) {
        if (residuals) residuals[j] = r.projected[j];
        if (true_residuals) true_residuals[j] = r.true_res[j];
        if (backward_errors) backward_errors[j] = r.backward[j];
    }
    if (!finite) throw Error(SBH_ERR_NUMERICAL, "block GMRES generated NaN or infinity in residual metrics");
    return SBH_OK;
}

}  // namespace sbh
