// GCRO-DR: GMRES with a recycled subspace carried from one system to the next.
//
// Follows gcro_dr_core (crates/sonium-solvers/crates/sonium-solvers-iterative/
// src/gcro_dr.rs:292-760) with build_c_r (:770-860) and the harmonic Ritz
// update (:905-1050), whose small dense eigensolver is gmres_dr.rs's
// Householder reduction (:1857-1950) and shifted QR (:1306-1460,1589-1650).
// Parks, de Sturler, Mackey, Johnson, Maiti, SIAM J. Sci. Comput. 28(5), 2006.
//
// Every n-vector (U, C, the Krylov basis, the iterate) stays on the device;
// the projections against C and the Arnoldi basis are the one-read CGS passes
// of the GMRES driver.  The (k + m) Ritz problem runs on the host in FP64.
//
// Kept from the reference on purpose: its Hessenberg reduction uses a real
// alpha (-sign(Re x0) ||x||) and then zeroes everything below the subdiagonal,
// and its un-rotation of the Givens-rotated Hessenberg uses conj(s) for the
// lower row, so for complex data the eigenproblem it solves is a perturbation
// of the augmented matrix; the same subspace is selected here.  Departures:
// numerically real trailing 2x2 blocks with a conjugate pair continue with the
// complex single-shift iteration (the reference switches to a real Francis
// double shift and a quasi-triangular eigenvector extraction), and the lifted
// Ritz vectors are orthonormalised by two classical Gram-Schmidt passes
// instead of one modified pass (the same vectors up to rounding).

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <vector>

#include "krylov.hpp"

namespace sbh {
using namespace krylov;

namespace {

// Dense complex matrix, row-major.
struct Mat {
    int rows = 0, cols = 0;
    std::vector<cplx> a;
    Mat() = default;
    Mat(int r, int c) : rows(r), cols(c), a((size_t)r * c, cplx(0.0, 0.0)) {}
    cplx& operator()(int i, int j) { return a[(size_t)i * cols + j]; }
    cplx operator()(int i, int j) const { return a[(size_t)i * cols + j]; }
};

inline double norm_sqr(cplx z) { return z.real() * z.real() + z.imag() * z.imag(); }

std::pair<cplx, cplx> givens(cplx a, cplx b) {
    if (cnorm(b) < 1e-30) return {cplx(1.0, 0.0), cplx(0.0, 0.0)};
    if (cnorm(a) < 1e-30) return {cplx(0.0, 0.0), cplx(1.0, 0.0)};
    const double r = std::sqrt(norm_sqr(a) + norm_sqr(b));
    return {a * (1.0 / r), b * (1.0 / r)};
}

// R y = g for the leading k x k upper triangle; pivots below 1e-30 give 0.
std::vector<cplx> upper_solve(const Mat& r, const std::vector<cplx>& g, int k) {
    std::vector<cplx> y(k, cplx(0.0, 0.0));
    for (int i = k - 1; i >= 0; --i) {
        cplx s = g[i];
        for (int j = i + 1; j < k; ++j) s -= cm(r(i, j), y[j]);
        if (cnorm(r(i, i)) > 1e-30) y[i] = cm(s, cinv(r(i, i)));
    }
    return y;
}

// The raw Hessenberg block from the rotated one (gcro_dr.rs:884-905).
Mat unapply_givens(const Mat& hr, const std::vector<cplx>& cs, const std::vector<cplx>& sn, int steps) {
    Mat h = hr;
    for (int j = steps - 1; j >= 0; --j) {
        if (j >= (int)cs.size()) continue;
        const cplx c = cs[j], s = sn[j];
        for (int col = j; col < steps; ++col) {
            const cplx a = h(j, col), b = h(j + 1, col);
            h(j, col) = cm(c, a) - cm(std::conj(s), b);
            h(j + 1, col) = cm(std::conj(s), a) + cm(std::conj(c), b);
        }
    }
    return h;
}

// Householder reduction with the real alpha, then the sub-subdiagonal zeroed.
void to_hessenberg(Mat& h, Mat& q) {
    const int n = h.rows;
    q = Mat(n, n);
    for (int i = 0; i < n; ++i) q(i, i) = 1.0;
    std::vector<cplx> v(n);
    for (int j = 0; j + 2 < n; ++j) {
        double xs = 0.0;
        for (int i = j + 1; i < n; ++i) xs += norm_sqr(h(i, j));
        const double xn = std::sqrt(xs);
        if (xn < 1e-30) continue;
        const cplx x0 = h(j + 1, j);
        const double alpha = x0.real() >= 0.0 ? -xn : xn;
        const int len = n - j - 1;
        v[0] = x0 - alpha;
        for (int i = 1; i < len; ++i) v[i] = h(j + 1 + i, j);
        double vs = 0.0;
        for (int i = 0; i < len; ++i) vs += norm_sqr(v[i]);
        if (vs < 1e-30) continue;
        const double beta = 2.0 / vs;
        for (int col = j; col < n; ++col) {
            cplx d(0.0, 0.0);
            for (int l = 0; l < len; ++l) d += std::conj(v[l]) * h(j + 1 + l, col);
            for (int l = 0; l < len; ++l) h(j + 1 + l, col) -= beta * v[l] * d;
        }
        for (int row = 0; row < n; ++row) {
            cplx d(0.0, 0.0);
            for (int l = 0; l < len; ++l) d += h(row, j + 1 + l) * v[l];
            for (int l = 0; l < len; ++l) h(row, j + 1 + l) -= beta * d * std::conj(v[l]);
            cplx e(0.0, 0.0);
            for (int l = 0; l < len; ++l) e += q(row, j + 1 + l) * v[l];
            for (int l = 0; l < len; ++l) q(row, j + 1 + l) -= beta * e * std::conj(v[l]);
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = j + 2; i < n; ++i) h(i, j) = 0.0;
}

cplx wilkinson(cplx a, cplx b, cplx c, cplx d) {
    const cplx th = (a + d) * 0.5;
    const cplx dp = ((a - d) * 0.5) * ((a - d) * 0.5) + b * c;
    const cplx disc = std::sqrt(dp);
    const cplx l1 = th + disc, l2 = th - disc;
    return std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
}

// Eigenpairs of a Hessenberg matrix: shifted QR with the deflation rule of
// the reference (the largest i with a negligible t[i][i-1] closes the active
// window at i), eigenvectors of the triangular form by back substitution.
void hessenberg_eigen(const Mat& h, std::vector<cplx>& vals, Mat& vecs) {
    const int m = h.rows;
    Mat t = h, q(m, m);
    for (int i = 0; i < m; ++i) q(i, i) = 1.0;
    int p = m;
    std::vector<cplx> gc(m), gs(m);
    for (int it = 0; it < 200 * m; ++it) {
        if (p <= 1) break;
        bool deflated = false;
        for (int i = p - 1; i >= 1; --i) {
            const double sub = std::abs(t(i, i - 1));
            const double ds = std::abs(t(i, i)) + std::abs(t(i - 1, i - 1));
            if (sub < 1e-14 * (ds > 0.0 ? ds : 1.0)) {
                t(i, i - 1) = 0.0;
                p = i;
                deflated = true;
                break;
            }
        }
        if (deflated) continue;
        const cplx sigma = wilkinson(t(p - 2, p - 2), t(p - 2, p - 1), t(p - 1, p - 2), t(p - 1, p - 1));
        for (int i = 0; i < p; ++i) t(i, i) -= sigma;
        for (int i = 0; i + 1 < p; ++i) {
            auto cs = givens(t(i, i), t(i + 1, i));
            gc[i] = cs.first;
            gs[i] = cs.second;
            for (int col = i; col < m; ++col) {
                const cplx a = t(i, col), b = t(i + 1, col);
                t(i, col) = std::conj(gc[i]) * a + std::conj(gs[i]) * b;
                t(i + 1, col) = -gs[i] * a + gc[i] * b;
            }
        }
        for (int i = 0; i + 1 < p; ++i) {
            const cplx c = gc[i], s = gs[i];
            const int rows = std::min(p, i + 3);
            for (int row = 0; row < rows; ++row) {
                const cplx a = t(row, i), b = t(row, i + 1);
                t(row, i) = a * c + b * s;
                t(row, i + 1) = a * (-std::conj(s)) + b * std::conj(c);
            }
            for (int row = 0; row < m; ++row) {
                const cplx a = q(row, i), b = q(row, i + 1);
                q(row, i) = a * c + b * s;
                q(row, i + 1) = a * (-std::conj(s)) + b * std::conj(c);
            }
        }
        for (int i = 0; i < p; ++i) t(i, i) += sigma;
    }
    vals.resize(m);
    for (int i = 0; i < m; ++i) vals[i] = t(i, i);
    Mat sv(m, m);
    for (int j = 0; j < m; ++j) {
        sv(j, j) = 1.0;
        for (int i = j - 1; i >= 0; --i) {
            cplx s(0.0, 0.0);
            for (int l = i + 1; l <= j; ++l) s += t(i, l) * sv(l, j);
            const cplx diag = t(i, i) - t(j, j);
            sv(i, j) = std::abs(diag) > 1e-30 ? -s / diag : -s / (diag + 1e-10);
        }
    }
    vecs = Mat(m, m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            cplx s(0.0, 0.0);
            for (int l = 0; l < m; ++l) s += q(i, l) * sv(l, j);
            vecs(i, j) = s;
        }
}

// A w = e by LU with partial pivoting; false when a pivot is below
// eps * n * max|a| (the reference's singularity test).
bool lu_solve(Mat a, std::vector<cplx>& w) {
    const int n = a.rows;
    double scale = 0.0;
    for (const cplx& z : a.a) scale = std::max(scale, std::abs(z));
    const double tol = std::numeric_limits<double>::epsilon() * std::max(n, 1) *
                       std::max(scale, std::numeric_limits<double>::min());
    std::vector<int> piv(n);
    std::iota(piv.begin(), piv.end(), 0);
    for (int k = 0; k < n; ++k) {
        int best = k;
        double bv = std::abs(a(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::abs(a(i, k)) > bv) {
                bv = std::abs(a(i, k));
                best = i;
            }
        if (bv <= tol) return false;
        if (best != k) {
            for (int j = 0; j < n; ++j) std::swap(a(k, j), a(best, j));
            std::swap(w[k], w[best]);
        }
        for (int i = k + 1; i < n; ++i) {
            const cplx f = a(i, k) / a(k, k);
            for (int j = k + 1; j < n; ++j) a(i, j) -= f * a(k, j);
            w[i] -= f * w[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        cplx s = w[i];
        for (int j = i + 1; j < n; ++j) s -= a(i, j) * w[j];
        w[i] = s / a(i, i);
    }
    return true;
}

// Coefficient vectors (columns, length k_in + steps) of the harmonic Ritz
// pairs with the smallest |theta|, ascending (gcro_dr.rs:922-1000).
Mat harmonic_ritz(const Mat& r, const Mat& bk, const Mat& hrot, const std::vector<cplx>& cs,
                  const std::vector<cplx>& sn, int steps, int k_in, int k_target) {
    if (k_target == 0 || steps == 0) return Mat(k_in + steps, 0);
    const int total = k_in + steps;
    const Mat hraw = unapply_givens(hrot, cs, sn, steps);
    const cplx rho = hraw(steps, steps - 1);
    Mat g(total, total);
    for (int i = 0; i < k_in; ++i) {
        for (int j = 0; j < k_in; ++j) g(i, j) = r(i, j);
        for (int j = 0; j < steps; ++j) g(i, k_in + j) = bk(i, j);
    }
    for (int i = 0; i < steps; ++i)
        for (int j = 0; j < steps; ++j) g(k_in + i, k_in + j) = hraw(i, j);
    const cplx rho_sq = rho * std::conj(rho);
    Mat gh(total, total);
    for (int i = 0; i < total; ++i)
        for (int j = 0; j < total; ++j) gh(i, j) = std::conj(g(j, i));
    std::vector<cplx> w(total, cplx(0.0, 0.0));
    w[total - 1] = 1.0;
    if (!lu_solve(gh, w)) {
        for (int i = 0; i < total; ++i) gh(i, i) += 1e-12;
        std::fill(w.begin(), w.end(), cplx(0.0, 0.0));
        w[total - 1] = 1.0;
        if (!lu_solve(gh, w)) throw Error(SBH_ERR_NUMERICAL, "GCRO-DR: singular harmonic Ritz system");
    }
    Mat aug = g;
    for (int i = 0; i < total; ++i) aug(i, total - 1) += rho_sq * w[i];
    Mat qh;
    to_hessenberg(aug, qh);
    std::vector<cplx> vals;
    Mat vh;
    hessenberg_eigen(aug, vals, vh);
    std::vector<int> order(total);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return std::abs(vals[a]) < std::abs(vals[b]); });
    const int take = std::min(k_target, total);
    Mat y(total, take);
    for (int s = 0; s < take; ++s)
        for (int i = 0; i < total; ++i) {
            cplx acc(0.0, 0.0);
            for (int l = 0; l < total; ++l) acc += qh(i, l) * vh(l, order[s]);
            y(i, s) = acc;
        }
    return y;
}

}  // namespace

sbh_status run_gcro_dr(sbh_operator* op, sbh_precond* M, const double2* d_b, const double2* d_x0,
                       sbh_gcro_subspace* sub, const sbh_gcro_dr_config& cfg, double2* d_x,
                       sbh_gmres_solution* sol) {
    sbh_context* ctx = op->ctx;
    const int64_t n = op->n;
    const int m = (int)std::max<int64_t>(1, std::min<int64_t>(cfg.restart, std::max<int64_t>(n, 1)));
    const int kmax = sub->k;
    *sol = sbh_gmres_solution{};
    Krylov K(ctx, op, M, n, std::max(m, kmax));
    admit_device(ctx, (int64_t)sizeof(double2) * ((int64_t)(m + 1 + 2 * kmax + 3) * n),
                 "the GCRO-DR Krylov basis and recycled vectors", false);
    DeviceArray<double2> V((size_t)(m + 1) * n), w(n), opb(n);
    DeviceArray<double2> C((size_t)std::max(kmax, 1) * n), Unew((size_t)std::max(kmax, 1) * n);
    if (d_x0)
        SBH_HIP(hipMemcpyAsync(d_x, d_x0, sizeof(double2) * n, hipMemcpyDeviceToDevice, ctx->stream));
    else
        SBH_HIP(hipMemsetAsync(d_x, 0, sizeof(double2) * n, ctx->stream));
    auto vec = [&](const DeviceArray<double2>& A, int i) { return A.ptr + (size_t)i * n; };
    int64_t total = 0, restarts = 0, reorth = 0, exact_checks = 0;
    auto finish = [&](double res, double est, bool conv, int status) {
        sol->iterations = total;
        sol->restarts = restarts;
        sol->residual = res;
        sol->estimated_residual = est;
        sol->backward_error = res;
        sol->converged = conv ? 1 : 0;
        sol->status = status;
        sol->spmv_count = K.spmv;
        sol->reorthogonalizations = reorth;
        sol->exact_residual_checks = exact_checks;
    };
    const double tol = cfg.tolerance;
    K.precond(d_b, w.ptr);
    const double b_norm = K.norm(w.ptr);
    const double true_b = K.norm(d_b);
    const double public_b = true_b < 1e-15 ? 1.0 : true_b;
    // the working recycled count; the subspace object changes only on success
    int ucount = sub->count;
    const double2* U = sub->U.ptr;
    bool replaced = false;
    if (b_norm < 1e-15) {
        // M^-1 b = 0: x0 (or zero) is returned as converged and finalised
        finish(0.0, 0.0, true, SBH_SOLVER_CONVERGED);
    } else {
        // C = orth(M^-1 A U), R upper triangular (build_c_r, gcro_dr.rs:770-860)
        Mat R(std::max(ucount, 1), std::max(ucount, 1));
        int k = 0;
        std::vector<cplx> coeffs, re;
        for (int j = 0; j < ucount; ++j) {
            K.apply(U + (size_t)j * n, opb.ptr);
            K.precond(opb.ptr, w.ptr);
            double nz;
            if (k > 0) {
                double unused = 0.0;
                K.dots_project(C.ptr, k, w.ptr, coeffs, &unused);
                for (int i = 0; i < k; ++i) R(i, j) = coeffs[i];
                nz = K.dots_project(C.ptr, k, w.ptr, re, &unused);
                for (int i = 0; i < k; ++i) R(i, j) += re[i];
            } else {
                nz = K.norm(w.ptr);
            }
            if (nz < 1e-12) break;
            R(j, j) = nz;
            scale_copy(ctx, w.ptr, 1.0 / nz, vec(C, k), n);
            k += 1;
        }
        ucount = k;  // a rank-deficient U is truncated (an empty one cleared)

        Mat H(m + 1, m), Bk(std::max(k, 1), m);
        std::vector<cplx> g(m + 1), cs, sn;
        auto exact_rel = [&]() {
            exact_checks += 1;
            K.apply(d_x, opb.ptr);
            return K.residual(d_b, opb.ptr) / public_b;
        };
        auto add_u = [&](const std::vector<cplx>& z) {
            K.upload_coeffs(z, k);
            basis_update(ctx, U, n, k, K.coeff.ptr, d_x, n);
        };
        // new U from the selected Ritz coefficient vectors: lift, then CGS2
        auto refresh = [&](int steps) {
            const Mat Y = harmonic_ritz(R, Bk, H, cs, sn, steps, k, kmax);
            int accepted = 0;
            std::vector<cplx> yu(k), yv(steps), unused_c;
            for (int s = 0; s < Y.cols; ++s) {
                double2* psi = vec(Unew, accepted);
                SBH_HIP(hipMemsetAsync(psi, 0, sizeof(double2) * n, ctx->stream));
                if (k > 0) {
                    for (int i = 0; i < k; ++i) yu[i] = Y(i, s);
                    K.upload_coeffs(yu, k);
                    basis_update(ctx, U, n, k, K.coeff.ptr, psi, n);
                }
                for (int j = 0; j < steps; ++j) yv[j] = Y(k + j, s);
                K.upload_coeffs(yv, steps);
                basis_update(ctx, V.ptr, n, steps, K.coeff.ptr, psi, n);
                double nrm;
                if (accepted > 0) {
                    double unused = 0.0;
                    K.dots_project(Unew.ptr, accepted, psi, unused_c, &unused);
                    nrm = K.dots_project(Unew.ptr, accepted, psi, unused_c, &unused);
                } else {
                    nrm = K.norm(psi);
                }
                if (nrm > 1e-12) {
                    scale_copy(ctx, psi, 1.0 / nrm, psi, n);
                    accepted += 1;
                }
            }
            if (accepted > 0) {
                ucount = accepted;
                replaced = true;
            }
        };

        bool returned = false;
        for (int64_t outer = 0; outer < cfg.max_iterations && !returned; ++outer) {
            K.apply(d_x, opb.ptr);
            K.residual(d_b, opb.ptr);  // opb = b - A x
            K.precond(opb.ptr, w.ptr);
            if (k > 0) {
                double unused = 0.0;
                K.dots_project(C.ptr, k, w.ptr, coeffs, &unused);
                add_u(upper_solve(R, coeffs, k));
            }
            const double beta = K.norm(w.ptr);
            const double rel0 = beta / b_norm;
            if (rel0 < tol) {
                const double t = exact_rel();
                if (t < tol || t < 1e-20) {
                    finish(t, rel0, true, SBH_SOLVER_CONVERGED);
                    returned = true;
                    break;
                }
            }
            scale_copy(ctx, w.ptr, 1.0 / beta, vec(V, 0), n);
            std::fill(H.a.begin(), H.a.end(), cplx(0.0, 0.0));
            std::fill(Bk.a.begin(), Bk.a.end(), cplx(0.0, 0.0));
            cs.clear();
            sn.clear();
            std::fill(g.begin(), g.end(), cplx(0.0, 0.0));
            g[0] = cplx(beta, 0.0);
            bool inner_converged = false;
            int converged_at = -1;
            for (int j = 0; j < m; ++j) {
                total += 1;
                K.apply(vec(V, j), opb.ptr);
                K.precond(opb.ptr, w.ptr);
                if (k > 0) {
                    double unused = 0.0;
                    K.dots_project(C.ptr, k, w.ptr, coeffs, &unused);
                    for (int i = 0; i < k; ++i) Bk(i, j) = coeffs[i];
                    K.dots_project(C.ptr, k, w.ptr, re, &unused);
                    for (int i = 0; i < k; ++i) Bk(i, j) += re[i];
                }
                double before = 0.0;
                double wn = K.dots_project(V.ptr, j + 1, w.ptr, coeffs, &before);
                for (int i = 0; i <= j; ++i) H(i, j) = coeffs[i];
                if (wn < 0.707 * before) {
                    reorth += 1;
                    double unused = 0.0;
                    wn = K.dots_project(V.ptr, j + 1, w.ptr, re, &unused);
                    for (int i = 0; i <= j; ++i) H(i, j) += re[i];
                }
                H(j + 1, j) = cplx(wn, 0.0);
                if (wn < 1e-20)
                    inner_converged = true;
                else
                    scale_copy(ctx, w.ptr, 1.0 / wn, vec(V, j + 1), n);
                for (int i = 0; i < j; ++i) {
                    const cplx temp = cm(std::conj(cs[i]), H(i, j)) + cm(std::conj(sn[i]), H(i + 1, j));
                    H(i + 1, j) = (cplx(0.0, 0.0) - cm(sn[i], H(i, j))) + cm(cs[i], H(i + 1, j));
                    H(i, j) = temp;
                }
                auto rot = givens(H(j, j), H(j + 1, j));
                cs.push_back(rot.first);
                sn.push_back(rot.second);
                const cplx c = rot.first, s = rot.second;
                H(j, j) = cm(std::conj(c), H(j, j)) + cm(std::conj(s), H(j + 1, j));
                H(j + 1, j) = 0.0;
                const cplx temp = cm(std::conj(c), g[j]) + cm(std::conj(s), g[j + 1]);
                g[j + 1] = (cplx(0.0, 0.0) - cm(s, g[j])) + cm(c, g[j + 1]);
                g[j] = temp;
                const double abs_res = cnorm(g[j + 1]);
                if (abs_res / b_norm < tol || abs_res < 1e-20 || inner_converged) {
                    converged_at = j;
                    break;
                }
            }
            const int steps = converged_at < 0 ? m : converged_at + 1;
            const std::vector<cplx> y = upper_solve(H, g, steps);
            K.upload_coeffs(y, steps);
            basis_update(ctx, V.ptr, n, steps, K.coeff.ptr, d_x, n);
            if (k > 0) {
                std::vector<cplx> by(k);
                for (int i = 0; i < k; ++i) {
                    cplx s(0.0, 0.0);
                    for (int jj = 0; jj < steps; ++jj) s += cm(Bk(i, jj), y[jj]);
                    by[i] = cplx(0.0, 0.0) - s;
                }
                add_u(upper_solve(R, by, k));
            }
            const double rel_after = (steps ? cnorm(g[steps]) : 0.0) / b_norm;
            if (converged_at >= 0 || rel_after < tol) {
                const double t = exact_rel();
                if (t < tol || t < 1e-20) {
                    refresh(steps);
                    finish(t, rel_after, true, SBH_SOLVER_CONVERGED);
                    returned = true;
                    break;
                }
                if (inner_converged) {
                    finish(t, rel_after, false, SBH_SOLVER_BREAKDOWN);
                    returned = true;
                    break;
                }
                restarts += 1;
                continue;
            }
            restarts += 1;
        }
        if (!returned) {
            refresh(m);
            const double t = exact_rel();
            finish(t, t, false, SBH_SOLVER_MAX_ITERATIONS);
        }
    }
    // finalize_gmres_solution (gmres.rs:159-250)
    K.apply(d_x, opb.ptr);
    const double ax_norm = K.norm(opb.ptr);
    const double xn = K.norm(d_x);
    const double rn = K.residual(d_b, opb.ptr);
    if (!std::isfinite(xn) || !std::isfinite(ax_norm) || !std::isfinite(rn) ||
        !std::isfinite(sol->estimated_residual))
        throw Error(SBH_ERR_NUMERICAL, "GCRO-DR: operator or preconditioner produced a non-finite result");
    const double true_res = true_b > 0.0 ? rn / true_b : rn;
    sol->residual = true_res;
    sol->backward_error = rn / std::max(ax_norm + true_b, 2.2250738585072014e-308);
    sol->spmv_count = K.spmv;
    const bool conv = true_res <= tol;
    sol->converged = conv ? 1 : 0;
    if (conv)
        sol->status = SBH_SOLVER_CONVERGED;
    else if (sol->status == SBH_SOLVER_CONVERGED)
        sol->status = SBH_SOLVER_MAX_ITERATIONS;
    // commit the subspace: the accepted solve's Ritz vectors, or U truncated
    // to the columns that survived build_c_r
    if (replaced) {
        SBH_HIP(hipMemcpyAsync(sub->U.ptr, Unew.ptr, sizeof(double2) * ucount * n, hipMemcpyDeviceToDevice,
                               ctx->stream));
    }
    SBH_HIP(hipStreamSynchronize(ctx->stream));
    sub->count = ucount;
    return SBH_OK;
}

}  // namespace sbh
