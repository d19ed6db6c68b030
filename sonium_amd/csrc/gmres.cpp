// Restarted GMRES over device-resident vectors.
//
// Follows gmres_core (crates/sonium-solvers/crates/sonium-solvers-iterative/
// src/gmres.rs:631-1057) — convergence checked every convergence_check_interval
// inner steps, adaptive restart when min_restart and max_restart are set — then
// finalize_gmres_solution (gmres.rs:159-250).  Left preconditioning,
// classical Gram–Schmidt with one DGKS reorthogonalisation pass when the
// projected norm drops below 0.707 of the unprojected one (gmres.rs:799-829),
// complex Givens rotations (gmres_common.rs:91-105) and back substitution
// (gmres_common.rs:108-127).
//
// The Krylov basis, the iterate and every vector operation stay on the
// device; only the (j+1) Gram–Schmidt coefficients and norms come back to the
// host each step, where the small Hessenberg recurrence runs in FP64.

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "krylov.hpp"

namespace sbh {
using namespace krylov;

sbh_status run_gmres(sbh_operator* op, sbh_precond* M, const double2* d_b, const double2* d_x0,
                     const sbh_gmres_config& cfg, double2* d_x, sbh_gmres_solution* sol) {
    sbh_context* ctx = op->ctx;
    const int64_t n = op->n;
    *sol = sbh_gmres_solution{};
    auto finish = [&](int64_t it, int64_t rs, double res, double est, double bw, bool conv, int status) {
        sol->iterations = it;
        sol->restarts = rs;
        sol->residual = res;
        sol->estimated_residual = est;
        sol->backward_error = bw;
        sol->converged = conv ? 1 : 0;
        sol->status = status;
    };
    const int64_t nmax = std::max<int64_t>(n, 1);
    const bool adaptive = cfg.min_restart > 0 && cfg.max_restart > 0;
    int m = std::max(1, (int)std::min<int64_t>(cfg.restart, nmax));
    // the basis is allocated for the largest restart the cycle can reach
    const int alloc_m = adaptive ? std::max(m, (int)std::min<int64_t>(cfg.max_restart, nmax)) : m;
    const int min_m = adaptive ? (int)std::min<int64_t>(cfg.min_restart, nmax) : m;
    const int max_m = alloc_m;
    const int64_t check_every = std::max<int64_t>(1, cfg.convergence_check_interval);
    Krylov K(ctx, op, M, n, alloc_m);
    admit_device(ctx, (int64_t)sizeof(double2) * ((int64_t)(alloc_m + 1) * n + 2 * n), "the GMRES Krylov basis", false);
    DeviceArray<double2> V((size_t)(alloc_m + 1) * n), w(n), opb(n);
    // Initial guess.
    if (d_x0)
        SBH_HIP(hipMemcpyAsync(d_x, d_x0, sizeof(double2) * n, hipMemcpyDeviceToDevice, ctx->stream));
    else
        SBH_HIP(hipMemsetAsync(d_x, 0, sizeof(double2) * n, ctx->stream));

    // invalid_gmres_input (gmres.rs:252-275)
    double bnorm_check = K.norm(d_b);
    double xnorm_check = d_x0 ? K.norm(d_x0) : 0.0;
    if (cfg.max_iterations <= 0 || cfg.restart <= 0 || !std::isfinite(cfg.tolerance) ||
        cfg.tolerance <= 0.0 || !std::isfinite(bnorm_check) || !std::isfinite(xnorm_check) ||
        cfg.convergence_check_interval < 0 || cfg.exact_residual_check_interval < 0 ||
        cfg.min_restart < 0 || cfg.max_restart < 0) {
        finish(0, 0, INFINITY, INFINITY, INFINITY, false, SBH_SOLVER_NUMERICAL_ERROR);
        sol->spmv_count = K.spmv;
        return SBH_OK;
    }
    const double tol = cfg.tolerance;
    const double tol_threshold = 1e-15;
    const double raw_b = bnorm_check;
    const bool rhs_zero = raw_b < tol_threshold;
    const double true_b_norm = rhs_zero ? 1.0 : raw_b;
    K.precond(d_b, w.ptr);
    const double raw_pb = (!M || M->kind == 0) ? raw_b : K.norm(w.ptr);
    int64_t total = 0, restarts = 0, reorth = 0, exact_checks = 0;
    auto done = [&]() {
        sol->spmv_count = K.spmv;
        sol->reorthogonalizations = reorth;
        sol->exact_residual_checks = exact_checks;
        return SBH_OK;
    };
    if (!rhs_zero && raw_pb < tol_threshold) {
        finish(0, 0, 1.0, 1.0, 1.0, false, SBH_SOLVER_BREAKDOWN);
        return done();
    }
    const double pb_norm = rhs_zero ? 1.0 : raw_pb;

    std::vector<cplx> H((size_t)(alloc_m + 1) * alloc_m), g(alloc_m + 1), cs, sn, coeffs, re;
    auto h = [&](int i, int j) -> cplx& { return H[(size_t)i * alloc_m + j]; };
    auto vec = [&](int i) { return V.ptr + (size_t)i * n; };

    // exact_relative_residual (gmres_common.rs:16-40)
    auto exact_rel = [&]() {
        exact_checks += 1;
        K.apply(d_x, opb.ptr);
        double rn = K.residual(d_b, opb.ptr);
        return rn / std::max(true_b_norm, 2.2250738585072014e-308);
    };
    auto solve_update = [&](int kk) {
        std::vector<cplx> y(kk, cplx(0.0, 0.0));
        for (int i = kk - 1; i >= 0; --i) {
            cplx sum = g[i];
            for (int j = i + 1; j < kk; ++j) sum -= cm(h(i, j), y[j]);
            if (cnorm(h(i, i)) > 1e-30) y[i] = cm(sum, cinv(h(i, i)));
        }
        K.upload_coeffs(y, kk);
        basis_update(ctx, V.ptr, n, kk, K.coeff.ptr, d_x, n);
    };

    bool returned = false;
    for (int64_t outer = 0; outer < cfg.max_iterations && !returned; ++outer) {
        K.apply(d_x, opb.ptr);
        double rn = K.residual(d_b, opb.ptr);
        double true_rel = rn / true_b_norm;
        double beta;
        if (!M || M->kind == 0) {
            beta = rn;
            SBH_HIP(hipMemcpyAsync(w.ptr, opb.ptr, sizeof(double2) * n, hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            K.precond(opb.ptr, w.ptr);
            beta = K.norm(w.ptr);
        }
        double rel = beta / pb_norm;
        if (rel < tol && true_rel < tol) {
            finish(total, restarts, true_rel, rel, true_rel, true, SBH_SOLVER_CONVERGED);
            returned = true;
            break;
        }
        const double rel_before_cycle = rel;
        if (!std::isfinite(beta) || beta == 0.0) {
            finish(total, restarts, true_rel, rel, true_rel, false, SBH_SOLVER_BREAKDOWN);
            returned = true;
            break;
        }
        scale_copy(ctx, w.ptr, 1.0 / beta, vec(0), n);
        std::fill(H.begin(), H.end(), cplx(0.0, 0.0));
        cs.clear();
        sn.clear();
        std::fill(g.begin(), g.end(), cplx(0.0, 0.0));
        g[0] = cplx(beta, 0.0);
        bool inner_converged = false;
        bool restart_now = false;
        for (int j = 0; j < m; ++j) {
            total += 1;
            if (!M || M->kind == 0) {
                K.apply(vec(j), w.ptr);
            } else {
                K.apply(vec(j), opb.ptr);
                K.precond(opb.ptr, w.ptr);
            }
            double wnb = 0.0;
            double wn = K.dots_project(V.ptr, j + 1, w.ptr, coeffs, &wnb);
            for (int i = 0; i <= j; ++i) h(i, j) = coeffs[i];
            if (wn < 0.707 * wnb) {
                reorth += 1;
                double unused = 0.0;
                wn = K.dots_project(V.ptr, j + 1, w.ptr, re, &unused);
                for (int i = 0; i <= j; ++i) h(i, j) += re[i];
            }
            h(j + 1, j) = cplx(wn, 0.0);
            if (wn < 1e-20) {
                inner_converged = true;
            } else {
                scale_copy(ctx, w.ptr, 1.0 / wn, vec(j + 1), n);
            }
            for (int i = 0; i < j; ++i) {
                cplx temp = cm(std::conj(cs[i]), h(i, j)) + cm(std::conj(sn[i]), h(i + 1, j));
                h(i + 1, j) = (cplx(0.0, 0.0) - cm(sn[i], h(i, j))) + cm(cs[i], h(i + 1, j));
                h(i, j) = temp;
            }
            // givens_rotation (gmres_common.rs:91-105)
            cplx a = h(j, j), b = h(j + 1, j), c, s;
            if (cnorm(b) < 1e-30) {
                c = cplx(1.0, 0.0);
                s = cplx(0.0, 0.0);
            } else if (cnorm(a) < 1e-30) {
                c = cplx(0.0, 0.0);
                s = cplx(1.0, 0.0);
            } else {
                double r = std::sqrt((a.real() * a.real() + a.imag() * a.imag()) +
                                     (b.real() * b.real() + b.imag() * b.imag()));
                c = a * (1.0 / r);
                s = b * (1.0 / r);
            }
            cs.push_back(c);
            sn.push_back(s);
            h(j, j) = cm(std::conj(c), h(j, j)) + cm(std::conj(s), h(j + 1, j));
            h(j + 1, j) = cplx(0.0, 0.0);
            cplx temp = cm(std::conj(c), g[j]) + cm(std::conj(s), g[j + 1]);
            g[j + 1] = (cplx(0.0, 0.0) - cm(s, g[j])) + cm(c, g[j + 1]);
            g[j] = temp;
            double abs_res = cnorm(g[j + 1]);
            double rrel = abs_res / pb_norm;
            bool converged = rrel < tol || (!inner_converged && abs_res < 1e-20);
            const bool should_check = j % check_every == 0 || j == m - 1 || inner_converged;
            if (should_check && converged) {
                solve_update(j + 1);
                double t = exact_rel();
                if (t >= tol && t >= 1e-20) {
                    restarts += 1;
                    restart_now = true;
                    break;
                }
                finish(total, restarts, t, rrel, t, true, SBH_SOLVER_CONVERGED);
                returned = true;
                break;
            }
            if (inner_converged) {
                solve_update(j + 1);
                double t = exact_rel();
                finish(total, restarts, t, rrel, t, t < tol, t < tol ? SBH_SOLVER_CONVERGED : SBH_SOLVER_BREAKDOWN);
                returned = true;
                break;
            }
        }
        if (returned || restart_now) continue;
        solve_update(m);
        double t = exact_rel();
        if (t < tol) {
            finish(total, restarts, t, t, t, true, SBH_SOLVER_CONVERGED);
            returned = true;
            break;
        }
        if (adaptive) {
            // gmres.rs:1013-1029: stalling cycles double the restart, fast ones halve it
            const double ratio = (cnorm(g[m]) / pb_norm) / rel_before_cycle;
            if (ratio > 0.9 && m < max_m)
                m = std::min(m * 2, max_m);
            else if (ratio < 0.1 && m > min_m)
                m = std::max(m / 2, min_m);
        }
        restarts += 1;
    }
    if (!returned) {
        double t = exact_rel();
        finish(total, restarts, t, t, t, false, SBH_SOLVER_MAX_ITERATIONS);
    }
    // finalize_gmres_solution (gmres.rs:159-250)
    if (sol->status == SBH_SOLVER_NUMERICAL_ERROR) return done();
    if (!std::isfinite(sol->estimated_residual) || sol->estimated_residual < 0.0) {
        finish(sol->iterations, sol->restarts, INFINITY, sol->estimated_residual, INFINITY, false,
               SBH_SOLVER_NUMERICAL_ERROR);
        return done();
    }
    K.apply(d_x, opb.ptr);
    double ax_norm = K.norm(opb.ptr);
    double xn = K.norm(d_x);
    double rn = K.residual(d_b, opb.ptr);
    if (!std::isfinite(xn) || !std::isfinite(ax_norm) || !std::isfinite(rn)) {
        finish(sol->iterations, sol->restarts, INFINITY, sol->estimated_residual, INFINITY, false,
               SBH_SOLVER_NUMERICAL_ERROR);
        return done();
    }
    double true_res = raw_b > 0.0 ? rn / raw_b : rn;
    double denom = std::max(ax_norm + raw_b, 2.2250738585072014e-308);
    sol->residual = true_res;
    sol->backward_error = rn / denom;
    bool conv = true_res <= tol;
    sol->converged = conv ? 1 : 0;
    if (conv)
        sol->status = SBH_SOLVER_CONVERGED;
    else if (sol->status == SBH_SOLVER_CONVERGED)
        sol->status = SBH_SOLVER_MAX_ITERATIONS;
    return done();
}

}  // namespace sbh
