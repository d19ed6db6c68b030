// Device-resident Krylov helpers shared by the GMRES and GCRO-DR drivers:
// operator and preconditioner application, norms and residuals summed on the
// host in chunk order, and one-read classical Gram-Schmidt passes.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "internal.hpp"

namespace sbh {
namespace krylov {

using cplx = std::complex<double>;

// num-complex arithmetic in the reference's order.
inline cplx cm(cplx a, cplx b) {
    return cplx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cplx cinv(cplx a) {
    double n = a.real() * a.real() + a.imag() * a.imag();
    return cplx(a.real() / n, -a.imag() / n);
}
inline double cnorm(cplx a) { return std::hypot(a.real(), a.imag()); }

struct Krylov {
    sbh_context* ctx;
    sbh_operator* op;
    sbh_precond* M;
    int64_t n;
    DeviceArray<double2> part;   // reduction partials
    DeviceArray<double2> coeff;  // (m+1) coefficients
    double* host = nullptr;      // pinned
    int64_t spmv = 0;

    Krylov(sbh_context* c, sbh_operator* o, sbh_precond* p, int64_t nn, int m)
        : ctx(c), op(o), M(p), n(nn) {
        int chunks = reduction_chunks(n);
        size_t need = std::max<size_t>((size_t)(m + 2) * chunks, (size_t)((n + 255) / 256)) + 8;
        part.allocate(need);
        coeff.allocate((size_t)m + 2);
        host = ctx->pinned_rw(need * 2);
    }

    void apply(const double2* x, double2* y) {
        operator_apply_device(op, x, y);
        spmv += 1;
    }
    void precond(const double2* r, double2* y) {
        if (!M || M->kind == 0) {
            if (r != y) SBH_HIP(hipMemcpyAsync(y, r, sizeof(double2) * n, hipMemcpyDeviceToDevice, ctx->stream));
        } else if (M->kind == 1) {
            diag_scale(ctx, M->d_dinv.ptr, r, y, n);
        } else {
            callback_apply_device(ctx, M->callback, M->user, M->callback_memory, r, y, n, "preconditioner");
        }
    }
    double fetch_sum(int count) {
        SBH_HIP(hipMemcpyAsync(host, part.ptr, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        upload_pending = false;
        double s = 0.0;
        for (int i = 0; i < count; ++i) s += host[i];
        return s;
    }
    double norm(const double2* x) {
        int nb = norm_partials(ctx, x, n, reinterpret_cast<double*>(part.ptr));
        return std::sqrt(fetch_sum(nb));
    }
    // r = b - r, returns ||r||
    double residual(const double2* b, double2* r) {
        int nb = residual_and_norm(ctx, b, r, n, reinterpret_cast<double*>(part.ptr));
        return std::sqrt(fetch_sum(nb));
    }
    // The pinned host buffer is written here and read by an asynchronous
    // copy: a second upload first waits for the previous one to complete.
    bool upload_pending = false;
    void upload_coeffs(const std::vector<cplx>& c, int count) {
        if (upload_pending) SBH_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < count; ++i) {
            host[2 * i] = c[i].real();
            host[2 * i + 1] = c[i].imag();
        }
        SBH_HIP(hipMemcpyAsync(coeff.ptr, host, sizeof(double2) * count, hipMemcpyHostToDevice, ctx->stream));
        upload_pending = true;
    }
    // One classical Gram-Schmidt pass with one host read: the coefficients
    // <v_i, w> are summed on the device in chunk order (the order the host
    // used), w -= sum c_i v_i runs with them, and coefficients, ||w|| before
    // and the norm partials after come back together.  Returns ||w|| after.
    double dots_project(const double2* V, int nvec, double2* w, std::vector<cplx>& out, double* before) {
        const int chunks = dots_and_norm_partials(ctx, V, n, nvec, w, n, part.ptr);
        sum_chunk_partials(ctx, part.ptr, nvec + 1, chunks, coeff.ptr);
        const int nb = project_and_norm(ctx, V, n, nvec, coeff.ptr, w, n, reinterpret_cast<double*>(part.ptr));
        SBH_HIP(hipMemcpyAsync(host, coeff.ptr, sizeof(double2) * (nvec + 1), hipMemcpyDeviceToHost, ctx->stream));
        SBH_HIP(hipMemcpyAsync(host + 2 * (nvec + 1), part.ptr, sizeof(double) * nb, hipMemcpyDeviceToHost,
                               ctx->stream));
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        upload_pending = false;
        out.assign(nvec, cplx(0.0, 0.0));
        for (int v = 0; v < nvec; ++v) out[v] = cplx(host[2 * v], host[2 * v + 1]);
        *before = std::sqrt(host[2 * nvec]);
        double s = 0.0;
        for (int i = 0; i < nb; ++i) s += host[2 * (nvec + 1) + i];
        return std::sqrt(s);
    }
};

}  // namespace krylov
}  // namespace sbh
