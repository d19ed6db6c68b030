// C ABI of libsbh (declarations and reference citations in include/sbh.h).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "internal.hpp"

namespace sbh {
sbh_status run_gmres_block(sbh_operator* op, const double2* d_B, double2* d_X, int k, const sbh_gmres_config& cfg,
                           bool x_zero, sbh_gmres_solution* sols);
int bk_frob_partials(sbh_context* ctx, const double2* X, int64_t count, double* part);
sbh_status run_block_gmres(sbh_operator* op, const sbh_precond* M, const double2* d_B, double2* d_X, int k,
                           bool x_zero, const sbh_block_gmres_config& cfg, sbh_block_gmres_info* info,
                           double* residuals, double* true_residuals, double* backward_errors);
sbh_status run_gmres_ir(sbh_operator* op, sbh_lowp_matrix* lp, const double2* d_B, double2* d_X, int k,
                        const sbh_gmres_ir_config& cfg, sbh_gmres_ir_status* st);
void zgemm_f64(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X, int64_t k, double2* Y);
int64_t lowp_padded_rows(int64_t n);
void lowp_gemm(sbh_context* ctx, const uint16_t* A, int64_t ld, int64_t n, int format, const double* row_inv,
               const double2* X, int k, uint16_t* XT, unsigned* colmax, double* col_inv, float* Ypart, double2* Y);
sbh_status run_gmres(sbh_operator* op, sbh_precond* M, const double2* d_b, const double2* d_x0,
                     const sbh_gmres_config& cfg, double2* d_x, sbh_gmres_solution* sol);
sbh_status run_gcro_dr(sbh_operator* op, sbh_precond* M, const double2* d_b, const double2* d_x0,
                       sbh_gcro_subspace* sub, const sbh_gcro_dr_config& cfg, double2* d_x,
                       sbh_gmres_solution* sol);
void launch_sincos_probe(sbh_context* ctx, const double* x, int64_t n, double* s, double* c);
void launch_rsqrt_probe(sbh_context* ctx, const double* x, int64_t n, double* y);

namespace {
thread_local std::string g_last_error;

inline void use_device(sbh_context* ctx) { SBH_HIP(hipSetDevice(ctx->device)); }

void require_dense(const sbh_operator* op, const char* what) {
    if (!op->dense()) throw Error(SBH_ERR_UNSUPPORTED, std::string(what) + " needs a dense (stored) operator");
}

// Device copy of a host or device input vector.
struct DeviceInput {
    DeviceArray<double2> owned;
    const double2* ptr = nullptr;
    DeviceInput(sbh_context* ctx, const sbh_c64* p, size_t n, int memory) {
        if (memory == SBH_MEM_DEVICE) {
            ptr = reinterpret_cast<const double2*>(p);
        } else {
            owned.allocate(n);
            injection_check(INJECT_TRANSFER_H2D);
            if (n) SBH_HIP(hipMemcpyAsync(owned.ptr, p, n * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
            ptr = owned.ptr;
        }
    }
};

struct DeviceOutput {
    DeviceArray<double2> owned;
    double2* ptr = nullptr;
    sbh_c64* host = nullptr;
    size_t n = 0;
    sbh_context* ctx;
    DeviceOutput(sbh_context* c, sbh_c64* p, size_t nn, int memory) : n(nn), ctx(c) {
        if (memory == SBH_MEM_DEVICE) {
            ptr = reinterpret_cast<double2*>(p);
        } else {
            owned.allocate(n);
            ptr = owned.ptr;
            host = p;
        }
    }
    void finish() {
        if (host && n) {
            injection_check(INJECT_TRANSFER_D2H);
            SBH_HIP(hipMemcpyAsync(host, ptr, n * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        }
        injection_check(INJECT_SYNCHRONIZATION);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
    }
};

void check_memory(int memory) {
    SBH_REQUIRE(memory == SBH_MEM_HOST || memory == SBH_MEM_DEVICE, "memory must be SBH_MEM_HOST or SBH_MEM_DEVICE");
}

bool all_finite(const sbh_c64* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i].re) || !std::isfinite(p[i].im)) return false;
    return true;
}

}  // namespace

// y = A x on device vectors, for either kind of operator.
void callback_apply_device(sbh_context* ctx, sbh_apply_fn fn, void* user, int memory, const double2* x, double2* y,
                           int64_t n, const char* what) {
    SBH_HIP(hipStreamSynchronize(ctx->stream));
    int rc = 0;
    if (memory == SBH_MEM_DEVICE) {
        rc = fn(user, reinterpret_cast<const sbh_c64*>(x), reinterpret_cast<sbh_c64*>(y), n, SBH_MEM_DEVICE);
    } else {
        std::vector<sbh_c64> hx((size_t)n), hy((size_t)n);
        SBH_HIP(hipMemcpyAsync(hx.data(), x, sizeof(double2) * n, hipMemcpyDeviceToHost, ctx->stream));
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        rc = fn(user, hx.data(), hy.data(), n, SBH_MEM_HOST);
        if (rc == 0) {
            SBH_HIP(hipMemcpyAsync(y, hy.data(), sizeof(double2) * n, hipMemcpyHostToDevice, ctx->stream));
            SBH_HIP(hipStreamSynchronize(ctx->stream));
        }
    }
    if (rc != 0) throw Error(SBH_ERR_BACKEND, std::string(what) + " callback failed with status " + std::to_string(rc));
}

void operator_apply_device(const sbh_operator* op, const double2* x, double2* y) {
    injection_check(INJECT_LAUNCH);
    if (op->callback)
        callback_apply_device(op->ctx, op->callback, op->user, op->callback_memory, x, y, op->n, "operator");
    else if (op->exact())
        launch_exact_apply(op, x, y);
    else
        zgemv_n(op->ctx, op->sys->d_A.ptr, op->sys->lda, op->n, op->n, x, y);
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

thread_local sbh_context* g_current_ctx = nullptr;
std::atomic<int64_t> g_live_device_bytes{0};

void injection_check(int point) {
    sbh_context* ctx = g_current_ctx;
    if (!ctx) return;
    int expected = point;
    if (!ctx->armed_failure.compare_exchange_strong(expected, INJECT_NONE)) return;
    switch (point) {
        case INJECT_ALLOCATION:
            throw Error(SBH_ERR_OUT_OF_MEMORY, "injected allocation failure");
        case INJECT_LAUNCH:
            throw Error(SBH_ERR_BACKEND, "injected launch failure");
        case INJECT_TRANSFER_H2D:
            throw Error(SBH_ERR_BACKEND, "injected host-to-device transfer failure");
        case INJECT_TRANSFER_D2H:
            throw Error(SBH_ERR_BACKEND, "injected device-to-host transfer failure");
        default:
            throw Error(SBH_ERR_BACKEND, "injected synchronization failure");
    }
}

void admit_device(sbh_context* ctx, int64_t bytes, const char* what, bool reserve) {
    if (bytes <= 0) return;
    const int64_t budget = ctx->device_budget.load();
    const int64_t held = ctx->device_reserved.load();
    if (budget > 0 && held + bytes > budget)
        throw Error(SBH_ERR_OUT_OF_MEMORY, std::string("admission refused: ") + what + " needs " +
                                               std::to_string(bytes) + " device bytes; the context's budget is " +
                                               std::to_string(budget) + " with " + std::to_string(held) + " held");
    size_t free_b = 0, total_b = 0;
    SBH_HIP(hipMemGetInfo(&free_b, &total_b));
    if ((uint64_t)bytes > (uint64_t)free_b)
        throw Error(SBH_ERR_OUT_OF_MEMORY, std::string("admission refused: ") + what + " needs " +
                                               std::to_string(bytes) + " device bytes; the device has " +
                                               std::to_string(free_b) + " free of " + std::to_string(total_b));
    if (reserve) ctx->device_reserved += bytes;
}

void release_device(sbh_context* ctx, int64_t bytes) {
    if (bytes > 0) ctx->device_reserved -= bytes;
}

void context_retain(sbh_context* ctx) { ctx->refs.fetch_add(1); }

void context_release(sbh_context* ctx) {
    if (ctx->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& t : ctx->timing)
        for (auto& e : t.events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // namespace sbh

using namespace sbh;

double* sbh_context::pinned_rw(size_t count) {
    if (pinned_count < count) {
        if (pinned) (void)hipHostFree(pinned);
        pinned = nullptr;
        pinned_count = 0;
        SBH_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned), count * sizeof(double), hipHostMallocDefault));
        pinned_count = count;
    }
    return pinned;
}

double2* sbh_context::device_scratch(size_t count) {
    if (scratch_count < count) {
        SBH_HIP(hipStreamSynchronize(stream));
        if (scratch) (void)hipFree(scratch);
        scratch = nullptr;
        scratch_count = 0;
        SBH_HIP(hipMalloc(reinterpret_cast<void**>(&scratch), count * sizeof(double2)));
        scratch_count = count;
    }
    return scratch;
}

hipEvent_t sbh_context::take_event() {
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    SBH_HIP(hipEventCreate(&e));
    return e;
}

void sbh_context::begin_timed(int cls, hipEvent_t* start) {
    *start = nullptr;
    if (!profiling) return;
    *start = take_event();
    SBH_HIP(hipEventRecord(*start, stream));
}

void sbh_context::end_timed(int cls, hipEvent_t start) {
    if (!profiling || !start) return;
    hipEvent_t stop = take_event();
    SBH_HIP(hipEventRecord(stop, stream));
    std::lock_guard<std::mutex> lk(mu);
    timing[cls].events.emplace_back(start, stop);
    timing[cls].launches += 1;
}

extern "C" {

const char* sbh_last_error(void) { return g_last_error.c_str(); }

const char* sbh_version(void) { return "sonium_amd-sbh 0.1 (gfx950)"; }

sbh_status sbh_physics_new(double frequency, double speed_of_sound, double density, int is_internal,
                           sbh_physics* out) {
    return guard([&] {
        SBH_REQUIRE(out, "out is null");
        SBH_REQUIRE(std::isfinite(frequency) && frequency > 0.0, "frequency must be positive and finite");
        SBH_REQUIRE(std::isfinite(speed_of_sound) && speed_of_sound > 0.0, "speed of sound must be positive");
        SBH_REQUIRE(std::isfinite(density) && density > 0.0, "density must be positive");
        const double PI = 3.14159265358979323846;
        sbh_physics p;
        p.frequency = frequency;
        p.speed_of_sound = speed_of_sound;
        p.density = density;
        p.omega = 2.0 * PI * frequency;
        p.wave_number = p.omega / speed_of_sound;
        p.wave_length = speed_of_sound / frequency;
        p.harmonic_factor = 1.0;
        p.pressure_factor = density * p.omega * p.harmonic_factor;
        p.tau = is_internal ? -1.0 : 1.0;
        *out = p;
    });
}

sbh_status sbh_physics_from_wavenumber(double wave_number, double speed_of_sound, double density,
                                       int is_internal, sbh_physics* out) {
    return guard([&] {
        SBH_REQUIRE(out, "out is null");
        SBH_REQUIRE(std::isfinite(wave_number) && wave_number > 0.0, "wave number must be positive and finite");
        SBH_REQUIRE(std::isfinite(speed_of_sound) && speed_of_sound > 0.0, "speed of sound must be positive");
        SBH_REQUIRE(std::isfinite(density) && density > 0.0, "density must be positive");
        const double PI = 3.14159265358979323846;
        sbh_physics p;
        p.speed_of_sound = speed_of_sound;
        p.density = density;
        p.wave_number = wave_number;
        p.omega = wave_number * speed_of_sound;
        p.frequency = p.omega / (2.0 * PI);
        p.wave_length = 2.0 * PI / wave_number;
        p.harmonic_factor = 1.0;
        p.pressure_factor = density * p.omega * p.harmonic_factor;
        p.tau = is_internal ? -1.0 : 1.0;
        *out = p;
    });
}

sbh_c64 sbh_burton_miller_beta_scaled(const sbh_physics* ph, double scale) {
    if (!ph || !(ph->tau > 0.0)) return sbh_c64{0.0, 0.0};
    return sbh_c64{0.0, ph->harmonic_factor * scale / ph->wave_number};
}

sbh_status sbh_context_create(int device, sbh_context** out) {
    return guard([&] {
        SBH_REQUIRE(out, "out is null");
        *out = nullptr;
        int count = 0;
        SBH_HIP(hipGetDeviceCount(&count));
        SBH_REQUIRE(device >= 0 && device < count, "device index out of range");
        std::unique_ptr<sbh_context> ctx(new sbh_context());
        ctx->device = device;
        SBH_HIP(hipSetDevice(device));
        SBH_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->pinned_rw(1 << 16);
        upload_quadrature_tables(ctx.get());
        noalloc_scratch_init(device);
        *out = ctx.release();
    });
}

void sbh_context_destroy(sbh_context* ctx) {
    if (ctx) context_release(ctx);
}

sbh_status sbh_context_synchronize(sbh_context* ctx) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx, "ctx is null");
        use_device(ctx);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

namespace {
bool mul_ok(int64_t a, int64_t b, int64_t* r) { return !__builtin_mul_overflow(a, b, r); }
bool add_ok(int64_t a, int64_t b, int64_t* r) { return !__builtin_add_overflow(a, b, r); }
}  // namespace

sbh_status sbh_admission_estimate(int64_t n, int64_t rhs, int64_t freq, int64_t restart,
                                  sbh_resource_estimate* out) {
    return guard([&] {
        SBH_REQUIRE(out, "out is null");
        SBH_REQUIRE(n > 0, "invalid request: active_dofs must be greater than zero");
        SBH_REQUIRE(rhs > 0, "invalid request: rhs_count must be greater than zero");
        SBH_REQUIRE(freq > 0, "invalid request: frequency_count must be greater than zero");
        SBH_REQUIRE(restart > 0, "invalid request: restart must be greater than zero");
        const int64_t C = 16;  // bytes of a complex double
        auto prod = [](std::initializer_list<int64_t> v, const char* what) {
            int64_t r = 1;
            for (int64_t x : v)
                if (!mul_ok(r, x, &r)) throw Error(SBH_ERR_INVALID_ARGUMENT, std::string("arithmetic overflow: ") + what);
            return r;
        };
        auto sum = [](std::initializer_list<int64_t> v, const char* what) {
            int64_t r = 0;
            for (int64_t x : v)
                if (!add_ok(r, x, &r)) throw Error(SBH_ERR_INVALID_ARGUMENT, std::string("arithmetic overflow: ") + what);
            return r;
        };
        sbh_resource_estimate e{};
        e.operator_bytes = prod({n, n, C}, "dense operator size");
        const int64_t rhs_bytes = prod({n, rhs, C}, "right-hand-side size");
        e.active_work = prod({freq, rhs}, "frequency/RHS work count");
        e.scratch_bytes = prod({n, 4 * C + 8}, "assembly scratch size");
        e.rhs_bytes = prod({rhs_bytes, freq}, "frequency right-hand-side size");
        e.host_bytes = sum({e.operator_bytes, e.rhs_bytes, e.scratch_bytes}, "peak host allocation");
        // this library: system (matrix, one right-hand side, row maxima of the
        // assembly), right-hand sides and solutions, Krylov basis (shared
        // block basis for rhs > 1)
        const int64_t basis_vectors = rhs == 1 ? sum({restart, 3}, "basis size")
                                               : prod({sum({restart, 2}, "basis size"), rhs}, "basis size");
        e.device_bytes = sum({e.operator_bytes, prod({n, C}, "rhs"), prod({n, (n + 255) / 256 + 1, 8}, "row maxima"),
                              prod({2, rhs_bytes}, "blocks"),
                              prod({basis_vectors, n, C}, "Krylov basis"),
                              prod({6, rhs_bytes}, "block scratch")},
                             "peak device allocation");
        *out = e;
    });
}

sbh_status sbh_admission_admit(const sbh_resource_budget* budget, int64_t n, int64_t rhs, int64_t freq,
                               int64_t restart, sbh_resource_estimate* out) {
    return guard([&] {
        SBH_REQUIRE(budget && out, "null argument");
        SBH_REQUIRE(budget->max_host_bytes > 0, "invalid budget: max_host_bytes must be greater than zero");
        SBH_REQUIRE(budget->max_active_work > 0, "invalid budget: max_active_work must be greater than zero");
        sbh_resource_estimate e{};
        const sbh_status st = sbh_admission_estimate(n, rhs, freq, restart, &e);
        if (st != SBH_OK) throw Error(st, sbh_last_error());
        auto exceeded = [](const char* res, int64_t req, int64_t lim) {
            throw Error(SBH_ERR_OUT_OF_MEMORY, std::string("limit exceeded: ") + res + " requested " +
                                                   std::to_string(req) + ", limit " + std::to_string(lim));
        };
        if (e.host_bytes > budget->max_host_bytes) exceeded("host memory", e.host_bytes, budget->max_host_bytes);
        if (e.device_bytes > budget->max_device_bytes)
            exceeded("device memory", e.device_bytes, budget->max_device_bytes);
        if (e.active_work > budget->max_active_work) exceeded("active work", e.active_work, budget->max_active_work);
        *out = e;
    });
}

sbh_status sbh_context_set_device_budget(sbh_context* ctx, int64_t bytes) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx, "ctx is null");
        SBH_REQUIRE(bytes >= 0, "the device budget must be >= 0");
        ctx->device_budget = bytes;
    });
}

sbh_status sbh_context_device_budget(sbh_context* ctx, int64_t* budget, int64_t* held) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx, "ctx is null");
        if (budget) *budget = ctx->device_budget.load();
        if (held) *held = ctx->device_reserved.load();
    });
}

sbh_status sbh_context_inject_failure(sbh_context* ctx, int point) {
    return guard([&] {
        SBH_REQUIRE(ctx, "ctx is null");
        SBH_REQUIRE(point >= SBH_INJECT_NONE && point <= SBH_INJECT_SYNCHRONIZATION, "unknown injection point");
        ctx->armed_failure = point;
    });
}

int64_t sbh_live_device_bytes(void) { return g_live_device_bytes.load(); }

sbh_status sbh_context_set_profiling(sbh_context* ctx, int enabled) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx, "ctx is null");
        ctx->profiling = enabled != 0;
        if (ctx->profiling) {
            // warm the event pool so that timed regions create no events
            use_device(ctx);
            std::lock_guard<std::mutex> lk(ctx->mu);
            while (ctx->event_pool.size() < 4096) {
                hipEvent_t e = nullptr;
                SBH_HIP(hipEventCreate(&e));
                ctx->event_pool.push_back(e);
            }
        }
    });
}

sbh_status sbh_context_kernel_stats(sbh_context* ctx, int cls, int64_t* launches, double* total_ms) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx, "ctx is null");
        SBH_REQUIRE(cls >= 0 && cls < SBH_KERNEL_CLASS_COUNT, "kernel class out of range");
        use_device(ctx);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        std::lock_guard<std::mutex> lk(ctx->mu);
        KernelTiming& t = ctx->timing[cls];
        for (auto& e : t.events) {
            float ms = 0.0f;
            SBH_HIP(hipEventElapsedTime(&ms, e.first, e.second));
            t.total_ms += ms;
            ctx->event_pool.push_back(e.first);
            ctx->event_pool.push_back(e.second);
        }
        t.events.clear();
        if (launches) *launches = t.launches;
        if (total_ms) *total_ms = t.total_ms;
    });
}

sbh_status sbh_context_reset_stats(sbh_context* ctx) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx, "ctx is null");
        use_device(ctx);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (auto& t : ctx->timing) {
            for (auto& e : t.events) {
                ctx->event_pool.push_back(e.first);
                ctx->event_pool.push_back(e.second);
            }
            t.events.clear();
            t.launches = 0;
            t.total_ms = 0.0;
        }
    });
}

sbh_status sbh_vector_create(sbh_context* ctx, int64_t length, sbh_vector** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && out && length >= 1, "invalid vector arguments");
        use_device(ctx);
        std::unique_ptr<sbh_vector> v(new sbh_vector());
        v->ctx = ctx;
        v->length = length;
        v->d.allocate((size_t)length);
        SBH_HIP(hipMemsetAsync(v->d.ptr, 0, v->d.bytes(), ctx->stream));
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        context_retain(ctx);
        *out = v.release();
    });
}

void sbh_vector_destroy(sbh_vector* v) {
    if (!v) return;
    sbh_context* ctx = v->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete v;
    context_release(ctx);
}

int64_t sbh_vector_length(const sbh_vector* v) { return v ? v->length : 0; }

sbh_c64* sbh_vector_data(sbh_vector* v) { return v ? reinterpret_cast<sbh_c64*>(v->d.ptr) : nullptr; }

sbh_status sbh_vector_upload(sbh_vector* v, const sbh_c64* host, int64_t offset, int64_t count) {
    return guard_on(ctx_of(v), [&] {
        SBH_REQUIRE(v && host, "vector/host is null");
        SBH_REQUIRE(offset >= 0 && count >= 0 && offset + count <= v->length, "range out of bounds");
        use_device(v->ctx);
        SBH_HIP(hipMemcpyAsync(v->d.ptr + offset, host, sizeof(double2) * count, hipMemcpyHostToDevice,
                               v->ctx->stream));
        SBH_HIP(hipStreamSynchronize(v->ctx->stream));
    });
}

sbh_status sbh_vector_download(const sbh_vector* v, sbh_c64* host, int64_t offset, int64_t count) {
    return guard_on(ctx_of(v), [&] {
        SBH_REQUIRE(v && host, "vector/host is null");
        SBH_REQUIRE(offset >= 0 && count >= 0 && offset + count <= v->length, "range out of bounds");
        use_device(v->ctx);
        SBH_HIP(hipMemcpyAsync(host, v->d.ptr + offset, sizeof(double2) * count, hipMemcpyDeviceToHost,
                               v->ctx->stream));
        SBH_HIP(hipStreamSynchronize(v->ctx->stream));
    });
}

sbh_status sbh_mesh_create(sbh_context* ctx, const double* nodes, int64_t num_nodes,
                           const int64_t* triangles, int64_t num_elements, const int32_t* bc_type,
                           const sbh_c64* bc_value, const sbh_c64* admittance,
                           const uint8_t* has_admittance, sbh_mesh** out) {
    return sbh_mesh_create_oriented(ctx, nodes, num_nodes, triangles, num_elements, bc_type, bc_value, admittance,
                                    has_admittance, SBH_NORMALS_OUTWARD, out);
}

sbh_status sbh_mesh_create_oriented(sbh_context* ctx, const double* nodes, int64_t num_nodes,
                                    const int64_t* triangles, int64_t num_elements, const int32_t* bc_type,
                                    const sbh_c64* bc_value, const sbh_c64* admittance,
                                    const uint8_t* has_admittance, int normal_rule, sbh_mesh** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && out, "ctx/out is null");
        SBH_REQUIRE(normal_rule == SBH_NORMALS_OUTWARD || normal_rule == SBH_NORMALS_VERTEX_ORDER,
                    "unknown normal rule");
        *out = nullptr;
        SBH_REQUIRE(num_nodes >= 3 && nodes, "mesh needs at least 3 nodes");
        SBH_REQUIRE(num_elements >= 1 && triangles, "mesh needs at least one element");
        SBH_REQUIRE(num_elements < (int64_t)1 << 31, "too many elements");
        for (int64_t i = 0; i < 3 * num_nodes; ++i)
            SBH_REQUIRE(std::isfinite(nodes[i]), "node coordinates must be finite");
        for (int64_t i = 0; i < 3 * num_elements; ++i)
            SBH_REQUIRE(triangles[i] >= 0 && triangles[i] < num_nodes, "triangle node index out of range");
        use_device(ctx);
        std::unique_ptr<sbh_mesh> m(new sbh_mesh());
        m->ctx = ctx;
        m->n = num_elements;
        m->num_nodes = num_nodes;
        m->nodes.assign(nodes, nodes + 3 * num_nodes);
        m->tris.assign(triangles, triangles + 3 * num_elements);
        m->vertex_order_normals = normal_rule == SBH_NORMALS_VERTEX_ORDER;
        build_mesh_geometry(m.get());
        m->bc.assign(num_elements, PanelBc{});
        bool general = false;
        for (int64_t e = 0; e < num_elements; ++e) {
            PanelBc& b = m->bc[e];
            b.type = bc_type ? bc_type[e] : 0;
            SBH_REQUIRE(b.type == 0 || b.type == 1, "bc_type must be 0 (velocity) or 1 (pressure)");
            b.value[0] = bc_value ? bc_value[e].re : 0.0;
            b.value[1] = bc_value ? bc_value[e].im : 0.0;
            SBH_REQUIRE(std::isfinite(b.value[0]) && std::isfinite(b.value[1]), "BC values must be finite");
            b.has_adm = (admittance && has_admittance && has_admittance[e]) ? 1 : 0;
            if (b.has_adm) {
                b.adm[0] = admittance[e].re;
                b.adm[1] = admittance[e].im;
                SBH_REQUIRE(std::isfinite(b.adm[0]) && std::isfinite(b.adm[1]), "admittance must be finite");
            }
            // has_nonzero_bc (tbem.rs:1240-1242)
            if (std::hypot(b.value[0], b.value[1]) > 1e-15) m->nonzero_bc.push_back((int32_t)e);
            if (b.type != 0 || b.has_adm) general = true;
        }
        if (!m->nonzero_bc.empty()) general = true;
        m->general = general;
        const double FOUR_PI = 12.566370614359172;
        std::vector<FarPanel> far(num_elements);
        std::vector<Source> src(num_elements);
        for (int64_t e = 0; e < num_elements; ++e) {
            const PanelGeom& g = m->geom[e];
            FarPanel& f = far[e];
            for (int d = 0; d < 3; ++d) {
                f.v0[d] = g.c[d];
                f.e1[d] = g.c[3 + d] - g.c[d];
                f.e2[d] = g.c[6 + d] - g.c[d];
                f.nf[d] = g.nf[d];
                src[e].c[d] = m->centers[3 * e + d];
                src[e].n[d] = m->normals[3 * e + d];
            }
            f.wj = 0.5 * g.jac / FOUR_PI;
            f.pad = 0.0;
        }
        m->d_far.allocate(num_elements);
        m->d_geom.allocate(num_elements);
        m->d_src.allocate(num_elements);
        m->d_bc.allocate(num_elements);
        copy_on_stream(ctx, m->d_far.ptr, far.data(), m->d_far.bytes(), hipMemcpyHostToDevice);
        copy_on_stream(ctx, m->d_geom.ptr, m->geom.data(), m->d_geom.bytes(), hipMemcpyHostToDevice);
        copy_on_stream(ctx, m->d_src.ptr, src.data(), m->d_src.bytes(), hipMemcpyHostToDevice);
        copy_on_stream(ctx, m->d_bc.ptr, m->bc.data(), m->d_bc.bytes(), hipMemcpyHostToDevice);
        if (!m->nonzero_bc.empty()) {
            m->d_nonzero_bc.allocate(m->nonzero_bc.size());
            copy_on_stream(ctx, m->d_nonzero_bc.ptr, m->nonzero_bc.data(), m->d_nonzero_bc.bytes(),
                           hipMemcpyHostToDevice);
        }
        context_retain(ctx);
        *out = m.release();
    });
}

void sbh_mesh_destroy(sbh_mesh* mesh) {
    if (!mesh) return;
    sbh_context* ctx = mesh->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete mesh;
    context_release(ctx);
}

int64_t sbh_mesh_num_elements(const sbh_mesh* mesh) { return mesh ? mesh->n : 0; }

sbh_status sbh_mesh_geometry(const sbh_mesh* m, double* centers, double* normals, double* areas) {
    return guard_on(ctx_of(m), [&] {
        SBH_REQUIRE(m, "mesh is null");
        if (centers) std::memcpy(centers, m->centers.data(), sizeof(double) * 3 * m->n);
        if (normals) std::memcpy(normals, m->normals.data(), sizeof(double) * 3 * m->n);
        if (areas) std::memcpy(areas, m->areas.data(), sizeof(double) * m->n);
    });
}

sbh_status sbh_workspace_create(sbh_context* ctx, const sbh_mesh* mesh, int route_threads,
                                sbh_workspace** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && mesh && out, "ctx/mesh/out is null");
        SBH_REQUIRE(mesh->ctx == ctx, "mesh belongs to another context");
        *out = nullptr;
        use_device(ctx);
        auto t0 = std::chrono::steady_clock::now();
        RouteTable rt = build_routes(mesh, route_threads);
        std::unique_ptr<sbh_workspace> ws(new sbh_workspace());
        ws->ctx = ctx;
        ws->mesh = mesh;
        ws->near_pairs = (int64_t)rt.pairs.size();
        ws->leaves = (int64_t)rt.leaves.size();
        if (ws->near_pairs) {
            ws->d_pairs.allocate(rt.pairs.size());
            copy_on_stream(ctx, ws->d_pairs.ptr, rt.pairs.data(), ws->d_pairs.bytes(), hipMemcpyHostToDevice);
        }
        if (ws->leaves) {
            ws->d_leaves.allocate(rt.leaves.size());
            copy_on_stream(ctx, ws->d_leaves.ptr, rt.leaves.data(), ws->d_leaves.bytes(), hipMemcpyHostToDevice);
        }
        ws->d_row_ptr.allocate(rt.row_ptr.size());
        copy_on_stream(ctx, ws->d_row_ptr.ptr, rt.row_ptr.data(), ws->d_row_ptr.bytes(), hipMemcpyHostToDevice);
        ws->build_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        context_retain(ctx);
        *out = ws.release();
    });
}

void sbh_workspace_destroy(sbh_workspace* ws) {
    if (!ws) return;
    sbh_context* ctx = ws->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ws;
    context_release(ctx);
}

sbh_status sbh_workspace_stats(const sbh_workspace* ws, int64_t* near_pairs, int64_t* leaves,
                               double* build_ms) {
    return guard_on(ctx_of(ws), [&] {
        SBH_REQUIRE(ws, "workspace is null");
        if (near_pairs) *near_pairs = ws->near_pairs;
        if (leaves) *leaves = ws->leaves;
        if (build_ms) *build_ms = ws->build_ms;
    });
}

sbh_status sbh_system_create(sbh_context* ctx, int64_t num_dofs, sbh_system** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && out, "ctx/out is null");
        SBH_REQUIRE(num_dofs >= 1, "num_dofs must be positive");
        *out = nullptr;
        use_device(ctx);
        SBH_REQUIRE(num_dofs <= ((int64_t)1 << 31), "num_dofs too large");
        const int64_t stride = (num_dofs + 255) / 256 + 1;
        const int64_t bytes = (int64_t)sizeof(double2) * (num_dofs * num_dofs + num_dofs) +
                              (int64_t)sizeof(double) * num_dofs * stride;
        admit_device(ctx, bytes, "a dense system", false);
        std::unique_ptr<sbh_system> s(new sbh_system());
        s->ctx = ctx;
        s->n = num_dofs;
        s->lda = num_dofs;
        s->d_A.allocate((size_t)num_dofs * (size_t)num_dofs);
        s->d_rhs.allocate((size_t)num_dofs);
        s->d_rowmax.allocate((size_t)num_dofs * stride);
        s->rowmax_stride = stride;
        SBH_HIP(hipMemsetAsync(s->d_rhs.ptr, 0, s->d_rhs.bytes(), ctx->stream));
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        ctx->device_reserved += bytes;
        s->admitted = bytes;
        context_retain(ctx);
        *out = s.release();
    });
}

static void system_release(sbh_system* sys) {
    if (sys->refs.fetch_sub(1) != 1) return;
    sbh_context* ctx = sys->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_device(ctx, sys->admitted);
    delete sys;
    context_release(ctx);
}

void sbh_system_destroy(sbh_system* sys) {
    if (sys) system_release(sys);
}

int64_t sbh_system_num_dofs(const sbh_system* sys) { return sys ? sys->n : 0; }

sbh_status sbh_system_device_pointers(sbh_system* sys, sbh_c64** matrix, int64_t* lda, sbh_c64** rhs) {
    return guard_on(ctx_of(sys), [&] {
        SBH_REQUIRE(sys, "system is null");
        if (matrix) {
            // the caller may now write the matrix behind the library's back
            sys->matrix_exposed = true;
            sys->rowmax_valid = false;
            *matrix = reinterpret_cast<sbh_c64*>(sys->d_A.ptr);
        }
        if (lda) *lda = sys->lda;
        if (rhs) *rhs = reinterpret_cast<sbh_c64*>(sys->d_rhs.ptr);
    });
}

sbh_status sbh_system_download(sbh_system* sys, int64_t row_begin, int64_t row_end, sbh_c64* rows,
                               sbh_c64* rhs) {
    return guard_on(ctx_of(sys), [&] {
        SBH_REQUIRE(sys, "system is null");
        SBH_REQUIRE(row_begin >= 0 && row_begin <= row_end && row_end <= sys->n, "row range out of bounds");
        use_device(sys->ctx);
        const int64_t nr = row_end - row_begin;
        if (rows && nr > 0)
            SBH_HIP(hipMemcpy2DAsync(rows, sizeof(double2) * sys->n, sys->d_A.ptr + row_begin * sys->lda,
                                     sizeof(double2) * sys->lda, sizeof(double2) * sys->n, (size_t)nr,
                                     hipMemcpyDeviceToHost, sys->ctx->stream));
        if (rhs)
            SBH_HIP(hipMemcpyAsync(rhs, sys->d_rhs.ptr, sizeof(double2) * sys->n, hipMemcpyDeviceToHost,
                                   sys->ctx->stream));
        SBH_HIP(hipStreamSynchronize(sys->ctx->stream));
    });
}

sbh_status sbh_system_set_rhs(sbh_system* sys, const sbh_c64* rhs, int memory) {
    return guard_on(ctx_of(sys), [&] {
        SBH_REQUIRE(sys && rhs, "system/rhs is null");
        check_memory(memory);
        use_device(sys->ctx);
        SBH_HIP(hipMemcpyAsync(sys->d_rhs.ptr, rhs, sizeof(double2) * sys->n,
                               memory == SBH_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               sys->ctx->stream));
        SBH_HIP(hipStreamSynchronize(sys->ctx->stream));
    });
}

sbh_status sbh_workspace_assemble(sbh_workspace* ws, const sbh_physics* ph, sbh_c64 beta, sbh_system* sys) {
    return guard_on(ctx_of(ws), [&] {
        SBH_REQUIRE(ws && ph && sys, "workspace/physics/system is null");
        SBH_REQUIRE(sys->ctx == ws->ctx, "system belongs to another context");
        SBH_REQUIRE(sys->n == ws->mesh->n, "system size does not match the mesh");
        SBH_REQUIRE(std::isfinite(ph->wave_number) && ph->wave_number > 0.0, "wave number must be positive");
        SBH_REQUIRE(std::isfinite(beta.re) && std::isfinite(beta.im), "beta must be finite");
        use_device(ws->ctx);
        launch_assembly(ws, *ph, beta, sys);
    });
}

sbh_status sbh_workspace_assemble_derivative(sbh_workspace* ws, const sbh_physics* ph, sbh_c64 beta,
                                             sbh_c64 beta_derivative, double sign, double wavenumber_scale,
                                             double pressure_factor_derivative, double admittance_scale,
                                             sbh_system* sys) {
    return guard_on(ctx_of(ws), [&] {
        SBH_REQUIRE(ws && ph && sys, "workspace/physics/system is null");
        SBH_REQUIRE(sys->ctx == ws->ctx, "system belongs to another context");
        SBH_REQUIRE(sys->n == ws->mesh->n, "system size does not match the mesh");
        SBH_REQUIRE(std::isfinite(ph->wave_number) && ph->wave_number > 0.0, "wave number must be positive");
        SBH_REQUIRE(std::isfinite(beta.re) && std::isfinite(beta.im) && std::isfinite(beta_derivative.re) &&
                        std::isfinite(beta_derivative.im),
                    "beta and its derivative must be finite");
        SBH_REQUIRE(std::isfinite(sign) && std::isfinite(wavenumber_scale) &&
                        std::isfinite(pressure_factor_derivative) && std::isfinite(admittance_scale),
                    "derivative parameters must be finite");
        use_device(ws->ctx);
        launch_derivative_assembly(ws, *ph, beta, beta_derivative, sign, wavenumber_scale,
                                   pressure_factor_derivative, admittance_scale, sys);
    });
}

sbh_status sbh_build_tbem_system(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* ph,
                                 sbh_system* sys) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && mesh && ph && sys, "null argument");
        sbh_workspace* ws = nullptr;
        sbh_status st = sbh_workspace_create(ctx, mesh, 0, &ws);
        if (st != SBH_OK) throw Error(st, g_last_error);
        std::unique_ptr<sbh_workspace, void (*)(sbh_workspace*)> hold(ws, sbh_workspace_destroy);
        st = sbh_workspace_assemble(ws, ph, sbh_burton_miller_beta_scaled(ph, 1.0), sys);
        if (st != SBH_OK) throw Error(st, g_last_error);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

sbh_status sbh_incident_plane_wave_rhs(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* ph,
                                       sbh_c64 beta, double sign, const double* directions,
                                       const sbh_c64* amplitudes, int64_t num_rhs, sbh_c64* rhs,
                                       int memory, int accumulate) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && mesh && ph && directions && amplitudes && rhs, "null argument");
        SBH_REQUIRE(num_rhs >= 1, "num_rhs must be positive");
        check_memory(memory);
        use_device(ctx);
        std::vector<double> dirs(3 * num_rhs);
        for (int64_t c = 0; c < num_rhs; ++c) {
            // IncidentField::plane_wave (incident.rs:185-198) normalizes
            const double* d = directions + 3 * c;
            double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            SBH_REQUIRE(std::isfinite(len), "directions must be finite");
            if (len > 1e-10) {
                dirs[3 * c] = d[0] / len;
                dirs[3 * c + 1] = d[1] / len;
                dirs[3 * c + 2] = d[2] / len;
            } else {
                dirs[3 * c] = 0.0;
                dirs[3 * c + 1] = 0.0;
                dirs[3 * c + 2] = -1.0;
            }
        }
        DeviceArray<double> d_dirs(3 * num_rhs);
        DeviceArray<double2> d_amps(num_rhs);
        SBH_HIP(hipMemcpyAsync(d_dirs.ptr, dirs.data(), d_dirs.bytes(), hipMemcpyHostToDevice, ctx->stream));
        SBH_HIP(hipMemcpyAsync(d_amps.ptr, amplitudes, d_amps.bytes(), hipMemcpyHostToDevice, ctx->stream));
        const size_t total = (size_t)mesh->n * (size_t)num_rhs;
        if (memory == SBH_MEM_DEVICE) {
            launch_incident_rhs(ctx, mesh, *ph, beta, sign, d_dirs.ptr, d_amps.ptr, num_rhs,
                                reinterpret_cast<double2*>(rhs), accumulate != 0);
            SBH_HIP(hipStreamSynchronize(ctx->stream));
        } else {
            DeviceArray<double2> tmp(total);
            if (accumulate)
                SBH_HIP(hipMemcpyAsync(tmp.ptr, rhs, tmp.bytes(), hipMemcpyHostToDevice, ctx->stream));
            launch_incident_rhs(ctx, mesh, *ph, beta, sign, d_dirs.ptr, d_amps.ptr, num_rhs, tmp.ptr,
                                accumulate != 0);
            SBH_HIP(hipMemcpyAsync(rhs, tmp.ptr, tmp.bytes(), hipMemcpyDeviceToHost, ctx->stream));
            SBH_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

// IncidentField::plane_wave (incident.rs:185-198) normalizes the direction;
// a zero direction becomes -z.
static std::vector<double> normalized_directions(const double* directions, int64_t count) {
    std::vector<double> dirs(3 * count);
    for (int64_t c = 0; c < count; ++c) {
        const double* d = directions + 3 * c;
        double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        SBH_REQUIRE(std::isfinite(len), "directions must be finite");
        if (len > 1e-10) {
            dirs[3 * c] = d[0] / len;
            dirs[3 * c + 1] = d[1] / len;
            dirs[3 * c + 2] = d[2] / len;
        } else {
            dirs[3 * c] = 0.0;
            dirs[3 * c + 1] = 0.0;
            dirs[3 * c + 2] = -1.0;
        }
    }
    return dirs;
}

// Validation and staging shared by the field entry points; sources are host memory.
static void field_pressure_impl(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* ph, const sbh_c64* solution,
                                const sbh_c64* surface_velocity, const std::vector<IncidentSource>& sources,
                                const double* points, int64_t num_points, sbh_c64* scattered, sbh_c64* incident,
                                sbh_c64* total, int memory) {
    SBH_REQUIRE(ctx && mesh && ph && solution, "null argument");
    SBH_REQUIRE(num_points >= 0, "counts must be non-negative");
    SBH_REQUIRE(num_points == 0 || (points && scattered), "points/scattered is null");
    SBH_REQUIRE(std::isfinite(ph->wave_number) && std::isfinite(ph->harmonic_factor) &&
                    std::isfinite(ph->pressure_factor),
                "physics parameters must be finite");
    check_memory(memory);
    use_device(ctx);
    if (num_points == 0) return;
    const int64_t n = mesh->n;
    for (const IncidentSource& src : sources) {
        SBH_REQUIRE(src.kind == SBH_SOURCE_PLANE_WAVE || src.kind == SBH_SOURCE_POINT, "unknown source kind");
        SBH_REQUIRE(std::isfinite(src.v[0]) && std::isfinite(src.v[1]) && std::isfinite(src.v[2]) &&
                        std::isfinite(src.amp[0]) && std::isfinite(src.amp[1]),
                    "incident sources must be finite");
    }
    // validate_field_inputs (pressure.rs:627-631)
    std::vector<double> pts_copy;
    const double* host_pts = points;
    if (memory == SBH_MEM_DEVICE) {
        pts_copy.resize((size_t)(3 * num_points));
        copy_on_stream(ctx, pts_copy.data(), points, pts_copy.size() * sizeof(double), hipMemcpyDeviceToHost);
        host_pts = pts_copy.data();
    }
    for (int64_t i = 0; i < 3 * num_points; ++i)
        if (!std::isfinite(host_pts[i]))
            throw Error(SBH_ERR_INVALID_ARGUMENT,
                        "evaluation point " + std::to_string(i / 3) + " has non-finite coordinates");
    // NonFiniteSurfaceSample (validate_field_inputs, pressure.rs:690-733)
    for (const sbh_c64* samples : {solution, surface_velocity}) {
        if (!samples) continue;
        std::vector<sbh_c64> copy;
        const sbh_c64* h = samples;
        if (memory == SBH_MEM_DEVICE) {
            copy.resize((size_t)n);
            copy_on_stream(ctx, copy.data(), samples, copy.size() * sizeof(sbh_c64), hipMemcpyDeviceToHost);
            h = copy.data();
        }
        for (int64_t j = 0; j < n; ++j)
            if (!std::isfinite(h[j].re) || !std::isfinite(h[j].im))
                throw Error(SBH_ERR_INVALID_ARGUMENT, "surface sample " + std::to_string(j) + " is not finite");
    }
    DeviceArray<IncidentSource> d_src(sources.size());
    if (!sources.empty())
        SBH_HIP(hipMemcpyAsync(d_src.ptr, sources.data(), d_src.bytes(), hipMemcpyHostToDevice, ctx->stream));
    const double2* d_sol = reinterpret_cast<const double2*>(solution);
    const double2* d_vel = reinterpret_cast<const double2*>(surface_velocity);
    const double* d_pts = points;
    double2* outs[3] = {reinterpret_cast<double2*>(scattered), reinterpret_cast<double2*>(incident),
                        reinterpret_cast<double2*>(total)};
    DeviceArray<double2> h_sol, h_vel, h_out[3];
    DeviceArray<double> h_pts;
    if (memory == SBH_MEM_HOST) {
        h_sol.allocate((size_t)n);
        h_pts.allocate((size_t)(3 * num_points));
        SBH_HIP(hipMemcpyAsync(h_sol.ptr, solution, h_sol.bytes(), hipMemcpyHostToDevice, ctx->stream));
        SBH_HIP(hipMemcpyAsync(h_pts.ptr, points, h_pts.bytes(), hipMemcpyHostToDevice, ctx->stream));
        if (surface_velocity) {
            h_vel.allocate((size_t)n);
            SBH_HIP(hipMemcpyAsync(h_vel.ptr, surface_velocity, h_vel.bytes(), hipMemcpyHostToDevice, ctx->stream));
            d_vel = h_vel.ptr;
        }
        d_sol = h_sol.ptr;
        d_pts = h_pts.ptr;
        for (int o = 0; o < 3; ++o)
            if (outs[o]) {
                h_out[o].allocate((size_t)num_points);
                outs[o] = h_out[o].ptr;
            }
    }
    launch_field(ctx, mesh, *ph, d_sol, d_vel, d_pts, num_points, d_src.ptr, (int64_t)sources.size(), outs[0],
                 outs[1], outs[2]);
    // validate_finite_results (pressure.rs:735-742): every returned value must be finite
    sbh_c64* user[3] = {scattered, incident, total};
    std::vector<sbh_c64> host_copy((size_t)num_points);
    for (int o = 0; o < 3; ++o) {
        if (!outs[o]) continue;
        copy_on_stream(ctx, host_copy.data(), outs[o], host_copy.size() * sizeof(sbh_c64), hipMemcpyDeviceToHost);
        for (int64_t i = 0; i < num_points; ++i)
            if (!std::isfinite(host_copy[i].re) || !std::isfinite(host_copy[i].im))
                throw Error(SBH_ERR_NUMERICAL, "field value " + std::to_string(i) + " is not finite");
        if (memory == SBH_MEM_HOST) std::memcpy(user[o], host_copy.data(), host_copy.size() * sizeof(sbh_c64));
    }
}

sbh_status sbh_field_pressure(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* ph, const sbh_c64* solution,
                              const sbh_c64* surface_velocity, const double* directions, const sbh_c64* amplitudes,
                              int64_t num_waves, const double* points, int64_t num_points, sbh_c64* scattered,
                              sbh_c64* incident, sbh_c64* total, int memory) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(num_waves >= 0, "counts must be non-negative");
        SBH_REQUIRE(num_waves == 0 || (directions && amplitudes), "directions/amplitudes is null");
        const std::vector<double> dirs = normalized_directions(directions, num_waves);
        std::vector<IncidentSource> src((size_t)num_waves);
        for (int64_t w = 0; w < num_waves; ++w) {
            src[w].kind = SBH_SOURCE_PLANE_WAVE;
            src[w].pad = 0;
            for (int d = 0; d < 3; ++d) src[w].v[d] = dirs[3 * w + d];
            src[w].amp[0] = amplitudes[w].re;
            src[w].amp[1] = amplitudes[w].im;
        }
        field_pressure_impl(ctx, mesh, ph, solution, surface_velocity, src, points, num_points, scattered, incident,
                            total, memory);
    });
}

sbh_status sbh_field_pressure_sources(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* ph,
                                      const sbh_c64* solution, const sbh_c64* surface_velocity,
                                      const sbh_incident_source* sources, int64_t num_sources, const double* points,
                                      int64_t num_points, sbh_c64* scattered, sbh_c64* incident, sbh_c64* total,
                                      int memory) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(num_sources >= 0 && (num_sources == 0 || sources), "sources is null");
        const IncidentSource* s = reinterpret_cast<const IncidentSource*>(sources);
        std::vector<IncidentSource> src(s, s + num_sources);
        field_pressure_impl(ctx, mesh, ph, solution, surface_velocity, src, points, num_points, scattered, incident,
                            total, memory);
    });
}

sbh_status sbh_incident_rhs(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* ph, sbh_c64 beta, double sign,
                            const sbh_incident_source* sources, int64_t num_sources, sbh_c64* rhs, int memory,
                            int accumulate) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && mesh && ph && rhs, "null argument");
        SBH_REQUIRE(num_sources >= 0 && (num_sources == 0 || sources), "sources is null");
        check_memory(memory);
        use_device(ctx);
        const IncidentSource* s = reinterpret_cast<const IncidentSource*>(sources);
        for (int64_t w = 0; w < num_sources; ++w) {
            SBH_REQUIRE(s[w].kind == SBH_SOURCE_PLANE_WAVE || s[w].kind == SBH_SOURCE_POINT, "unknown source kind");
            SBH_REQUIRE(std::isfinite(s[w].v[0]) && std::isfinite(s[w].v[1]) && std::isfinite(s[w].v[2]) &&
                            std::isfinite(s[w].amp[0]) && std::isfinite(s[w].amp[1]),
                        "incident sources must be finite");
        }
        DeviceArray<IncidentSource> d_src((size_t)num_sources);
        if (num_sources)
            SBH_HIP(hipMemcpyAsync(d_src.ptr, s, d_src.bytes(), hipMemcpyHostToDevice, ctx->stream));
        if (memory == SBH_MEM_DEVICE) {
            launch_source_rhs(ctx, mesh, *ph, beta, sign, d_src.ptr, num_sources, reinterpret_cast<double2*>(rhs),
                              accumulate != 0);
            SBH_HIP(hipStreamSynchronize(ctx->stream));
        } else {
            DeviceArray<double2> tmp((size_t)mesh->n);
            if (accumulate) SBH_HIP(hipMemcpyAsync(tmp.ptr, rhs, tmp.bytes(), hipMemcpyHostToDevice, ctx->stream));
            launch_source_rhs(ctx, mesh, *ph, beta, sign, d_src.ptr, num_sources, tmp.ptr, accumulate != 0);
            SBH_HIP(hipMemcpyAsync(rhs, tmp.ptr, tmp.bytes(), hipMemcpyDeviceToHost, ctx->stream));
            SBH_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

sbh_status sbh_dense_operator_create(sbh_system* sys, sbh_operator** out) {
    return guard_on(ctx_of(sys), [&] {
        SBH_REQUIRE(sys && out, "system/out is null");
        sbh_operator* op = new sbh_operator();
        op->ctx = sys->ctx;
        op->sys = sys;
        op->n = sys->n;
        context_retain(op->ctx);
        *out = op;
    });
}

sbh_status sbh_exact_operator_create(sbh_workspace* ws, const sbh_physics* physics, sbh_c64 beta, double sign,
                                     sbh_operator** out) {
    return guard_on(ctx_of(ws), [&] {
        SBH_REQUIRE(ws && physics && out, "workspace/physics/out is null");
        SBH_REQUIRE(std::isfinite(sign) && std::isfinite(beta.re) && std::isfinite(beta.im),
                    "sign and beta must be finite");
        const sbh_mesh* m = ws->mesh;
        SBH_REQUIRE(std::isfinite(physics->wave_number), "wave number is not finite");
        SBH_REQUIRE(std::fabs(physics->wave_number) * m->diameter < 5.0e5,
                    "k * mesh diameter exceeds the supported phase range (5e5 rad)");
        if (physics->harmonic_factor != 1.0)
            throw Error(SBH_ERR_UNSUPPORTED, "only harmonic_factor = +1 (ACOUSTIC_CONVENTION) is supported");
        use_device(ws->ctx);
        std::unique_ptr<sbh_operator> op(new sbh_operator());
        op->ctx = ws->ctx;
        op->ws = ws;
        op->n = m->n;
        op->physics = *physics;
        op->beta = make_double2(beta.re, beta.im);
        op->sign = sign;
        op->d_far.allocate((size_t)m->n);
        op->d_self.allocate((size_t)m->n);
        op->d_near.allocate((size_t)std::max<int64_t>(ws->near_pairs, 1));
        context_retain(op->ctx);
        *out = op.release();
    });
}

void sbh_operator_destroy(sbh_operator* op) {
    if (!op) return;
    sbh_context* ctx = op->ctx;
    delete op;
    context_release(ctx);
}
int64_t sbh_operator_num_rows(const sbh_operator* op) { return op ? op->n : 0; }
int64_t sbh_operator_num_cols(const sbh_operator* op) { return op ? op->n : 0; }
int sbh_operator_transpose_capability(const sbh_operator* op) {
    return !op ? SBH_ADJOINT_UNKNOWN : !op->dense() ? SBH_ADJOINT_UNSUPPORTED : SBH_ADJOINT_EXACT;
}
int sbh_operator_hermitian_capability(const sbh_operator* op) {
    return !op ? SBH_ADJOINT_UNKNOWN : !op->dense() ? SBH_ADJOINT_UNSUPPORTED : SBH_ADJOINT_EXACT;
}


static sbh_status operator_apply(sbh_operator* op, const sbh_c64* x, sbh_c64* y, int memory, int mode) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && x && y, "operator/x/y is null");
        check_memory(memory);
        if (mode != 0) require_dense(op, "the transpose/Hermitian product");
        const int64_t n = op->n;
        if (memory == SBH_MEM_HOST && !all_finite(x, (size_t)n))
            throw Error(SBH_ERR_NUMERICAL, "operator input contains non-finite values");
        use_device(op->ctx);
        DeviceInput in(op->ctx, x, (size_t)n, memory);
        DeviceOutput outv(op->ctx, y, (size_t)n, memory);
        if (mode == 0)
            operator_apply_device(op, in.ptr, outv.ptr);
        else
            zgemv_t(op->ctx, op->sys->d_A.ptr, op->sys->lda, n, n, in.ptr, outv.ptr, mode == 2);
        outv.finish();
    });
}

sbh_status sbh_operator_apply(sbh_operator* op, const sbh_c64* x, sbh_c64* y, int memory) {
    return operator_apply(op, x, y, memory, 0);
}
sbh_status sbh_operator_apply_transpose(sbh_operator* op, const sbh_c64* x, sbh_c64* y, int memory) {
    return operator_apply(op, x, y, memory, 1);
}
sbh_status sbh_operator_apply_hermitian(sbh_operator* op, const sbh_c64* x, sbh_c64* y, int memory) {
    return operator_apply(op, x, y, memory, 2);
}

sbh_status sbh_operator_apply_block(sbh_operator* op, const sbh_c64* X, sbh_c64* Y, int64_t k, int memory) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && X && Y, "operator/X/Y is null");
        SBH_REQUIRE(k >= 1, "block width must be positive");
        check_memory(memory);
        const int64_t n = op->n;
        if (memory == SBH_MEM_HOST && !all_finite(X, (size_t)(n * k)))
            throw Error(SBH_ERR_NUMERICAL, "operator input contains non-finite values");
        use_device(op->ctx);
        DeviceInput in(op->ctx, X, (size_t)(n * k), memory);
        DeviceOutput outv(op->ctx, Y, (size_t)(n * k), memory);
        if (k == 1) {
            operator_apply_device(op, in.ptr, outv.ptr);
        } else if (!op->dense()) {
            // column by column (the trait's default apply_block): strided gather, product, scatter
            DeviceArray<double2> xc((size_t)n), yc((size_t)n);
            for (int64_t c = 0; c < k; ++c) {
                SBH_HIP(hipMemcpy2DAsync(xc.ptr, sizeof(double2), in.ptr + c, sizeof(double2) * k, sizeof(double2),
                                         n, hipMemcpyDeviceToDevice, op->ctx->stream));
                operator_apply_device(op, xc.ptr, yc.ptr);
                SBH_HIP(hipMemcpy2DAsync(outv.ptr + c, sizeof(double2) * k, yc.ptr, sizeof(double2), sizeof(double2),
                                         n, hipMemcpyDeviceToDevice, op->ctx->stream));
            }
        } else {
            zgemm_f64(op->ctx, op->sys->d_A.ptr, op->sys->lda, n, in.ptr, k, outv.ptr);
        }
        outv.finish();
    });
}

sbh_status sbh_operator_apply_block_noalloc(sbh_operator* op, const sbh_c64* X, sbh_c64* Y, int64_t k) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && X && Y, "operator/X/Y is null");
        SBH_REQUIRE(k >= 1, "block width must be positive");
        if (!op->dense())
            throw Error(SBH_ERR_UNSUPPORTED, "the allocation-free block product needs a dense (stored) operator");
        const int64_t n = op->n;
        if (k >= 2) {
            const size_t need = noalloc_block_scratch_need(n, k);
            if (need > NOALLOC_SCRATCH_DOUBLES)
                throw Error(SBH_ERR_INVALID_ARGUMENT,
                            "n x k too large for the allocation-free block product: its partial sums need " +
                                std::to_string(need * sizeof(double) >> 20) + " MiB, the fixed scratch holds " +
                                std::to_string(NOALLOC_SCRATCH_DOUBLES * sizeof(double) >> 20) + " MiB");
        }
        const double2* x = reinterpret_cast<const double2*>(X);
        double2* y = reinterpret_cast<double2*>(Y);
        if (k == 1)
            zgemv_n_untimed(op->ctx, op->sys->d_A.ptr, op->sys->lda, n, n, x, y);
        else
            zgemm_f64_noalloc(op->ctx, op->sys->d_A.ptr, op->sys->lda, n, x, k, y);
        SBH_HIP(hipStreamSynchronize(op->ctx->stream));
    });
}

sbh_status sbh_precond_identity(sbh_context* ctx, int64_t n, sbh_precond** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && out && n >= 1, "invalid identity preconditioner arguments");
        sbh_precond* p = new sbh_precond();
        p->ctx = ctx;
        p->n = n;
        p->kind = 0;
        context_retain(ctx);
        *out = p;
    });
}

sbh_status sbh_precond_jacobi(sbh_operator* op, sbh_precond** out) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && out, "operator/out is null");
        require_dense(op, "the Jacobi preconditioner");
        use_device(op->ctx);
        std::unique_ptr<sbh_precond> p(new sbh_precond());
        p->ctx = op->ctx;
        p->n = op->n;
        p->kind = 1;
        p->d_dinv.allocate(p->n);
        extract_inverse_diagonal(op->ctx, op->sys->d_A.ptr, op->sys->lda, p->n, p->d_dinv.ptr);
        SBH_HIP(hipStreamSynchronize(op->ctx->stream));
        context_retain(op->ctx);
        *out = p.release();
    });
}

sbh_status sbh_precond_create_callback(sbh_context* ctx, int64_t n, sbh_apply_fn apply, void* user, int memory,
                                       sbh_precond** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && apply && out && n >= 1, "invalid callback preconditioner arguments");
        check_memory(memory);
        std::unique_ptr<sbh_precond> p(new sbh_precond());
        p->ctx = ctx;
        p->n = n;
        p->kind = 2;
        p->callback = apply;
        p->user = user;
        p->callback_memory = memory;
        context_retain(ctx);
        *out = p.release();
    });
}

sbh_status sbh_operator_create_callback(sbh_context* ctx, int64_t n, sbh_apply_fn apply, void* user, int memory,
                                        sbh_operator** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && apply && out && n >= 1, "invalid callback operator arguments");
        check_memory(memory);
        std::unique_ptr<sbh_operator> op(new sbh_operator());
        op->ctx = ctx;
        op->n = n;
        op->callback = apply;
        op->user = user;
        op->callback_memory = memory;
        context_retain(ctx);
        *out = op.release();
    });
}

void sbh_precond_destroy(sbh_precond* p) {
    if (!p) return;
    sbh_context* ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete p;
    context_release(ctx);
}

sbh_status sbh_precond_apply(sbh_precond* p, const sbh_c64* r, sbh_c64* y, int memory) {
    return guard_on(ctx_of(p), [&] {
        SBH_REQUIRE(p && r && y, "preconditioner/r/y is null");
        check_memory(memory);
        use_device(p->ctx);
        DeviceInput in(p->ctx, r, (size_t)p->n, memory);
        DeviceOutput outv(p->ctx, y, (size_t)p->n, memory);
        if (p->kind == 0)
            SBH_HIP(hipMemcpyAsync(outv.ptr, in.ptr, sizeof(double2) * p->n, hipMemcpyDeviceToDevice, p->ctx->stream));
        else if (p->kind == 1)
            diag_scale(p->ctx, p->d_dinv.ptr, in.ptr, outv.ptr, p->n);
        else
            callback_apply_device(p->ctx, p->callback, p->user, p->callback_memory, in.ptr, outv.ptr, p->n,
                                  "preconditioner");
        outv.finish();
    });
}

sbh_status sbh_gmres(sbh_operator* op, sbh_precond* precond, const sbh_c64* b, const sbh_c64* x0,
                     const sbh_gmres_config* config, sbh_c64* x, int memory,
                     sbh_gmres_solution* solution) {
    sbh_status st = SBH_OK;
    sbh_status g = guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && b && config && x && solution, "null argument");
        check_memory(memory);
        if (precond) SBH_REQUIRE(precond->n == op->n, "preconditioner size does not match the operator");
        const int64_t n = op->n;
        use_device(op->ctx);
        DeviceInput db(op->ctx, b, (size_t)n, memory);
        std::unique_ptr<DeviceInput> dx0;
        if (x0) dx0.reset(new DeviceInput(op->ctx, x0, (size_t)n, memory));
        DeviceOutput dx(op->ctx, x, (size_t)n, memory);
        st = run_gmres(op, precond, db.ptr, dx0 ? dx0->ptr : nullptr, *config, dx.ptr, solution);
        dx.finish();
    });
    return g != SBH_OK ? g : st;
}

sbh_status sbh_gmres_block(sbh_operator* op, const sbh_c64* B, const sbh_c64* X0, int64_t k,
                           const sbh_gmres_config* config, sbh_c64* X, int memory,
                           sbh_gmres_solution* solutions) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && B && config && X && solutions, "null argument");
        SBH_REQUIRE(k >= 1 && k <= 64, "block width must be in 1..64");
        check_memory(memory);
        require_dense(op, "block GMRES");
        if (config->convergence_check_interval > 1 || config->min_restart != 0 || config->max_restart != 0)
            throw Error(SBH_ERR_UNSUPPORTED, "block GMRES checks convergence every step with a fixed restart");
        const int64_t n = op->n;
        use_device(op->ctx);
        DeviceInput db(op->ctx, B, (size_t)(n * k), memory);
        DeviceOutput dx(op->ctx, X, (size_t)(n * k), memory);
        if (X0) {
            DeviceInput dx0(op->ctx, X0, (size_t)(n * k), memory);
            SBH_HIP(hipMemcpyAsync(dx.ptr, dx0.ptr, sizeof(double2) * n * k, hipMemcpyDeviceToDevice, op->ctx->stream));
            SBH_HIP(hipStreamSynchronize(op->ctx->stream));
        } else {
            SBH_HIP(hipMemsetAsync(dx.ptr, 0, sizeof(double2) * n * k, op->ctx->stream));
        }
        sbh_status st = run_gmres_block(op, db.ptr, dx.ptr, (int)k, *config, X0 == nullptr, solutions);
        dx.finish();
        if (st != SBH_OK) throw Error(st, "block GMRES failed");
    });
}

sbh_status sbh_block_gmres(sbh_operator* op, sbh_precond* precond, const sbh_c64* B, const sbh_c64* X0, int64_t k,
                           const sbh_block_gmres_config* config, sbh_c64* X, int memory, sbh_block_gmres_info* info,
                           double* residuals, double* true_residuals, double* backward_errors) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && B && config && X, "null argument");
        SBH_REQUIRE(k >= 1 && k <= 64, "block width must be in 1..64");
        check_memory(memory);
        // BlockGmresConfig::validate (block_gmres.rs:91-115)
        SBH_REQUIRE(config->max_iterations > 0, "max_iterations must be greater than zero");
        SBH_REQUIRE(config->restart > 0, "restart must be greater than zero");
        SBH_REQUIRE(std::isfinite(config->tolerance) && config->tolerance > 0.0,
                    "tolerance must be finite and positive");
        SBH_REQUIRE(config->exact_residual_check_interval >= 0, "exact_residual_check_interval must be >= 0");
        if (precond) {
            SBH_REQUIRE(precond->ctx == op->ctx && precond->n == op->n, "preconditioner does not match the operator");
            if (precond->kind == 2)
                throw Error(SBH_ERR_UNSUPPORTED, "block GMRES takes the identity or Jacobi preconditioner");
        }
        const int64_t n = op->n;
        if (memory == SBH_MEM_HOST) {
            if (!all_finite(B, (size_t)(n * k)))
                throw Error(SBH_ERR_NUMERICAL, "right-hand side contains NaN or infinity");
            if (X0 && !all_finite(X0, (size_t)(n * k)))
                throw Error(SBH_ERR_NUMERICAL, "initial guess contains NaN or infinity");
        }
        use_device(op->ctx);
        DeviceInput db(op->ctx, B, (size_t)(n * k), memory);
        DeviceOutput dx(op->ctx, X, (size_t)(n * k), memory);
        if (X0) {
            DeviceInput dx0(op->ctx, X0, (size_t)(n * k), memory);
            SBH_HIP(hipMemcpyAsync(dx.ptr, dx0.ptr, sizeof(double2) * n * k, hipMemcpyDeviceToDevice, op->ctx->stream));
            SBH_HIP(hipStreamSynchronize(op->ctx->stream));
        } else {
            SBH_HIP(hipMemsetAsync(dx.ptr, 0, sizeof(double2) * n * k, op->ctx->stream));
        }
        if (memory == SBH_MEM_DEVICE) {
            // device inputs: the same finiteness checks through one norm each
            // (bk_frob_partials writes at most 256 partials)
            DeviceArray<double> part(256 + 16);
            auto finite_norm = [&](const double2* v) {
                const int chunks = bk_frob_partials(op->ctx, v, n * k, part.ptr);
                std::vector<double> h((size_t)chunks);
                SBH_HIP(hipMemcpyAsync(h.data(), part.ptr, sizeof(double) * chunks, hipMemcpyDeviceToHost,
                                       op->ctx->stream));
                SBH_HIP(hipStreamSynchronize(op->ctx->stream));
                double acc = 0.0;
                for (double x : h) acc += x;
                return std::isfinite(acc);
            };
            if (!finite_norm(db.ptr)) throw Error(SBH_ERR_NUMERICAL, "right-hand side contains NaN or infinity");
            if (X0 && !finite_norm(dx.ptr)) throw Error(SBH_ERR_NUMERICAL, "initial guess contains NaN or infinity");
        }
        sbh_status st = run_block_gmres(op, precond, db.ptr, dx.ptr, (int)k, X0 == nullptr, *config, info, residuals,
                                        true_residuals, backward_errors);
        dx.finish();
        if (st != SBH_OK) throw Error(st, "block GMRES failed");
    });
}

sbh_status sbh_gcro_subspace_create(sbh_context* ctx, int64_t n, int64_t k, sbh_gcro_subspace** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && out, "context/out is null");
        SBH_REQUIRE(n >= 1, "n must be positive");
        SBH_REQUIRE(k >= 0 && k <= 256, "the recycled subspace size must be in 0..256");
        use_device(ctx);
        admit_device(ctx, (int64_t)sizeof(double2) * std::max<int64_t>(k, 1) * n, "a GCRO-DR recycled subspace", false);
        std::unique_ptr<sbh_gcro_subspace> g(new sbh_gcro_subspace());
        g->ctx = ctx;
        g->n = n;
        g->k = (int)k;
        g->U.allocate((size_t)std::max<int64_t>(k, 1) * n);
        context_retain(ctx);
        *out = g.release();
    });
}

void sbh_gcro_subspace_destroy(sbh_gcro_subspace* g) {
    if (!g) return;
    sbh_context* ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete g;
    context_release(ctx);
}

int64_t sbh_gcro_subspace_size(const sbh_gcro_subspace* g) { return g ? g->count : -1; }

sbh_status sbh_gcro_subspace_reset(sbh_gcro_subspace* g) {
    return guard_on(ctx_of(g), [&] {
        SBH_REQUIRE(g, "subspace is null");
        g->count = 0;
    });
}

sbh_status sbh_gcro_subspace_download(const sbh_gcro_subspace* g, sbh_c64* u) {
    return guard_on(ctx_of(g), [&] {
        SBH_REQUIRE(g && u, "null argument");
        use_device(g->ctx);
        injection_check(INJECT_TRANSFER_D2H);
        copy_on_stream(g->ctx, u, g->U.ptr, sizeof(double2) * (size_t)g->count * (size_t)g->n,
                       hipMemcpyDeviceToHost);
    });
}

sbh_status sbh_gcro_dr(sbh_operator* op, sbh_precond* precond, const sbh_c64* b, const sbh_c64* x0,
                       sbh_gcro_subspace* subspace, const sbh_gcro_dr_config* config, sbh_c64* x, int memory,
                       sbh_gmres_solution* solution) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && b && subspace && config && x && solution, "null argument");
        check_memory(memory);
        // validate_gcro_request (gcro_dr.rs:240-290)
        SBH_REQUIRE(config->max_iterations > 0 && config->restart > 0 && std::isfinite(config->tolerance) &&
                        config->tolerance > 0.0,
                    "invalid GCRO-DR configuration");
        SBH_REQUIRE(subspace->ctx == op->ctx && subspace->n == op->n, "the recycled subspace does not match the operator");
        if (precond) SBH_REQUIRE(precond->ctx == op->ctx && precond->n == op->n, "preconditioner does not match the operator");
        const int64_t n = op->n;
        if (memory == SBH_MEM_HOST) {
            if (!all_finite(b, (size_t)n)) throw Error(SBH_ERR_NUMERICAL, "right-hand side contains NaN or infinity");
            if (x0 && !all_finite(x0, (size_t)n)) throw Error(SBH_ERR_NUMERICAL, "initial guess contains NaN or infinity");
        }
        use_device(op->ctx);
        DeviceInput db(op->ctx, b, (size_t)n, memory);
        std::unique_ptr<DeviceInput> dx0;
        if (x0) dx0.reset(new DeviceInput(op->ctx, x0, (size_t)n, memory));
        DeviceOutput dx(op->ctx, x, (size_t)n, memory);
        if (memory == SBH_MEM_DEVICE) {
            DeviceArray<double> part(256 + 16);
            auto finite_norm = [&](const double2* v) {
                const int chunks = bk_frob_partials(op->ctx, v, n, part.ptr);
                std::vector<double> h((size_t)chunks);
                SBH_HIP(hipMemcpyAsync(h.data(), part.ptr, sizeof(double) * chunks, hipMemcpyDeviceToHost,
                                       op->ctx->stream));
                SBH_HIP(hipStreamSynchronize(op->ctx->stream));
                double acc = 0.0;
                for (double v2 : h) acc += v2;
                return std::isfinite(acc);
            };
            if (!finite_norm(db.ptr)) throw Error(SBH_ERR_NUMERICAL, "right-hand side contains NaN or infinity");
            if (dx0 && !finite_norm(dx0->ptr)) throw Error(SBH_ERR_NUMERICAL, "initial guess contains NaN or infinity");
        }
        run_gcro_dr(op, precond, db.ptr, dx0 ? dx0->ptr : nullptr, subspace, *config, dx.ptr, solution);
        dx.finish();
    });
}

sbh_status sbh_lowp_matrix_create(sbh_system* sys, int format, sbh_lowp_matrix** out) {
    return guard_on(ctx_of(sys), [&] {
        SBH_REQUIRE(sys && out, "system/out is null");
        SBH_REQUIRE(format == SBH_LOWP_FP16 || format == SBH_LOWP_BF16, "format must be FP16 or BF16");
        use_device(sys->ctx);
        const int64_t ld = lowp_padded_ld(sys->n);
        const int64_t bytes = 2 * lowp_padded_rows(sys->n) * 2 * ld + 8 * sys->n + 2 * 64 * 2 * ld +
                              4 * lowp_split_count() * sys->n * 128 + 64 * 12;
        admit_device(sys->ctx, bytes, "a low-precision matrix copy", false);
        std::unique_ptr<sbh_lowp_matrix> lp(new sbh_lowp_matrix());
        lp->ctx = sys->ctx;
        lp->sys = sys;
        lp->n = sys->n;
        lp->ld = ld;
        lp->format = format;
        lp->d_A.allocate((size_t)lowp_padded_rows(lp->n) * 2 * lp->ld);
        lp->d_row_inv.allocate((size_t)lp->n);
        lp->d_XT.allocate((size_t)64 * 2 * lp->ld);
        lp->d_colmax.allocate(64);
        lp->d_colinv.allocate(64);
        lp->d_ypart.allocate((size_t)lowp_split_count() * lp->n * 128);
        SBH_HIP(hipMemsetAsync(lp->d_XT.ptr, 0, lp->d_XT.bytes(), sys->ctx->stream));
        lowp_convert(sys->ctx, sys->d_A.ptr, sys->lda, lp->n, lp->ld, format, lp->d_A.ptr, lp->d_row_inv.ptr,
                     assembled_row_max(sys), sys->rowmax_stride);
        SBH_HIP(hipStreamSynchronize(sys->ctx->stream));
        sys->ctx->device_reserved += bytes;
        lp->admitted = bytes;
        context_retain(sys->ctx);
        sys->lowp_copies += 1;
        sys->refs.fetch_add(1);
        *out = lp.release();
    });
}

sbh_status sbh_lowp_matrix_update(sbh_lowp_matrix* lp) {
    return guard_on(ctx_of(lp), [&] {
        SBH_REQUIRE(lp, "matrix is null");
        use_device(lp->ctx);
        lowp_convert(lp->ctx, lp->sys->d_A.ptr, lp->sys->lda, lp->n, lp->ld, lp->format, lp->d_A.ptr,
                     lp->d_row_inv.ptr, assembled_row_max(lp->sys), lp->sys->rowmax_stride);
    });
}

void sbh_lowp_matrix_destroy(sbh_lowp_matrix* lp) {
    if (!lp) return;
    sbh_context* ctx = lp->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_device(ctx, lp->admitted);
    sbh_system* sys = lp->sys;
    sys->lowp_copies -= 1;
    delete lp;
    system_release(sys);
    context_release(ctx);
}

sbh_status sbh_lowp_matrix_apply(sbh_lowp_matrix* lp, const sbh_c64* X, sbh_c64* Y, int64_t k, int memory) {
    return guard_on(ctx_of(lp), [&] {
        SBH_REQUIRE(lp && X && Y, "null argument");
        SBH_REQUIRE(k >= 1 && k <= 64, "block width must be in 1..64");
        check_memory(memory);
        use_device(lp->ctx);
        const int64_t n = lp->n;
        DeviceInput in(lp->ctx, X, (size_t)(n * k), memory);
        DeviceOutput outv(lp->ctx, Y, (size_t)(n * k), memory);
        lowp_gemm(lp->ctx, lp->d_A.ptr, lp->ld, n, lp->format, lp->d_row_inv.ptr, in.ptr, (int)k, lp->d_XT.ptr,
                  lp->d_colmax.ptr, lp->d_colinv.ptr, lp->d_ypart.ptr, outv.ptr);
        outv.finish();
    });
}

sbh_status sbh_lowp_matrix_download(sbh_lowp_matrix* lp, uint16_t* halves, double* row_inv, int64_t* ld) {
    return guard_on(ctx_of(lp), [&] {
        SBH_REQUIRE(lp, "matrix is null");
        use_device(lp->ctx);
        SBH_HIP(hipStreamSynchronize(lp->ctx->stream));
        if (halves) {
            // the device copy is in MFMA fragment order (32-row tiles x 16 halves, 1 KB each);
            // hand it back row-major, n x 2*ld halves
            std::vector<uint16_t> packed(lp->d_A.count);
            copy_on_stream(lp->ctx, packed.data(), lp->d_A.ptr, lp->d_A.bytes(), hipMemcpyDeviceToHost);
            const int64_t steps = lp->ld / 8;
            for (int64_t row = 0; row < lp->n; ++row) {
                const int64_t rt = row / 32, r = row % 32;
                for (int64_t s = 0; s < steps; ++s)
                    for (int h = 0; h < 2; ++h) {
                        const uint16_t* src = packed.data() + ((rt * steps + s) * 64 + r + 32 * h) * 8;
                        std::copy(src, src + 8, halves + row * 2 * lp->ld + s * 16 + h * 8);
                    }
            }
        }
        if (row_inv) copy_on_stream(lp->ctx, row_inv, lp->d_row_inv.ptr, lp->d_row_inv.bytes(), hipMemcpyDeviceToHost);
        if (ld) *ld = lp->ld;
    });
}

sbh_status sbh_gmres_ir(sbh_operator* op, sbh_lowp_matrix* lp, const sbh_c64* B, int64_t k,
                        const sbh_gmres_ir_config* config, sbh_c64* X, int memory, sbh_gmres_ir_status* status) {
    return guard_on(ctx_of(op), [&] {
        SBH_REQUIRE(op && lp && B && config && X && status, "null argument");
        require_dense(op, "GMRES-IR");
        SBH_REQUIRE(lp->sys == op->sys, "the low-precision copy belongs to another system");
        SBH_REQUIRE(k >= 1 && k <= 64, "block width must be in 1..64");
        SBH_REQUIRE(config->max_outer_iterations >= 1 && config->max_inner_iterations >= 1 &&
                        std::isfinite(config->outer_tolerance) && config->outer_tolerance > 0.0 &&
                        std::isfinite(config->inner_tolerance) && config->inner_tolerance > 0.0,
                    "invalid GMRES-IR configuration");
        check_memory(memory);
        use_device(op->ctx);
        const int64_t n = op->sys->n;
        DeviceInput db(op->ctx, B, (size_t)(n * k), memory);
        DeviceOutput dx(op->ctx, X, (size_t)(n * k), memory);
        sbh_status st = run_gmres_ir(op, lp, db.ptr, dx.ptr, (int)k, *config, status);
        dx.finish();
        if (st != SBH_OK) throw Error(st, "GMRES-IR failed");
    });
}

sbh_status sbh_debug_subelements(const double* source_point, const double* element_coords,
                                 double element_area, double* leaves_out, int* count) {
    return guard([&] {
        SBH_REQUIRE(source_point && element_coords && leaves_out && count, "null argument");
        *count = subelements_for_pair(source_point, element_coords, element_area, leaves_out);
    });
}

sbh_status sbh_debug_hbm_read_rate(sbh_context* ctx, int64_t bytes, int reps, double* gbs) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && gbs && bytes >= 16 && reps >= 1, "invalid probe arguments");
        use_device(ctx);
        DeviceArray<double2> a(bytes / 16);
        DeviceArray<unsigned> sink(1);
        SBH_HIP(hipMemsetAsync(a.ptr, 0, a.bytes(), ctx->stream));
        *gbs = probe_stream_read(ctx, a.ptr, (int64_t)a.bytes(), reps, sink.ptr);
    });
}

sbh_status sbh_debug_fp64_fma_rate(sbh_context* ctx, double* tflops) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && tflops, "invalid probe arguments");
        use_device(ctx);
        DeviceArray<double> sink(1);
        *tflops = probe_fma_rate(ctx, sink.ptr);
    });
}

sbh_status sbh_debug_device_copy(sbh_context* ctx, int64_t bytes, int reps) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && bytes >= 16 && reps >= 1, "invalid copy arguments");
        use_device(ctx);
        const int64_t n = bytes / 16;
        DeviceArray<double2> a(n), b(n);
        SBH_HIP(hipMemsetAsync(a.ptr, 0, a.bytes(), ctx->stream));
        for (int r = 0; r < reps; ++r) device_copy(ctx, a.ptr, b.ptr, n);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

sbh_status sbh_debug_device_sincos(sbh_context* ctx, const double* x, int64_t n, double* s, double* c) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && x && s && c && n >= 0, "null argument");
        use_device(ctx);
        DeviceArray<double> dx(n), ds(n), dc(n);
        if (n) copy_on_stream(ctx, dx.ptr, x, sizeof(double) * n, hipMemcpyHostToDevice);
        launch_sincos_probe(ctx, dx.ptr, n, ds.ptr, dc.ptr);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        if (n) {
            copy_on_stream(ctx, s, ds.ptr, sizeof(double) * n, hipMemcpyDeviceToHost);
            copy_on_stream(ctx, c, dc.ptr, sizeof(double) * n, hipMemcpyDeviceToHost);
        }
    });
}

sbh_status sbh_debug_device_rsqrt(sbh_context* ctx, const double* x, int64_t n, double* y) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && x && y && n >= 0, "null argument");
        use_device(ctx);
        DeviceArray<double> dx(n), dy(n);
        if (n) copy_on_stream(ctx, dx.ptr, x, sizeof(double) * n, hipMemcpyHostToDevice);
        launch_rsqrt_probe(ctx, dx.ptr, n, dy.ptr);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        if (n) copy_on_stream(ctx, y, dy.ptr, sizeof(double) * n, hipMemcpyDeviceToHost);
    });
}

}  // extern "C"
