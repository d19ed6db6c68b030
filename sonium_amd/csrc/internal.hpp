// Internal definitions shared by the host sources and HIP kernels of libsbh.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbh.h"

namespace sbh {

struct Error : std::runtime_error {
    sbh_status code;
    Error(sbh_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

// Runs f, turning exceptions into a status code and the thread's last error.
template <typename F>
sbh_status guard(F&& f) {
    try {
        f();
        return SBH_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return SBH_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SBH_ERR_BACKEND;
    }
}

#define SBH_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t sbh_e_ = (expr);                                                           \
        if (sbh_e_ != hipSuccess) {                                                           \
            sbh_status sbh_c_ =                                                               \
                sbh_e_ == hipErrorOutOfMemory ? SBH_ERR_OUT_OF_MEMORY : SBH_ERR_BACKEND;      \
            throw ::sbh::Error(sbh_c_, std::string(#expr) + ": " + hipGetErrorString(sbh_e_)); \
        }                                                                                     \
    } while (0)

// One-shot failure injection (the reference's CudaFailureInjectionPoint,
// crates/sonium-compute/src/backend/mod.rs:120-150): a point armed on a
// context fails the next boundary of that kind reached by an entry point
// running on that context, before the boundary's effect.
enum InjectionPoint {
    INJECT_NONE = 0,
    INJECT_ALLOCATION = 1,
    INJECT_LAUNCH = 2,
    INJECT_TRANSFER_H2D = 3,
    INJECT_TRANSFER_D2H = 4,
    INJECT_SYNCHRONIZATION = 5
};
void injection_check(int point);
// Bytes held by the library's device allocations in this process.
extern std::atomic<int64_t> g_live_device_bytes;

#define SBH_REQUIRE(cond, msg)                                              \
    do {                                                                    \
        if (!(cond)) throw ::sbh::Error(SBH_ERR_INVALID_ARGUMENT, (msg));   \
    } while (0)

// Owning device allocation.
template <typename T>
struct DeviceArray {
    T* ptr = nullptr;
    size_t count = 0;
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { allocate(n); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DeviceArray& operator=(DeviceArray&& o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
        return *this;
    }
    ~DeviceArray() { release(); }
    void allocate(size_t n) {
        release();
        if (n == 0) return;
        injection_check(INJECT_ALLOCATION);
        SBH_HIP(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)));
        count = n;
        g_live_device_bytes += (int64_t)(n * sizeof(T));
    }
    void release() {
        if (ptr) {
            (void)hipFree(ptr);
            g_live_device_bytes -= (int64_t)(count * sizeof(T));
        }
        ptr = nullptr;
        count = 0;
    }
    size_t bytes() const { return count * sizeof(T); }
};

// Owning pinned host allocation.
template <typename T>
struct PinnedArray {
    T* ptr = nullptr;
    size_t count = 0;
    PinnedArray() = default;
    explicit PinnedArray(size_t n) { allocate(n); }
    PinnedArray(const PinnedArray&) = delete;
    PinnedArray& operator=(const PinnedArray&) = delete;
    ~PinnedArray() {
        if (ptr) (void)hipHostFree(ptr);
    }
    void allocate(size_t n) {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        count = 0;
        if (n == 0) return;
        SBH_HIP(hipHostMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T), hipHostMallocDefault));
        count = n;
    }
};

// Per-panel data for the far-field kernel: vertex 0, the two edge vectors
// (exactly v1-v0 and v2-v0, the reference's dx/ds and dx/dt), the
// vertex-order unit normal, and 0.5*J/(4*pi) so that a TR13 weight w gives
// wga/(4*pi) = w * wj.
struct FarPanel {
    double v0[3];
    double e1[3];
    double e2[3];
    double nf[3];
    double wj;
    double pad;
};

// Exact geometry for the near-pair and self-panel kernels, which evaluate in
// the reference's operation order.
struct PanelGeom {
    double c[9];    // node coordinates, row-major (node, xyz)
    double nf[3];   // vertex-order unit normal (compute_parameters)
    double jac;     // |(v1-v0) x (v2-v0)|
    double area;    // element area (compute_element_geometry)
    double size;    // estimate_element_size: max(sqrt(area), 1e-10)
    double pad;
};

// Collocation point and outward element normal of a source panel.
struct Source {
    double c[3];
    double n[3];
};

// Incident source on the device (same layout as sbh_incident_source).
struct IncidentSource {
    int32_t kind;  // 0 plane wave, 1 point source
    int32_t pad;
    double v[3];
    double amp[2];
};
static_assert(sizeof(IncidentSource) == sizeof(sbh_incident_source), "incident source layout");

// Boundary condition of a field panel (GENERAL assembly mode only).
struct PanelBc {
    int32_t type;       // 0 velocity, 1 pressure
    int32_t has_adm;
    double value[2];
    double adm[2];
};

// One near pair: leaves [begin, begin+count) of the leaf-code array.
struct NearPair {
    int32_t i, j;
    int32_t count;
    int32_t pad;
    int64_t begin;
};

// One instance of a batch of independent rigid scatterers.  Panel arrays of
// all instances are concatenated (instance panels at panel_off); its FP64
// matrix is n x n row-major at a_off; its right-hand sides and solutions are
// n x ncol row-major blocks at vec_off of the packed vectors; its complex
// FP16/BF16 copy has rows of ld complex entries at lp_off halves.
struct BatchInstance {
    int64_t panel_off;
    int64_t n;
    int64_t a_off;
    int64_t vec_off;
    int64_t lp_off;
    int64_t ld;
    int64_t dir_off;  // first direction (3 doubles each) in the direction array
    int32_t ncol;
    int32_t col_off;  // first global column
    double k;
    double tau;
    double sign;
    double beta[2];
    int32_t in_family;  // far entries written by the frequency-family kernel
    int32_t pad2;
};

// Instances on one mesh whose wavenumbers are k0 + f dk (f < count): the far
// entries of all of them come from one pass over the pair geometry, with the
// phase exp(i k_f r) advanced from f to f + 1 by one complex product.
constexpr int FAMILY_MAX = 16;
// Family kernels are compiled for 16, 8, 4 and 2 members; a family takes the
// smallest that holds it (class 0..3) and families are stored by class.
constexpr int FAMILY_CLASSES = 4;
inline int family_class(int count) { return count > 8 ? 0 : count > 4 ? 1 : count > 2 ? 2 : 3; }
struct FreqFamily {
    int64_t panel_off;  // the shared geometry (that of the first member)
    int64_t n;
    int64_t pair_begin;  // the first member's near pairs (global pair index)
    int64_t npairs;
    double k0, dk;
    int32_t count;
    int32_t pad;
    int32_t member[FAMILY_MAX];  // instance indices, in order of f
    double kf[FAMILY_MAX];       // the members' own wavenumbers
};

// A subdivision leaf is stored as its path from the root triangle: bits
// [2*l, 2*l+1] hold the child index chosen at level l+1 (0,1,2: corner
// children [v_c, m_c, m_(c-1)], 3: centre [m0, m1, m2]); bits 58..63 hold the
// depth.  Decoding reproduces the reference's midpoint arithmetic exactly.
constexpr int LEAF_DEPTH_SHIFT = 58;

struct KernelTiming {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    int64_t launches = 0;
    double total_ms = 0.0;
};

}  // namespace sbh

// Every object created on a context holds a reference to it, so objects may be
// destroyed in any order relative to their context.
struct sbh_context {
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t stream = nullptr;
    bool profiling = false;
    std::mutex mu;  // kernel-timing records
    // Held by every entry point that enqueues work on `stream` or touches the
    // shared scratch below (and by calls on objects made on this context), so
    // concurrent calls from several host threads run one after another.
    std::recursive_mutex op_mu;
    sbh::KernelTiming timing[SBH_KERNEL_CLASS_COUNT];
    // Kernel-timing events are reused: taken from this pool by begin/end_timed
    // and returned when their times are read, so a timed region records
    // events but creates none once the pool is warm.
    std::vector<hipEvent_t> event_pool;
    hipEvent_t take_event();
    // Device-memory admission (the reference's BemResourceBudget, device
    // side): an optional budget in bytes (0 = none) and the bytes held by this
    // context's long-lived objects (systems, low-precision copies, batches).
    std::atomic<int64_t> device_budget{0};
    std::atomic<int64_t> device_reserved{0};
    std::atomic<int> armed_failure{0};  // sbh::InjectionPoint, one shot
    // device scratch for partial sums of transposed products, grown on demand
    double2* scratch = nullptr;
    size_t scratch_count = 0;
    double2* device_scratch(size_t count);
    // small pinned staging buffer for reductions
    double* pinned = nullptr;
    size_t pinned_count = 0;
    double* pinned_rw(size_t count);
    void begin_timed(int cls, hipEvent_t* start);
    void end_timed(int cls, hipEvent_t start);
};

struct sbh_mesh {
    sbh_context* ctx = nullptr;
    int64_t n = 0;
    int64_t num_nodes = 0;
    std::vector<double> nodes;      // V x 3
    std::vector<int64_t> tris;      // N x 3
    std::vector<double> centers;    // N x 3 (collocation points)
    std::vector<double> normals;    // N x 3 (outward element normals)
    std::vector<double> areas;      // N
    std::vector<sbh::PanelGeom> geom;
    std::vector<sbh::PanelBc> bc;
    double diameter = 0.0;          // bounding-box diagonal of the nodes
    bool general = false;           // any pressure BC, admittance or nonzero BC value
    bool vertex_order_normals = false;  // keep (v1-v0)x(v2-v0) instead of orienting away from the origin
    std::vector<int32_t> nonzero_bc;  // panels whose BC value is nonzero (tbem.rs:1240-1242)
    sbh::DeviceArray<sbh::FarPanel> d_far;
    sbh::DeviceArray<sbh::PanelGeom> d_geom;
    sbh::DeviceArray<sbh::Source> d_src;
    sbh::DeviceArray<sbh::PanelBc> d_bc;
    sbh::DeviceArray<int32_t> d_nonzero_bc;
};

struct sbh_workspace {
    sbh_context* ctx = nullptr;
    const sbh_mesh* mesh = nullptr;
    int64_t near_pairs = 0;
    int64_t leaves = 0;
    double build_ms = 0.0;
    sbh::DeviceArray<sbh::NearPair> d_pairs;
    sbh::DeviceArray<uint64_t> d_leaves;
    sbh::DeviceArray<int64_t> d_row_ptr;  // CSR of near pairs by source row
};

struct sbh_system {
    sbh_context* ctx = nullptr;
    int64_t n = 0;
    int64_t lda = 0;
    int64_t admitted = 0;  // device bytes counted against the context's budget
    sbh::DeviceArray<double2> d_A;
    sbh::DeviceArray<double2> d_rhs;
    // Row maxima of max(|Re|, |Im|) left by the value assembly (n rows of
    // rowmax_stride partials), from which the low-precision copy takes its
    // row scales without reading the matrix.  Valid only right after that
    // assembly; any other write of the matrix clears it, and a matrix whose
    // device address was handed out never uses it.
    sbh::DeviceArray<double> d_rowmax;
    int64_t rowmax_stride = 0;
    bool rowmax_valid = false;
    bool matrix_exposed = false;
    // Live low-precision copies of this system; the assembly writes the row
    // maxima only while there is one to use them.  Each copy also holds a
    // reference, so the system is freed after the last of sbh_system_destroy
    // and its copies' destruction.
    std::atomic<int> lowp_copies{0};
    std::atomic<int> refs{1};
};

struct sbh_vector {
    sbh_context* ctx = nullptr;
    int64_t length = 0;
    sbh::DeviceArray<double2> d;
};

// Complex FP16/BF16 copy of a system matrix with power-of-two row scales,
// plus scratch for its products.
struct sbh_lowp_matrix {
    sbh_context* ctx = nullptr;
    sbh_system* sys = nullptr;
    int64_t n = 0;
    int64_t ld = 0;  // complex entries per row, multiple of 64
    int format = 0;  // 0 FP16, 1 BF16
    int64_t admitted = 0;  // device bytes counted against the context's budget
    sbh::DeviceArray<uint16_t> d_A;       // ceil(n/32)*32 x 2*ld halves, MFMA fragment order
    sbh::DeviceArray<double> d_row_inv;   // n
    sbh::DeviceArray<uint16_t> d_XT;      // 64 x 2*ld halves (real-part block operand)
    sbh::DeviceArray<unsigned> d_colmax;  // 64
    sbh::DeviceArray<double> d_colinv;    // 64
    sbh::DeviceArray<float> d_ypart;      // split-K partial sums, splits x n x 128
};

// A dense operator wraps a device-resident system; an exact operator
// evaluates the same entries on the fly from a workspace (O(N) memory).
struct sbh_operator {
    sbh_context* ctx = nullptr;
    int64_t n = 0;
    sbh_system* sys = nullptr;          // dense
    const sbh_workspace* ws = nullptr;  // exact
    sbh_physics physics{};
    double2 beta{};
    double sign = 1.0;
    sbh::DeviceArray<double2> d_far, d_near, d_self;  // exact-product partials
    sbh_apply_fn callback = nullptr;                  // caller-supplied
    void* user = nullptr;
    int callback_memory = 0;
    bool exact() const { return ws != nullptr; }
    bool dense() const { return sys != nullptr; }
};

struct sbh_precond {
    sbh_context* ctx = nullptr;
    int64_t n = 0;
    int kind = 0;  // 0 identity, 1 inverse diagonal, 2 caller-supplied
    sbh::DeviceArray<double2> d_dinv;
    sbh_apply_fn callback = nullptr;
    void* user = nullptr;
    int callback_memory = 0;
};

// Recycled GCRO-DR subspace: up to k orthonormal vectors of order n, rows of U.
struct sbh_gcro_subspace {
    sbh_context* ctx = nullptr;
    int64_t n = 0;
    int k = 0;
    int count = 0;
    sbh::DeviceArray<double2> U;
};

namespace sbh {

// Copies and fills ordered on the context's stream and waited for before
// returning.  The stream is created non-blocking, so plain hipMemcpy and
// hipMemset on the null stream would not be ordered with its kernels.
inline void copy_on_stream(sbh_context* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (!bytes) return;
    SBH_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    SBH_HIP(hipStreamSynchronize(ctx->stream));
}
inline void zero_on_stream(sbh_context* ctx, void* dst, size_t bytes) {
    if (!bytes) return;
    SBH_HIP(hipMemsetAsync(dst, 0, bytes, ctx->stream));
    SBH_HIP(hipStreamSynchronize(ctx->stream));
}

inline sbh_context* ctx_of(sbh_context* c) { return c; }
inline sbh_context* ctx_of(const sbh_vector* v) { return v ? v->ctx : nullptr; }
inline sbh_context* ctx_of(const sbh_mesh* m) { return m ? m->ctx : nullptr; }
inline sbh_context* ctx_of(const sbh_workspace* w) { return w ? w->ctx : nullptr; }
inline sbh_context* ctx_of(const sbh_system* s) { return s ? s->ctx : nullptr; }
inline sbh_context* ctx_of(const sbh_operator* o) { return o ? o->ctx : nullptr; }
inline sbh_context* ctx_of(const sbh_precond* p) { return p ? p->ctx : nullptr; }
inline sbh_context* ctx_of(const sbh_lowp_matrix* l) { return l ? l->ctx : nullptr; }
inline sbh_context* ctx_of(const sbh_gcro_subspace* g) { return g ? g->ctx : nullptr; }

// guard() with the context's operation lock held and its device current on
// the calling thread.  A null context runs f unlocked, so f reports its own
// null-argument error.
// The context whose entry point the calling thread is running (failure
// injection is attributed to it).
extern thread_local sbh_context* g_current_ctx;

template <typename F>
sbh_status guard_on(sbh_context* ctx, F&& f) {
    if (!ctx) return guard(std::forward<F>(f));
    std::lock_guard<std::recursive_mutex> lock(ctx->op_mu);
    struct Current {
        sbh_context* prev;
        explicit Current(sbh_context* c) : prev(g_current_ctx) { g_current_ctx = c; }
        ~Current() { g_current_ctx = prev; }
    } current(ctx);
    return guard([&] {
        SBH_HIP(hipSetDevice(ctx->device));
        f();
    });
}

// Admission before a device allocation of `bytes` (BemResourceRequest
// admission, crates/sonium-bem/src/core/admission.rs): refused with
// SBH_ERR_OUT_OF_MEMORY when it would exceed the context's budget (counting
// the bytes its long-lived objects hold) or the device's free memory, before
// anything is allocated.  With `reserve` the bytes are added to the
// context's held bytes (release them with release_device).
void admit_device(sbh_context* ctx, int64_t bytes, const char* what, bool reserve);
void release_device(sbh_context* ctx, int64_t bytes);

void context_retain(sbh_context* ctx);
void context_release(sbh_context* ctx);

// Host route builder (mesh_routes.cpp).
void build_mesh_geometry(sbh_mesh* mesh);
struct RouteTable {
    std::vector<NearPair> pairs;
    std::vector<uint64_t> leaves;
    std::vector<int64_t> row_ptr;
};
RouteTable build_routes(const sbh_mesh* mesh, int threads);
int subelements_for_pair(const double* src, const double* coords, double area, double* leaves_out);
void gauss_legendre_rule(int order, double* x, double* w);

// Assembly launchers (tbem_kernels.hip).
void upload_quadrature_tables(sbh_context* ctx);
void launch_assembly(sbh_workspace* ws, const sbh_physics& ph, sbh_c64 beta, sbh_system* sys);
// y = sign * A x without storing A (tbem_kernels.hip); x and y on the device.
void launch_exact_apply(const sbh_operator* op, const double2* x, double2* y);
// d/dp of the stabilized system into sys (tbem_kernels.hip): dk/dp =
// wavenumber_scale, d(pressure factor)/dp = pf_derivative, dY/dp =
// admittance_scale * Y; beta_d is the total derivative of beta.
void launch_derivative_assembly(sbh_workspace* ws, const sbh_physics& ph, sbh_c64 beta, sbh_c64 beta_d,
                                double sign, double wavenumber_scale, double pf_derivative, double admittance_scale,
                                sbh_system* sys);
// Pressure at points (tbem_kernels.hip): plane waves plus the scattered field
// of the panels' Cauchy data; v == nullptr derives it from the BCs.
void launch_field(sbh_context* ctx, const sbh_mesh* m, const sbh_physics& ph, const double2* sol, const double2* v,
                  const double* d_points, int64_t npts, const IncidentSource* d_sources, int64_t nsources,
                  double2* scattered, double2* incident, double2* total);
// rhs (+)= sign * -(gamma p_inc + beta tau dp_inc/dn) of the summed sources.
void launch_source_rhs(sbh_context* ctx, const sbh_mesh* m, const sbh_physics& ph, sbh_c64 beta, double sign,
                       const IncidentSource* d_sources, int64_t nsources, double2* rhs, bool accumulate);
// y = A x for a dense, exact or caller-supplied operator (api.cpp).
void operator_apply_device(const sbh_operator* op, const double2* x, double2* y);
// y = M^-1 r through a caller-supplied callback (api.cpp), device vectors.
void callback_apply_device(sbh_context* ctx, sbh_apply_fn fn, void* user, int memory, const double2* x, double2* y,
                           int64_t n, const char* what);
void launch_incident_rhs(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics& ph,
                         sbh_c64 beta, double sign, const double* d_dirs, const double2* d_amps,
                         int64_t num_rhs, double2* d_rhs, bool accumulate);

// Persistent assembly of a batch (tbem_kernels.hip): items are (instance,
// first row) pairs; each writes the far entries, near pairs, diagonal and
// plane-wave right-hand sides of rows_per_item rows.  counter holds 2 words.
struct BatchAssemblyArgs {
    const FreqFamily* families;
    int nfamilies;
    int64_t family_max_n;
    int64_t family_max_pairs;
    int family_class_end[FAMILY_CLASSES];  // families [end[c-1], end[c]) are of class c
    const BatchInstance* inst;
    const int2* items;
    int nitems;
    int rows_per_item;
    unsigned* counter;
    const FarPanel* far;
    const Source* src;
    const PanelGeom* geom;
    const NearPair* pairs;
    const int64_t* row_ptr;  // global rows -> global pair index
    const uint64_t* leaves;
    const double* dirs;
    double2* A;
    double2* rhs;
};
void launch_batch_assembly(sbh_context* ctx, const BatchAssemblyArgs& args, int grid);

// Dense BLAS launchers (blas_kernels.hip).
void zgemv_n(sbh_context* ctx, const double2* A, int64_t lda, int64_t rows, int64_t cols,
             const double2* x, double2* y);
void zgemv_t(sbh_context* ctx, const double2* A, int64_t lda, int64_t rows, int64_t cols,
             const double2* x, double2* y, bool conjugate);
void zgemm_block(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X,
                 double2* Y, int64_t k);
// Roofline ceiling probes (GB/s of a streaming read; FP64 FMA TFLOP/s).
double probe_stream_read(sbh_context* ctx, const void* x, int64_t bytes, int reps, unsigned* sink);
double probe_fma_rate(sbh_context* ctx, double* sink);
// Partial sums for dot products <v_i, w> (i < nvec) and ||w||^2 in one launch;
// out holds (nvec + 1) x chunks complex partials (the last row is the norm).
int dots_and_norm_partials(sbh_context* ctx, const double2* V, int64_t ldv, int nvec,
                           const double2* w, int64_t n, double2* out);
// s[v] = sum of the chunk partials of slot v (v < nv), in chunk order, on the device.
void sum_chunk_partials(sbh_context* ctx, const double2* part, int nv, int chunks, double2* s);
// w -= sum_i c_i v_i; writes ||w||^2 partials into norm_out (chunks).
int project_and_norm(sbh_context* ctx, const double2* V, int64_t ldv, int nvec, const double2* c,
                     double2* w, int64_t n, double* norm_out);
// x += sum_i c_i v_i
void basis_update(sbh_context* ctx, const double2* V, int64_t ldv, int nvec, const double2* c,
                  double2* x, int64_t n);
// y = alpha * x (alpha real)
void scale_copy(sbh_context* ctx, const double2* x, double alpha, double2* y, int64_t n);
// r = b - r ; norm partials of r into norm_out
int residual_and_norm(sbh_context* ctx, const double2* b, double2* r, int64_t n, double* norm_out);
// y = d .* x (elementwise complex)
void diag_scale(sbh_context* ctx, const double2* d, const double2* x, double2* y, int64_t n);
// d = 1 / diag(A)
void extract_inverse_diagonal(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, double2* d);
// dst = src (n complex values)
void device_copy(sbh_context* ctx, const double2* src, double2* dst, int64_t n);
// partials of ||x||^2
int norm_partials(sbh_context* ctx, const double2* x, int64_t n, double* out);
// Allocation-free block product (lowp_kernels.hip): the split-K partials sit
// in a fixed device array of NOALLOC_SCRATCH_DOUBLES doubles (64 MiB) per GPU,
// whose address noalloc_scratch_init takes once per device.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)8 << 20;
void noalloc_scratch_init(int device);
size_t noalloc_block_scratch_need(int64_t n, int64_t k);
void zgemm_f64_noalloc(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, const double2* X, int64_t k,
                       double2* Y);
// y = A x without kernel-timing events (blas_kernels.hip)
void zgemv_n_untimed(sbh_context* ctx, const double2* A, int64_t lda, int64_t rows, int64_t cols, const double2* x,
                     double2* y);
// Low-precision conversion (lowp_kernels.hip).  With rowmax non-null the row
// scales come from the assembly's row maxima (n rows of stride partials)
// instead of a pass over A.
void lowp_convert(sbh_context* ctx, const double2* A, int64_t lda, int64_t n, int64_t ld, int format,
                  uint16_t* out, double* row_inv, const double* rowmax, int64_t stride);
// The row maxima of sys's matrix when the last value assembly left them valid.
inline const double* assembled_row_max(const sbh_system* sys) {
    return sys->rowmax_valid && !sys->matrix_exposed ? sys->d_rowmax.ptr : nullptr;
}
int64_t lowp_padded_ld(int64_t n);
int64_t lowp_split_count();
int reduction_chunks(int64_t n);

}  // namespace sbh
