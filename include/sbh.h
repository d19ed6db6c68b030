/*
 * sbh.h — C ABI of sonium_amd: the dense Burton–Miller Helmholtz BEM path
 * (constant-panel collocation, Tri3) on one AMD Instinct MI355X (gfx950).
 *
 * Every entry point is a drop-in for one reference interface; each
 * declaration cites the reference item it replaces (paths relative to the
 * reference repository root).  Conventions:
 *   - complex numbers are interleaved {re, im} doubles (sbh_c64), the layout
 *     of num_complex::Complex64 (#[repr(C)]);
 *   - matrices are row-major; a block of k right-hand sides is an n x k
 *     row-major array, one right-hand side per column
 *     (crates/sonium-solvers/crates/sonium-solvers-core/src/traits.rs:2989-3000);
 *   - every function returns an sbh_status; on failure sbh_last_error()
 *     returns a message for the calling thread.  No function aborts;
 *   - the caller owns host arrays; the library owns device objects, which are
 *     released by the matching *_destroy call;
 *   - pointers tagged SBH_MEM_DEVICE are HIP device pointers on the
 *     context's device, SBH_MEM_HOST ordinary host memory.
 */
#ifndef SBH_H
#define SBH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double re, im;
} sbh_c64;

typedef enum {
    SBH_OK = 0,
    SBH_ERR_INVALID_ARGUMENT = 1, /* CudaTbemAssemblyError::InvalidInput, tbem.rs:3693-3718 */
    SBH_ERR_UNSUPPORTED = 2,      /* CudaTbemAssemblyError::UnsupportedElement */
    SBH_ERR_BACKEND = 3,          /* CudaTbemAssemblyError::Backend / BackendError, sonium-compute/src/error.rs:42-102 */
    SBH_ERR_OUT_OF_MEMORY = 4,
    SBH_ERR_NUMERICAL = 5         /* SolverStatus::NumericalError, gmres.rs:252-275 */
} sbh_status;

typedef enum { SBH_MEM_HOST = 0, SBH_MEM_DEVICE = 1 } sbh_memory;

typedef struct sbh_context sbh_context;
typedef struct sbh_mesh sbh_mesh;
typedef struct sbh_workspace sbh_workspace;
typedef struct sbh_system sbh_system;
typedef struct sbh_operator sbh_operator;
typedef struct sbh_precond sbh_precond;
typedef struct sbh_vector sbh_vector;

/* Thread-local message of the last failed call on this thread. */
const char* sbh_last_error(void);
/* Library version string. */
const char* sbh_version(void);

/* ---------------------------------------------------------------------------
 * Physics parameters — mirrors PhysicsParams
 * (crates/sonium-bem/src/core/types.rs:190-218, 306-487) under
 * ACOUSTIC_CONVENTION (crates/sonium-bem/src/core/conventions.rs:308-313). */
typedef struct {
    double frequency;
    double speed_of_sound;
    double density;
    double wave_number;
    double omega;
    double wave_length;
    double harmonic_factor; /* spatial sign, +1: G = exp(ikr)/(4 pi r) */
    double pressure_factor; /* density * omega * harmonic_factor */
    double tau;             /* +1 exterior, -1 interior */
} sbh_physics;

/* PhysicsParams::new_with_convention (types.rs:190-218). */
sbh_status sbh_physics_new(double frequency, double speed_of_sound, double density, int is_internal,
                           sbh_physics* out);
/* Same parameters from an exact wavenumber (frequency = k c / 2 pi); lets a
 * caller fix k bit-for-bit, as the reference's tests do by choosing f. */
sbh_status sbh_physics_from_wavenumber(double wave_number, double speed_of_sound, double density,
                                       int is_internal, sbh_physics* out);
/* PhysicsParams::burton_miller_beta_scaled (types.rs:418-424); scale 1 is
 * burton_miller_beta (types.rs:314-320). */
sbh_c64 sbh_burton_miller_beta_scaled(const sbh_physics* physics, double scale);

/* ---------------------------------------------------------------------------
 * Context: one HIP device, one stream, optional kernel timing. */
sbh_status sbh_context_create(int device, sbh_context** out);
void sbh_context_destroy(sbh_context* ctx);
sbh_status sbh_context_synchronize(sbh_context* ctx);
/* Kernel timing with HIP events on the context stream (off by default). */
sbh_status sbh_context_set_profiling(sbh_context* ctx, int enabled);
typedef enum {
    SBH_KERNEL_ASSEMBLE_FAR = 0,
    SBH_KERNEL_ASSEMBLE_NEAR = 1,
    SBH_KERNEL_ASSEMBLE_SELF = 2,
    SBH_KERNEL_RHS = 3,
    SBH_KERNEL_MATVEC = 4,
    SBH_KERNEL_KRYLOV_VECTOR = 5,
    SBH_KERNEL_BLOCK = 6,
    SBH_KERNEL_LOWP = 7,
    SBH_KERNEL_FIELD = 8,
    SBH_KERNEL_CLASS_COUNT = 9
} sbh_kernel_class;
/* Synchronizes, then reports launches and summed event time of one class
 * since the last reset. */
sbh_status sbh_context_kernel_stats(sbh_context* ctx, int kernel_class, int64_t* launches,
                                    double* total_ms);
sbh_status sbh_context_reset_stats(sbh_context* ctx);

/* Device-resident complex vector (or n x k block) owned by the library, so
 * callers can keep right-hand sides and solutions in HBM between calls. */
sbh_status sbh_vector_create(sbh_context* ctx, int64_t length, sbh_vector** out);
void sbh_vector_destroy(sbh_vector* v);
int64_t sbh_vector_length(const sbh_vector* v);
sbh_c64* sbh_vector_data(sbh_vector* v); /* device pointer */
sbh_status sbh_vector_upload(sbh_vector* v, const sbh_c64* host, int64_t offset, int64_t count);
sbh_status sbh_vector_download(const sbh_vector* v, sbh_c64* host, int64_t offset, int64_t count);

/* ---------------------------------------------------------------------------
 * Mesh: nodes + Tri3 elements + boundary conditions, host description of
 * `elements: &[Element], nodes: &[[f64;3]]` (types.rs:582-719).  Element
 * centre, area and outward normal are derived as compute_element_geometry
 * does (crates/sonium-bem/src/core/mesh/generators.rs:779-870).  DOF numbering
 * is the element order (dof_addresses = [i], as prepare_elements assigns for a
 * rigid body, bem_solver.rs:4564-4570).
 *   bc_type[i]   0 = Velocity, 1 = Pressure (get_bc_type_and_value, assembly/common.rs:273-307); NULL = all 0
 *   bc_value[i]  the panel's (average) BC value; NULL = all zero (rigid)
 *   admittance[i], has_admittance[i]  optional Robin admittance; NULL = none */
sbh_status sbh_mesh_create(sbh_context* ctx, const double* nodes, int64_t num_nodes,
                           const int64_t* triangles, int64_t num_elements, const int32_t* bc_type,
                           const sbh_c64* bc_value, const sbh_c64* admittance,
                           const uint8_t* has_admittance, sbh_mesh** out);
/* Normal rule of sbh_mesh_create_oriented: SBH_NORMALS_OUTWARD is the
 * generators' compute_element_geometry (vertex-order normal flipped to point
 * away from the origin, generators.rs:857-866), the rule sbh_mesh_create uses;
 * SBH_NORMALS_VERTEX_ORDER keeps (v1-v0)x(v2-v0) as compute_element_normal
 * (crates/sonium-bem/src/core/mesh/element.rs:368-392) does for meshes read
 * from files (load_project, crates/sonium-bem/src/numcalc/project.rs:78-105). */
typedef enum { SBH_NORMALS_OUTWARD = 0, SBH_NORMALS_VERTEX_ORDER = 1 } sbh_normal_rule;
sbh_status sbh_mesh_create_oriented(sbh_context* ctx, const double* nodes, int64_t num_nodes,
                                    const int64_t* triangles, int64_t num_elements, const int32_t* bc_type,
                                    const sbh_c64* bc_value, const sbh_c64* admittance,
                                    const uint8_t* has_admittance, int normal_rule, sbh_mesh** out);
void sbh_mesh_destroy(sbh_mesh* mesh);
int64_t sbh_mesh_num_elements(const sbh_mesh* mesh);
/* Host copies of centres, outward normals (N x 3) and areas (N); any may be NULL. */
sbh_status sbh_mesh_geometry(const sbh_mesh* mesh, double* centers, double* normals, double* areas);

/* ---------------------------------------------------------------------------
 * Resident assembly workspace — replaces CudaTbemAssemblyWorkspace::new
 * (crates/sonium-bem/src/core/assembly/tbem.rs:3779-4022): uploads the
 * geometry once and builds the frequency-invariant routes (near pairs and
 * their generate_subelements trees, singular.rs:1372-1528).  route_threads
 * host threads build the routes (0 = hardware concurrency). */
sbh_status sbh_workspace_create(sbh_context* ctx, const sbh_mesh* mesh, int route_threads,
                                sbh_workspace** out);
void sbh_workspace_destroy(sbh_workspace* ws);
/* Route statistics: near pairs, subdivision leaves, host build time. */
sbh_status sbh_workspace_stats(const sbh_workspace* ws, int64_t* near_pairs, int64_t* leaves,
                               double* build_ms);

/* Dense system on the device: row-major N x N matrix + N right-hand side
 * (TbemSystem, tbem.rs:134-141). */
sbh_status sbh_system_create(sbh_context* ctx, int64_t num_dofs, sbh_system** out);
void sbh_system_destroy(sbh_system* sys);
int64_t sbh_system_num_dofs(const sbh_system* sys);
/* Device pointers of the matrix (leading dimension in *lda, in elements) and rhs
 * (each output may be NULL). */
sbh_status sbh_system_device_pointers(sbh_system* sys, sbh_c64** matrix, int64_t* lda,
                                      sbh_c64** rhs);
/* Copies rows [row_begin, row_end) of the matrix (dense, N per row) and the
 * rhs to host memory; either destination may be NULL. */
sbh_status sbh_system_download(sbh_system* sys, int64_t row_begin, int64_t row_end,
                               sbh_c64* matrix_rows, sbh_c64* rhs);
/* Overwrites the rhs from host or device memory. */
sbh_status sbh_system_set_rhs(sbh_system* sys, const sbh_c64* rhs, int memory);

/* CudaTbemAssemblyWorkspace::assemble (tbem.rs:4230-4389) /
 * build_tbem_system_with_beta (tbem.rs:1196-1203 → 1284-1466): assembles
 * A and the BC part of b for one wavenumber into sys.
 *   A_ij = H_ij*gamma*tau + E_ij*beta (field BC 0), -(G_ij*gamma*tau + Ht_ij*beta) (BC 1),
 *   A_ii += -gamma/2 (BC 0) or +beta*tau/2 (BC 1), plus admittance terms. */
sbh_status sbh_workspace_assemble(sbh_workspace* ws, const sbh_physics* physics, sbh_c64 beta,
                                  sbh_system* sys);
/* d/dp of the stabilized system at fixed geometry —
 * build_tbem_system_stabilized_parameter_derivative (tbem.rs:1858-2019) and
 * CudaTbemAssemblyWorkspace::assemble_wavenumber_derivative (tbem.rs:4399-4591):
 * for a parameter p with dk/dp = wavenumber_scale, d(rho omega hf)/dp =
 * pressure_factor_derivative and dY/dp = admittance_scale * Y on admittance
 * panels, writes into sys
 *   dA = sign * [(dH gamma tau + dE beta) ws + E beta'] (BC 0, minus the
 *        admittance terms), -sign * [(dG gamma tau + dHt beta) ws + Ht beta'] (BC 1),
 *   plus the derivative free terms, and the derivative of the BC right-hand side.
 * beta is the coupling of the value system and beta_derivative its total
 * derivative (the reference's burton_miller_beta_wavenumber_derivative() *
 * beta_scale * wavenumber_scale).  The routes and self-panel tiers are those of
 * the value assembly at physics (held fixed, as in the reference).
 * wavenumber: ws = 1, pf' = density*c*hf (tbem.rs:1734-1752); sound speed:
 * ws = -k/c, pf' = 0 (:1754-1782); density: ws = 0, pf' = pf/density
 * (:1784-1805); impedance Y = 1/(rho c): admittance_scale = -1/density or
 * -1/c (:1807-1856). */
sbh_status sbh_workspace_assemble_derivative(sbh_workspace* ws, const sbh_physics* physics, sbh_c64 beta,
                                             sbh_c64 beta_derivative, double sign, double wavenumber_scale,
                                             double pressure_factor_derivative, double admittance_scale,
                                             sbh_system* sys);
/* build_tbem_system (tbem.rs:877-883): beta = burton_miller_beta(). */
sbh_status sbh_build_tbem_system(sbh_context* ctx, const sbh_mesh* mesh,
                                 const sbh_physics* physics, sbh_system* sys);

/* Plane-wave incident right-hand side:
 * IncidentField::compute_rhs_with_beta (crates/sonium-bem/src/core/incident.rs:807-834)
 * added with the formulation sign as add_incident_field_rhs_stabilized does
 * (bem_solver.rs:5005-5034):  rhs_col += sign * -(gamma p_inc + beta tau dp_inc/dn).
 * directions: num_rhs x 3 unit vectors; amplitudes: num_rhs complex.
 * rhs: N x num_rhs row-major (column c = direction c), device or host.
 * accumulate = 0 overwrites instead of adding. */
sbh_status sbh_incident_plane_wave_rhs(sbh_context* ctx, const sbh_mesh* mesh,
                                       const sbh_physics* physics, sbh_c64 beta, double sign,
                                       const double* directions, const sbh_c64* amplitudes,
                                       int64_t num_rhs, sbh_c64* rhs, int memory, int accumulate);

/* Pressure at points off the surface — compute_total_field(_checked) and
 * compute_scattered_field(_checked)
 * (crates/sonium-bem/src/core/postprocess/pressure.rs:250-280, 501-557, 868-894,
 * 2105-2145; integration 1388-1480, 1566-1620, 1987-2044) for constant Tri3
 * panels with the 7-point rule:
 *   scattered[i] = sum_j sum_q [p_j dG/dn_y - v_j i pf G] J_j w_q  at x_i,
 *   incident[i]  = sum_w amplitudes[w] exp(i hf k directions[w].x_i),
 *   total[i]     = incident[i] + scattered[i]   (FieldPoint, pressure.rs:32-41).
 * solution: the N solved unknowns.  surface_velocity: N values v_j (then
 * p_j = solution[j]), or NULL to derive (p_j, v_j) from the panel BCs as
 * constant_cauchy_data (pressure.rs:1481-1520) does, using the plane waves as
 * the incident field on pressure panels.  directions: num_waves x 3
 * (normalized as IncidentField::plane_wave), amplitudes: num_waves, host
 * memory; num_waves = 0 is the reference's "no incident field".
 * solution, surface_velocity, points (num_points x 3) and the outputs use
 * `memory`; incident and total may be NULL.  Non-finite points give
 * SBH_ERR_INVALID_ARGUMENT and a non-finite result SBH_ERR_NUMERICAL
 * (validate_field_inputs / validate_finite_results, pressure.rs:612-750). */
sbh_status sbh_field_pressure(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* physics,
                              const sbh_c64* solution, const sbh_c64* surface_velocity,
                              const double* directions, const sbh_c64* amplitudes, int64_t num_waves,
                              const double* points, int64_t num_points, sbh_c64* scattered,
                              sbh_c64* incident, sbh_c64* total, int memory);

/* One incident source: IncidentField::PlaneWave {direction, amplitude} or
 * IncidentField::PointSource {position, strength}
 * (crates/sonium-bem/src/core/incident.rs:17-39, 216-420):
 *   plane wave   p = A exp(i hf k d.x),            dp/dn = i hf k (d.n) p
 *   point source p = S exp(i hf k r) / (4 pi r),     dp/dn = (i hf k - 1/r) p (x-s).n / r
 * with r = |x - s|; a point within 1e-10 of the source gets 0 (as the
 * reference skips it).  Plane-wave directions are used as given here. */
typedef enum { SBH_SOURCE_PLANE_WAVE = 0, SBH_SOURCE_POINT = 1 } sbh_source_kind;
typedef struct {
    int32_t kind;
    int32_t pad;
    double vec[3];     /* direction (plane wave) or position (point source) */
    sbh_c64 amplitude; /* amplitude or strength */
} sbh_incident_source;
/* IncidentField::compute_rhs_with_beta for the sum of the sources
 * (MultiplePlaneWaves / MultiplePointSources, incident.rs:807-834):
 * rhs += sign * -(gamma p_inc + beta tau dp_inc/dn) at the panel centres, one
 * column of N values (device or host), accumulate = 0 overwrites. */
sbh_status sbh_incident_rhs(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* physics, sbh_c64 beta,
                            double sign, const sbh_incident_source* sources, int64_t num_sources, sbh_c64* rhs,
                            int memory, int accumulate);
/* sbh_field_pressure with a general incident field: the sources are summed
 * for the incident pressure and for the Cauchy data of pressure panels. */
sbh_status sbh_field_pressure_sources(sbh_context* ctx, const sbh_mesh* mesh, const sbh_physics* physics,
                                      const sbh_c64* solution, const sbh_c64* surface_velocity,
                                      const sbh_incident_source* sources, int64_t num_sources,
                                      const double* points, int64_t num_points, sbh_c64* scattered,
                                      sbh_c64* incident, sbh_c64* total, int memory);

/* ---------------------------------------------------------------------------
 * LinearOperator<Complex64> over a device-resident dense matrix —
 * DenseMatrixOperator (bem_solver.rs:6548-6593) / DenseOperator
 * (crates/sonium-bem/src/core/solver/fmm_interface.rs:26-387) with the trait
 * methods of traits.rs:2949-3290.  The operator borrows sys; sys must outlive it. */
typedef enum { SBH_ADJOINT_UNKNOWN = 0, SBH_ADJOINT_UNSUPPORTED = 1, SBH_ADJOINT_EXACT = 2 } sbh_adjoint_capability;
sbh_status sbh_dense_operator_create(sbh_system* sys, sbh_operator** out);
/* Matrix-free exact operator — apply_tbem_stabilized_exact
 * (crates/sonium-bem/src/core/assembly/hmatrix.rs:2363-2425): y = sign * A x
 * with A the matrix sbh_workspace_assemble(ws, physics, beta) would store,
 * evaluated on the fly in O(N) device memory (each product costs about one
 * assembly).  Usable with sbh_operator_apply(_block) and sbh_gmres; the
 * transpose/Hermitian products, the Jacobi preconditioner, block GMRES and
 * GMRES-IR need a dense operator and return SBH_ERR_UNSUPPORTED.  The
 * workspace (and its mesh) must outlive the operator. */
sbh_status sbh_exact_operator_create(sbh_workspace* workspace, const sbh_physics* physics, sbh_c64 beta,
                                     double sign, sbh_operator** out);
void sbh_operator_destroy(sbh_operator* op);
int64_t sbh_operator_num_rows(const sbh_operator* op);
int64_t sbh_operator_num_cols(const sbh_operator* op);
/* transpose_capability / hermitian_capability (traits.rs:3128-3137). */
int sbh_operator_transpose_capability(const sbh_operator* op);
int sbh_operator_hermitian_capability(const sbh_operator* op);
/* apply_into (traits.rs:2987): y = A x.  Checked like try_apply
 * (traits.rs:2964-2986): non-finite input is rejected with
 * SBH_ERR_NUMERICAL before any device work when x is host memory. */
sbh_status sbh_operator_apply(sbh_operator* op, const sbh_c64* x, sbh_c64* y, int memory);
/* apply_transpose_into (traits.rs:3145): y = A^T x. */
sbh_status sbh_operator_apply_transpose(sbh_operator* op, const sbh_c64* x, sbh_c64* y, int memory);
/* apply_hermitian (traits.rs:3174): y = A^H x. */
sbh_status sbh_operator_apply_hermitian(sbh_operator* op, const sbh_c64* x, sbh_c64* y, int memory);
/* apply_block (traits.rs:3000-3017; DenseMatrixOperator::apply_block →
 * DenseMatrix::matmul_into, dense.rs:450-488): Y = A X with X, Y n x k row-major. */
sbh_status sbh_operator_apply_block(sbh_operator* op, const sbh_c64* X, sbh_c64* Y, int64_t k,
                                    int memory);

/* Allocation-free variant of sbh_operator_apply_block.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Dense operator only; X and Y are device pointers (n x k row-major).  The
 * split-K partial sums of the block product live in a fixed 64 MiB device
 * array per GPU instead of the context's growable scratch, so the call makes
 * no hipMalloc/hipFree: it needs 8 * n * 16 * t doubles with t = 1, 2, 4, 8
 * for min(k, 64) up to 8, 16, 32, 64 columns (n <= 8192 at 64 columns), and
 * larger shapes are rejected with SBH_ERR_INVALID_ARGUMENT.  k = 1 uses the
 * matrix-vector kernel (no scratch).  Calls that overlap on one GPU (several
 * contexts or threads on one device) race on the shared array. */
sbh_status sbh_operator_apply_block_noalloc(sbh_operator* op, const sbh_c64* X, sbh_c64* Y, int64_t k);

/* ---------------------------------------------------------------------------
 * Failure injection — CudaFailureInjectionPoint (crates/sonium-compute/src/
 * backend/mod.rs:120-150) and live_allocation_bytes, for acceptance tests of
 * the failure paths (the reference's tbem.rs:6740-6794).  Inert until armed:
 * an armed point fails, once, the next boundary of its kind reached by an
 * entry point running on that context, before the boundary takes effect
 * (allocation: before hipMalloc, SBH_ERR_OUT_OF_MEMORY; launch: before the
 * assembly or operator kernels, transfers: before a host copy moves data,
 * synchronization: before the stream wait; SBH_ERR_BACKEND).  The call
 * returns its error, releases what it had allocated, and publishes nothing. */
typedef enum {
    SBH_INJECT_NONE = 0,
    SBH_INJECT_ALLOCATION = 1,
    SBH_INJECT_LAUNCH = 2,
    SBH_INJECT_TRANSFER_H2D = 3,
    SBH_INJECT_TRANSFER_D2H = 4,
    SBH_INJECT_SYNCHRONIZATION = 5
} sbh_injection_point;
sbh_status sbh_context_inject_failure(sbh_context* ctx, int point);
/* Bytes held by the library's device allocations in this process. */
int64_t sbh_live_device_bytes(void);

/* ---------------------------------------------------------------------------
 * Resource admission — BemResourceBudget / BemResourceRequest::estimate /
 * admit (crates/sonium-bem/src/core/admission.rs:17-330) for the dense
 * (Tbem) method: a decision before anything is allocated.  host_bytes and its
 * parts follow the reference's conservative estimate (dense operator n*n*16,
 * right-hand sides n*rhs*16 per frequency, scratch n*(4*16 + 8)); device_bytes
 * is this library's allocation for the same work (system, right-hand sides
 * and solutions, GMRES(restart) basis; the shared block basis for rhs > 1).
 * Errors: a zero request -> SBH_ERR_INVALID_ARGUMENT (InvalidRequest);
 * overflow -> SBH_ERR_INVALID_ARGUMENT (ArithmeticOverflow); over a limit
 * -> SBH_ERR_OUT_OF_MEMORY (LimitExceeded, the message names the resource);
 * a zero host or work budget -> SBH_ERR_INVALID_ARGUMENT (InvalidBudget).
 * Host-only: no device is touched. */
typedef struct {
    int64_t max_host_bytes;
    int64_t max_device_bytes;
    int64_t max_active_work;
} sbh_resource_budget;

typedef struct {
    int64_t host_bytes;
    int64_t device_bytes;
    int64_t active_work;
    int64_t operator_bytes;
    int64_t rhs_bytes;
    int64_t scratch_bytes;
} sbh_resource_estimate;

sbh_status sbh_admission_estimate(int64_t active_dofs, int64_t rhs_count, int64_t frequency_count, int64_t restart,
                                  sbh_resource_estimate* out);
sbh_status sbh_admission_admit(const sbh_resource_budget* budget, int64_t active_dofs, int64_t rhs_count,
                               int64_t frequency_count, int64_t restart, sbh_resource_estimate* out);

/* Device-side admission on a context: with a budget > 0 every large device
 * allocation (systems, low-precision copies, batch matrices, Krylov
 * workspaces) is refused with SBH_ERR_OUT_OF_MEMORY before it is made when the
 * bytes held by the context's objects plus the request exceed it; requests
 * beyond the device's free memory are refused the same way with or without a
 * budget.  0 removes the budget. */
sbh_status sbh_context_set_device_budget(sbh_context* ctx, int64_t max_device_bytes);
sbh_status sbh_context_device_budget(sbh_context* ctx, int64_t* max_device_bytes, int64_t* held_bytes);

/* A caller-supplied LinearOperator<Complex64> (traits.rs:2949-3290: `gmres`
 * is generic over any operator, gmres.rs:286).  apply computes y = A x for
 * vectors of length n in `memory` (SBH_MEM_HOST: the library stages them
 * through host buffers; SBH_MEM_DEVICE: device pointers, and the product
 * must be complete in device memory when apply returns).  The library's
 * stream is synchronized before each call.  A nonzero return fails the
 * calling entry point with SBH_ERR_BACKEND.  The operator works with
 * sbh_operator_apply(_block), sbh_gmres and sbh_block_gmres (one apply per
 * column); its transpose and Hermitian products are Unsupported. */
typedef int (*sbh_apply_fn)(void* user, const sbh_c64* x, sbh_c64* y, int64_t n, int memory);
sbh_status sbh_operator_create_callback(sbh_context* ctx, int64_t n, sbh_apply_fn apply, void* user, int memory,
                                        sbh_operator** out);

/* ---------------------------------------------------------------------------
 * Preconditioner<Complex64> (traits.rs:3645-3780): identity
 * (IdentityPreconditioner) and the inverse diagonal of a dense operator. */
sbh_status sbh_precond_identity(sbh_context* ctx, int64_t n, sbh_precond** out);
sbh_status sbh_precond_jacobi(sbh_operator* op, sbh_precond** out);
/* A caller-supplied Preconditioner<Complex64> (traits.rs:3645-3780): y = M^-1 r,
 * with the calling convention of sbh_operator_create_callback.  Accepted by
 * sbh_gmres and sbh_precond_apply (sbh_block_gmres takes identity or Jacobi). */
sbh_status sbh_precond_create_callback(sbh_context* ctx, int64_t n, sbh_apply_fn apply, void* user, int memory,
                                       sbh_precond** out);
void sbh_precond_destroy(sbh_precond* p);
sbh_status sbh_precond_apply(sbh_precond* p, const sbh_c64* r, sbh_c64* y, int memory);

/* ---------------------------------------------------------------------------
 * Restarted GMRES — GmresConfig / GmresSolution
 * (crates/sonium-solvers/crates/sonium-solvers-iterative/src/gmres.rs:26-130),
 * gmres / gmres_with_guess / gmres_preconditioned(_with_guess)
 * (gmres.rs:286-297, 564-626) → gmres_core (:631-1057) +
 * finalize_gmres_solution (:159-250).  Krylov basis and vectors live on the
 * device; the Hessenberg/Givens recurrence runs on the host in FP64. */
typedef struct {
    int64_t max_iterations; /* restart cycles */
    int64_t restart;
    double tolerance;
    /* Convergence is tested after inner step j when j % max(1, interval) == 0,
     * at the last step of a cycle and at breakdown (gmres.rs:871-876); 0 and 1
     * test every step (the default). */
    int64_t convergence_check_interval;
    /* exact_residual_check_interval (gmres.rs:37-42): every converged return is
     * confirmed with the recomputed residual whatever its value
     * (exact_residual_check_due, gmres_common.rs:8-14), so it changes no
     * result; negative values are rejected. */
    int64_t exact_residual_check_interval;
    /* Adaptive restart (min_restart/max_restart, gmres.rs:43-46, 1013-1029):
     * with both > 0, a full cycle whose projected residual fell by less than
     * 10 % doubles the restart (up to max_restart) and one that fell below a
     * tenth halves it (down to min_restart); 0 = None (fixed restart). */
    int64_t min_restart;
    int64_t max_restart;
} sbh_gmres_config;

typedef enum {
    SBH_SOLVER_CONVERGED = 0,
    SBH_SOLVER_MAX_ITERATIONS = 1,
    SBH_SOLVER_BREAKDOWN = 2,
    SBH_SOLVER_NUMERICAL_ERROR = 3
} sbh_solver_status;

typedef struct {
    int64_t iterations; /* total Arnoldi steps (matrix-vector products in the cycles) */
    int64_t restarts;
    double residual;           /* recomputed ||b - A x|| / ||b|| */
    double estimated_residual; /* projected residual of the last step */
    double backward_error;     /* ||b - A x|| / max(||A x|| + ||b||, min_positive) */
    int32_t converged;
    int32_t status;            /* sbh_solver_status */
    int64_t spmv_count;
    int64_t reorthogonalizations;
    int64_t exact_residual_checks; /* KrylovOpCounts::exact_residual_checks */
} sbh_gmres_solution;

/* b, x0 (may be NULL = zero guess) and x use `memory`; precond may be NULL
 * (identity). */
sbh_status sbh_gmres(sbh_operator* op, sbh_precond* precond, const sbh_c64* b, const sbh_c64* x0,
                     const sbh_gmres_config* config, sbh_c64* x, int memory,
                     sbh_gmres_solution* solution);

/* Many right-hand sides, column by column: B, X0 (NULL = zero) and X are
 * n x k row-major (1 <= k <= 64); solutions[k].  Each column follows
 * gmres_core exactly (the single-RHS GMRES of gmres.rs:631-1057); the columns
 * share one block product per Arnoldi step (apply_block, traits.rs:3000-3017).
 * The reference's own multi-RHS solver is the shared-basis block method,
 * sbh_block_gmres below. */
sbh_status sbh_gmres_block(sbh_operator* op, const sbh_c64* B, const sbh_c64* X0, int64_t k,
                           const sbh_gmres_config* config, sbh_c64* X, int memory,
                           sbh_gmres_solution* solutions);

/* ---------------------------------------------------------------------------
 * GCRO-DR Krylov recycling across a sequence of systems — GcroDrSubspace,
 * GcroDrConfig and try_gcro_dr_preconditioned
 * (crates/sonium-solvers/crates/sonium-solvers-iterative/src/gcro_dr.rs:
 * 63-157, 199-290, 292-760), the sweep accelerator of
 * frequency_sweep.rs:495-508,636-650.  The subspace holds up to k recycled
 * vectors U on the device; each solve rebuilds C = orth(M^-1 A U) for the
 * current operator (k products), keeps every Arnoldi vector orthogonal to C,
 * and after a solve accepted by the true residual replaces U with harmonic
 * Ritz vectors of the augmented (k + m) problem.  A failed solve leaves U
 * unchanged. */
typedef struct sbh_gcro_subspace sbh_gcro_subspace;

typedef struct {
    int64_t max_iterations; /* restart cycles (default 100) */
    int64_t restart;        /* Krylov dimension m per cycle (default 50) */
    double tolerance;       /* relative to ||M^-1 b|| (default 1e-6) */
} sbh_gcro_dr_config;

/* GcroDrSubspace::new(k) for systems of order n; k = 0 never recycles. */
sbh_status sbh_gcro_subspace_create(sbh_context* ctx, int64_t n, int64_t k, sbh_gcro_subspace** out);
void sbh_gcro_subspace_destroy(sbh_gcro_subspace* subspace);
/* num_vectors (gcro_dr.rs:82-85); -1 for a null handle. */
int64_t sbh_gcro_subspace_size(const sbh_gcro_subspace* subspace);
/* reset (gcro_dr.rs:92-96). */
sbh_status sbh_gcro_subspace_reset(sbh_gcro_subspace* subspace);
/* The recycled vectors, size() x n row-major, host memory. */
sbh_status sbh_gcro_subspace_download(const sbh_gcro_subspace* subspace, sbh_c64* u);
/* b, x0 (NULL = zero) and x use `memory`; precond NULL = identity.  Invalid
 * configurations, non-finite b or x0 and non-finite results are errors (the
 * checked reference boundary); the solution reports iterations (Arnoldi
 * steps, without the k products that rebuild C) and restarts as GMRES does. */
sbh_status sbh_gcro_dr(sbh_operator* op, sbh_precond* precond, const sbh_c64* b, const sbh_c64* x0,
                       sbh_gcro_subspace* subspace, const sbh_gcro_dr_config* config, sbh_c64* x, int memory,
                       sbh_gmres_solution* solution);

/* ---------------------------------------------------------------------------
 * Block GMRES with one shared block Krylov basis — BlockGmresConfig /
 * BlockGmresSolution and try_block_gmres_preconditioned(_with_guess)
 * (crates/sonium-solvers/crates/sonium-solvers-iterative/src/block_gmres.rs:
 * 58-148, 163-330, 666-1120), the solver BemSolver::solve_many routes dense
 * multi-RHS solves to (bem_solver.rs:2175-2181).  One block Arnoldi
 * recurrence on K_m(A, R0), block CGS with the DGKS trigger on the block
 * Frobenius norm, per-column CGS2 QR of each new block with rank-drop
 * deflation at sqrt(eps), a block-Hessenberg Givens least-squares solve
 * after every block step, and per-RHS true residuals at exit. */
typedef struct {
    int64_t max_iterations;               /* restart cycles (default 100) */
    int64_t restart;                      /* block Arnoldi steps per cycle (default 30) */
    double tolerance;                     /* per right-hand side (default 1e-6) */
    int32_t true_residual_check;          /* gate `converged` on the true residuals (default 1) */
    int32_t reserved;
    int64_t exact_residual_check_interval; /* true-residual checkpoint every N block steps (0 = off) */
} sbh_block_gmres_config;

typedef struct {
    int64_t iterations;            /* total block Arnoldi steps */
    int32_t converged;
    int32_t status;                /* sbh_solver_status */
    int64_t spmv_count;            /* KrylovOpCounts::spmv_count (columns multiplied) */
    int64_t exact_residual_checks; /* KrylovOpCounts::exact_residual_checks */
    int64_t restart_cycles;
    int64_t max_basis;             /* largest basis reached (vectors) */
    int64_t deflated;              /* directions dropped by rank-drop deflation */
    int64_t cholesky_qr_blocks;    /* block QRs done as Cholesky QR twice (clearly full rank) */
    int64_t chain_qr_blocks;       /* block QRs done column by column (CGS2 with rank-drop deflation) */
} sbh_block_gmres_info;

/* B, X0 (NULL = zero) and X are n x k row-major (1 <= k <= 64) in `memory`;
 * precond NULL = none (identity).  residuals, true_residuals and
 * backward_errors (k each, may be NULL) receive BlockGmresSolution's
 * per-RHS vectors.  Invalid configurations, non-finite B or X0 and
 * non-finite results are errors, as in the checked reference boundary. */
sbh_status sbh_block_gmres(sbh_operator* op, sbh_precond* precond, const sbh_c64* B, const sbh_c64* X0, int64_t k,
                           const sbh_block_gmres_config* config, sbh_c64* X, int memory, sbh_block_gmres_info* info,
                           double* residuals, double* true_residuals, double* backward_errors);

/* ---------------------------------------------------------------------------
 * Mixed precision.  A complex FP16 or BF16 copy of a system matrix, every row
 * scaled by a power of two (crates/sonium-compute/src/scaled_storage.rs:1-127
 * is the reference's explicit-scale storage precedent); products accumulate
 * in FP32 on MFMA for every k.  The device copy is kept in MFMA fragment
 * order (32-row tiles); sbh_lowp_matrix_download returns it row-major. */
typedef struct sbh_lowp_matrix sbh_lowp_matrix;
typedef enum { SBH_LOWP_FP16 = 0, SBH_LOWP_BF16 = 1 } sbh_lowp_format;
/* Converts the system's current matrix.  The copy holds a reference to sys
 * (sbh_system_destroy frees it after the last copy is destroyed); while a copy
 * exists, sbh_workspace_assemble into sys also leaves the row maxima from
 * which sbh_lowp_matrix_update takes its scales without a pass over the
 * matrix. */
sbh_status sbh_lowp_matrix_create(sbh_system* sys, int format, sbh_lowp_matrix** out);
/* Re-converts after the system was re-assembled. */
sbh_status sbh_lowp_matrix_update(sbh_lowp_matrix* lp);
void sbh_lowp_matrix_destroy(sbh_lowp_matrix* lp);
/* Y ~= A X in low precision, X and Y n x k row-major (1 <= k <= 64). */
sbh_status sbh_lowp_matrix_apply(sbh_lowp_matrix* lp, const sbh_c64* X, sbh_c64* Y, int64_t k, int memory);
/* Test hook: the stored halves, row-major (n x 2*ld, {re, im} per entry),
 * the row inverse scales (n) and ld. */
sbh_status sbh_lowp_matrix_download(sbh_lowp_matrix* lp, uint16_t* halves, double* row_inv, int64_t* ld);

/* FGMRES with iterative refinement — GmresIrConfig / GmresIrStatus and
 * gmres_ir_backend_preconditioned (crates/sonium-compute/src/gmres_ir.rs:
 * 536-665; Carson & Higham 2018, Alg. 5.3): FP64 outer flexible GMRES
 * (residual, Arnoldi, Givens, update) preconditioned by an inner GMRES on the
 * low-precision copy.  restart_size <= 0 means one cycle of
 * max_outer_iterations. */
typedef struct {
    int64_t max_outer_iterations; /* default 30 */
    double outer_tolerance;       /* default 1e-10 */
    int64_t max_inner_iterations; /* default 50 */
    double inner_tolerance;       /* default 1e-3 */
    int64_t restart_size;         /* default 0 (none) */
} sbh_gmres_ir_config;

typedef struct {
    int64_t outer_iterations;
    int64_t total_inner_iterations;
    double estimated_residual;
    double final_outer_residual; /* recomputed FP64 ||b - A x|| / ||b|| */
    double backward_error;
    int32_t converged;
    int32_t status; /* sbh_solver_status */
    int64_t restart_cycles;
    int64_t fp64_products; /* block products with the FP64 matrix */
    int64_t lowp_products; /* block products with the low-precision copy */
} sbh_gmres_ir_status;

sbh_status sbh_gmres_ir(sbh_operator* op, sbh_lowp_matrix* lp, const sbh_c64* B, int64_t k,
                        const sbh_gmres_ir_config* config, sbh_c64* X, int memory,
                        sbh_gmres_ir_status* status);

/* ---------------------------------------------------------------------------
 * Batches of independent rigid scatterers.  No reference counterpart: the
 * reference assembles and solves one system at a time
 * (build_tbem_system_with_beta_parallel, tbem.rs:1284-1466; solve_many,
 * bem_solver.rs:2032, for several right-hand sides of one system).  A batch
 * holds B meshes (velocity BC 0 on every panel, sbh_mesh_create with NULL BC
 * arrays), each with its own wavenumber, Burton-Miller coupling and 1..16
 * plane-wave directions, and assembles and solves all of them together:
 *   - sbh_batch_assemble: two persistent launches fill every instance's FP64
 *     matrix (the same entries as sbh_workspace_assemble) and its plane-wave
 *     right-hand sides (those of sbh_incident_plane_wave_rhs, amplitude 1);
 *   - sbh_batch_solve: all columns of all instances advance in lock step
 *     through restarted GMRES (format SBH_BATCH_FP64) or GMRES iterative
 *     refinement (FP16/BF16 inner Arnoldi on the row-scaled copy, FP64
 *     residuals), with the Hessenberg/Givens work on the device so that one
 *     restart cycle is a fixed launch sequence, captured once in a HIP graph
 *     and replayed (use_graph = 1).  A column is converged when its
 *     recomputed FP64 ||b - A x|| / ||b|| < tolerance. */
typedef struct sbh_batch sbh_batch;

enum { SBH_BATCH_FP64 = -1, SBH_BATCH_FP16_IR = 0, SBH_BATCH_BF16_IR = 1 };

sbh_status sbh_batch_create(sbh_context* ctx, int64_t num_instances, const sbh_mesh* const* meshes,
                            int route_threads, sbh_batch** out);
void sbh_batch_destroy(sbh_batch* batch);
/* physics[b], beta[b] and num_rhs[b] (1..16) per instance; directions holds
 * sum(num_rhs) unit vectors (3 doubles each) in instance order; sign is the
 * incident-field sign of add_incident_field_rhs_stabilized. */
sbh_status sbh_batch_set_problem(sbh_batch* batch, const sbh_physics* physics, const sbh_c64* beta,
                                 const int32_t* num_rhs, const double* directions, double sign);
sbh_status sbh_batch_assemble(sbh_batch* batch);
/* Row-scaled complex FP16 (0) or BF16 (1) copy of every instance matrix. */
sbh_status sbh_batch_convert(sbh_batch* batch, int format);

typedef struct {
    int32_t format;           /* SBH_BATCH_FP64, SBH_BATCH_FP16_IR, SBH_BATCH_BF16_IR */
    int32_t use_graph;        /* capture one restart cycle in a HIP graph */
    int64_t restart;          /* Arnoldi steps per cycle (1..200) */
    int64_t max_cycles;       /* restart cycles (FP64) or refinement steps (IR) */
    double tolerance;         /* on the recomputed FP64 relative residual */
    double inner_tolerance;   /* IR: inner estimate relative to the current residual (never below
                                 tolerance * ||b|| / 4: a cycle stops once it passes the outer target) */
    int32_t use_guess;        /* start from the solutions set by sbh_batch_set_guess (else zero) */
    int32_t pad;
} sbh_batch_solve_config;

typedef struct {
    int64_t iterations;             /* Arnoldi steps (operator products) */
    int64_t cycles;                 /* restarts (FP64) or refinement steps (IR) after the first cycle */
    int64_t reorthogonalizations;
    double residual;                /* recomputed FP64 ||b - A x|| / ||b|| */
    int32_t converged;
    int32_t pad;
    double completion_ms;           /* device clock: last sbh_batch_assemble start -> convergence */
} sbh_batch_column_status;

/* status has sum(num_rhs) entries, instance-major. */
sbh_status sbh_batch_solve(sbh_batch* batch, const sbh_batch_solve_config* config,
                           sbh_batch_column_status* status);
/* Y = A X for every instance at once with the batch's FP64 matrices (format
 * SBH_BATCH_FP64) or its FP16/BF16 copies (after sbh_batch_convert): X and Y
 * are the packed host vectors (instance b's n_b x num_rhs_b row-major block
 * after those of instances 0..b-1). */
sbh_status sbh_batch_apply(sbh_batch* batch, int format, const sbh_c64* X, sbh_c64* Y);
/* Initial guess of one instance (n x num_rhs row-major, host memory), used
 * by the next sbh_batch_solve with use_guess = 1 (gmres_with_guess,
 * gmres.rs:564-580; the relative residual stays relative to ||b||). */
sbh_status sbh_batch_set_guess(sbh_batch* batch, int64_t instance, const sbh_c64* x0);
/* The FP16/BF16 copy of one instance after sbh_batch_convert: n x 2*ld halves
 * row-major (entry j of row r as halves 2j, 2j+1 = re, im of A_rj / row_inv[r],
 * zero padding to ld entries) and the n power-of-two row scales; *ld may be
 * read alone by passing NULL arrays. */
sbh_status sbh_batch_download_lowp(sbh_batch* batch, int64_t instance, uint16_t* halves, double* row_inv,
                                   int64_t* ld);
/* Every instance's guess at once (the packed layout of sbh_batch_apply). */
sbh_status sbh_batch_set_guesses(sbh_batch* batch, const sbh_c64* X0);
/* Every instance's solutions (what = 0) or right-hand sides (what = 1), packed. */
sbh_status sbh_batch_download_all(sbh_batch* batch, int what, sbh_c64* out);
/* Device-resident sweep data.  Pointer tables hold one device pointer per
 * instance (of this batch's device), each addressing an n_b x num_rhs_b
 * row-major block; a null entry stands for a zero block.
 *   - sbh_batch_set_guesses_device: the initial guesses of the next
 *     sbh_batch_solve with use_guess = 1, copied from x0[b];
 *   - sbh_batch_store_solutions: each instance's current solution copied to
 *     dst[b] (null entries skipped);
 *   - sbh_batch_guess_gram: the minimal-residual data of m (1..6) candidate
 *     guesses per column: candidates[a * B + b] is instance b's candidate a;
 *     with P_a = A X_a (format as in sbh_batch_apply) it returns per column c
 *     gram[(c m + a) m + a'] = P_a^H P_a', projections[c m + a] = P_a^H b_c and
 *     rhs_norm2[c] = ||b_c||^2;
 *   - sbh_batch_guess_combine: x0 = sum_a coefficients[c m + a] X_a per column
 *     from the candidates of the last sbh_batch_guess_gram (same m), as the
 *     next initial guesses.
 * The previous-frequency guess of the reference's sweep
 * (frequency_sweep.rs:718-730) is the case m = 1 with coefficient 1. */
sbh_status sbh_batch_set_guesses_device(sbh_batch* batch, const sbh_c64* const* x0);
sbh_status sbh_batch_store_solutions(sbh_batch* batch, sbh_c64* const* dst);
sbh_status sbh_batch_guess_gram(sbh_batch* batch, int format, int m, const sbh_c64* const* candidates,
                                sbh_c64* gram, sbh_c64* projections, double* rhs_norm2);
sbh_status sbh_batch_guess_combine(sbh_batch* batch, int m, const sbh_c64* coefficients);
/* n x num_rhs row-major blocks of one instance (host memory). */
sbh_status sbh_batch_download_solution(sbh_batch* batch, int64_t instance, sbh_c64* x);
sbh_status sbh_batch_download_rhs(sbh_batch* batch, int64_t instance, sbh_c64* rhs);
sbh_status sbh_batch_download_matrix(sbh_batch* batch, int64_t instance, sbh_c64* a);
sbh_status sbh_batch_info(const sbh_batch* batch, int64_t* num_instances, int64_t* num_columns,
                          int64_t* total_panels, int64_t* matrix_entries, double* route_ms);

/* Launch shapes: rows per assembly item, persistent workgroups per CU,
 * elements per vector-kernel chunk. */
typedef struct {
    int32_t rows_per_item;   /* 1..64 */
    int32_t blocks_per_cu;   /* 1..16 */
    int32_t chunk_elems;     /* 256..65536 */
    int32_t pad;
} sbh_batch_tuning;

sbh_status sbh_batch_get_tuning(const sbh_batch* batch, sbh_batch_tuning* tuning);
sbh_status sbh_batch_set_tuning(sbh_batch* batch, const sbh_batch_tuning* tuning);
/* Times every candidate shape on this batch with HIP events (the assembly
 * once per (rows_per_item, blocks_per_cu) pair; the chunk size on one
 * Gram-Schmidt pass over 16 vectors), keeps the fastest, and reports the
 * candidates and their times (up to max_entries rows of 4 doubles:
 * rows_per_item, blocks_per_cu, chunk_elems, ms; unused fields 0). */
sbh_status sbh_batch_autotune(sbh_batch* batch, double* table, int max_entries, int* num_entries,
                              sbh_batch_tuning* best);

/* ---------------------------------------------------------------------------
 * Learned warm start.  No reference counterpart; the reference's warm start
 * is the previous frequency's solution (frequency_sweep.rs:718-730) passed to
 * gmres_with_guess.  A 4-layer MLP (64 features -> width -> width -> width ->
 * 2, SiLU) with BF16 weights and FP32 accumulation maps one panel's raw
 * input {c[3], n[3], d[3], k} (centroid, outward normal, incident direction,
 * wavenumber; 10 floats) to the incident-phase-free density
 * u = x exp(-i k d.c) as {Re, Im}; training minimises sum |u_pred - u|^2 / (2M)
 * with Adam on FP32 master weights.  Parameters are one FP32 vector: per layer
 * the row-major weights (out x in) followed by the bias (out). */
typedef struct sbh_mlp sbh_mlp;

sbh_status sbh_mlp_create(sbh_context* ctx, int32_t width, uint64_t seed, sbh_mlp** out);
void sbh_mlp_destroy(sbh_mlp* net);
int64_t sbh_mlp_num_params(const sbh_mlp* net);
sbh_status sbh_mlp_get_params(sbh_mlp* net, float* params);
sbh_status sbh_mlp_set_params(sbh_mlp* net, const float* params);
/* Loss and parameter gradients on M samples (M a multiple of 256; host arrays
 * M x 10 inputs, M x 2 targets; grads in the parameter layout). */
sbh_status sbh_mlp_loss_and_gradients(sbh_mlp* net, const float* inputs, const float* targets, int64_t M,
                                      double* loss, float* grads);
/* Training set (host arrays count x 10 and count x 2), copied to the device. */
sbh_status sbh_mlp_set_dataset(sbh_mlp* net, const float* inputs, const float* targets, int64_t count);
/* steps Adam steps on batches drawn from the training set (cosine learning
 * rate from lr_max to lr_min); losses (optional, steps entries) receive the
 * batch loss before each update. */
sbh_status sbh_mlp_train(sbh_mlp* net, int64_t steps, int64_t batch, double lr_max, double lr_min, uint64_t seed,
                         double* losses);
/* Predictions (M x 2 floats) for M raw inputs (M x 10 floats). */
sbh_status sbh_mlp_predict(sbh_mlp* net, const float* inputs, int64_t M, float* out, int memory);
/* The network's initial guess x0 = u exp(i k d.c) for every column of every
 * instance of a batch (panel centroids and outward normals, the instance's
 * wavenumber, the column's direction), computed on the device and written
 * to the device vector out (packed layout of sbh_batch_apply) or, with out
 * NULL, as the batch's initial guesses; net and batch share one context. */
sbh_status sbh_batch_learned_guesses(sbh_batch* batch, sbh_mlp* net, sbh_c64* out);

/* ---------------------------------------------------------------------------
 * Mesh2HRTF / NumCalc projects (host only).
 * sbh_nc_project_load — load_project (crates/sonium-bem/src/numcalc/project.rs:78-105):
 *   parse_nc_input (crates/sonium-bem/src/core/io/nc_format.rs:205-636), read the
 *   node and element files (load_nc_nodes / load_nc_elements, nc_format.rs:694-781;
 *   the first num_object_files pairs are the object, the rest evaluation grids),
 *   apply the BOUNDARY ranges (VELO, PRES, ADMI) over all elements in file order
 *   and expand the frequency curve (positive, finite, sorted, unique).
 * sbh_nc_parse — parse_nc_input_string (nc_format.rs:215): the NC.inp text only;
 *   file names are joined to base_dir and no mesh file is read.
 * Errors: SBH_ERR_INVALID_ARGUMENT with "parse error at line N: ...",
 *   "missing required section: ...", "IO error: ...", or "boundary spec covers
 *   element ..." (NcParseError / ProjectError, nc_format.rs:184-203,
 *   project.rs:18-33); an unknown BC keyword gives SBH_ERR_UNSUPPORTED. */
typedef struct sbh_nc_project sbh_nc_project;
typedef struct {
    int64_t num_nodes, num_elements, num_evaluation, num_frequencies, num_curve_points;
    int64_t num_plane_waves, num_point_sources, num_boundary_specs, num_node_files, num_element_files;
    int32_t mesh_loaded;
    int32_t main_params_i[9];   /* element type, nodes, elements, object files, evaluation files,
                                   BC type, method, FMM method, parallel */
    double main_params_ii[7];   /* as read; preconditioner and solver are entries 0 and 1 */
    double speed_of_sound, density, reference_pressure;
    int32_t has_symmetry, symmetry_flags[3];
    double symmetry_origin[3];
} sbh_nc_info;
typedef enum {
    SBH_NC_VERSION = 0, SBH_NC_DESCRIPTION = 1, SBH_NC_NODE_FILE = 2, SBH_NC_ELEMENT_FILE = 3,
    SBH_NC_BOUNDARY_TYPE = 4
} sbh_nc_string;
typedef enum {
    SBH_NC_CONTROL_I = 0, SBH_NC_CONTROL_II = 1, SBH_NC_MAIN_III = 2, SBH_NC_MAIN_IV_RESERVED = 3,
    SBH_NC_FREQUENCY_CURVE = 4,   /* (time, frequency, phase) triples */
    SBH_NC_FREQUENCIES = 5,       /* the expanded solve frequencies */
    SBH_NC_BOUNDARY_SPECS = 6,    /* (start, end, re, curve_re, im, curve_im) per record */
    SBH_NC_PLANE_WAVES = 7,       /* (dx, dy, dz, re, curve_re, im, curve_im) per record */
    SBH_NC_POINT_SOURCES = 8      /* (x, y, z, re, curve_re, im, curve_im) per record */
} sbh_nc_values;
sbh_status sbh_nc_project_load(const char* nc_inp_path, sbh_nc_project** out);
sbh_status sbh_nc_parse(const char* content, const char* base_dir, sbh_nc_project** out);
void sbh_nc_project_destroy(sbh_nc_project* project);
sbh_status sbh_nc_project_info(const sbh_nc_project* project, sbh_nc_info* info);
/* NUL-terminated copy (truncated to size - 1 bytes); index selects the file or record. */
sbh_status sbh_nc_project_string(const sbh_nc_project* project, int which, int64_t index, char* buf,
                                 int64_t size);
/* *count = number of doubles; they are written when out != NULL and capacity suffices. */
sbh_status sbh_nc_project_values(const sbh_nc_project* project, int which, double* out, int64_t capacity,
                                 int64_t* count);
/* All elements in file order (each array may be NULL): connectivity 4 per element
 * (-1 padded, node indices into the combined node list), counts = nodes per element,
 * is_evaluation, and the BC each element got (type 0/1, value, admittance). */
sbh_status sbh_nc_project_mesh(const sbh_nc_project* project, double* nodes, int64_t* connectivity,
                               int32_t* counts, uint8_t* is_evaluation, int32_t* bc_type, sbh_c64* bc_value,
                               sbh_c64* admittance, uint8_t* has_admittance);
/* Centroids of the evaluation elements (NumCalcProject::evaluation_centers,
 * project.rs:65-71), num_evaluation x 3. */
sbh_status sbh_nc_project_evaluation_centers(const sbh_nc_project* project, double* out);
/* The object's Tri3 elements with their BCs as an sbh_mesh with vertex-order
 * normals (run_project's surface_elements, runner.rs:63-68); Quad4 surface
 * elements give SBH_ERR_UNSUPPORTED. */
sbh_status sbh_nc_project_surface_mesh(const sbh_nc_project* project, sbh_context* ctx, sbh_mesh** out);

/* ---------------------------------------------------------------------------
 * Test hooks (no GPU work): the host route builder's view of one pair, so
 * tests can compare it with the CPU oracle's generate_subelements.
 * leaves_out: up to 110 x 6 doubles (xi,eta of 3 vertices); returns the
 * leaf count in *count (1 leaf with the full element means "far"). */
sbh_status sbh_debug_subelements(const double* source_point, const double* element_coords,
                                 double element_area, double* leaves_out, int* count);

/* Test hook: the device sin/cos used by the assembly kernels, evaluated at
 * host-supplied points (for accuracy checks against the C library). */
sbh_status sbh_debug_device_sincos(sbh_context* ctx, const double* x, int64_t n, double* s, double* c);
/* Test hook: the device 1/sqrt(x) used by the kernels, at host-supplied x > 0. */
sbh_status sbh_debug_device_rsqrt(sbh_context* ctx, const double* x, int64_t n, double* y);
/* Test hook: `reps` launches of a plain device copy of `bytes` bytes (reads
 * `bytes`, writes `bytes`), used to calibrate HBM byte counters. */
sbh_status sbh_debug_device_copy(sbh_context* ctx, int64_t bytes, int reps);
/* Roofline ceiling probes (no counterpart in the reference): the measured
 * rate of a streaming read of `bytes` bytes (GB/s: the best over launch
 * shapes of 1, 2, 4 and 8 workgroups per CU, each the average of `reps`
 * launches) and of independent FP64 FMAs on every CU (TFLOP/s, FMA = 2). */
sbh_status sbh_debug_hbm_read_rate(sbh_context* ctx, int64_t bytes, int reps, double* gbs);
sbh_status sbh_debug_fp64_fma_rate(sbh_context* ctx, double* tflops);
/* Test hook: sbh_mlp_predict through the layer-by-layer forward pass of
 * training (one GEMM launch per layer, activations in HBM) instead of the
 * fused inference kernel, to pin the two to the same bits. */
sbh_status sbh_debug_mlp_predict_layered(sbh_mlp* net, const float* inputs, int64_t M, float* out);

#ifdef __cplusplus
}
#endif

#endif /* SBH_H */
