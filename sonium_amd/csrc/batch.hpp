// Batched solver internals shared by batch_kernels.hip and batch.cpp.
#pragma once

#include "internal.hpp"

namespace sbh {

// Rows [r0, r0 + rows) of one instance: the unit of the vector kernels.
struct VecChunk {
    int32_t inst;
    int32_t rows;
    int64_t r0;
};

// Everything the vector kernels need about the batch layout (device pointers).
struct RaggedLayout {
    int B, C, nchunks, MV, m;
    int64_t L;
    const BatchInstance* inst;
    const VecChunk* chunks;
    const int2* chunk_range;  // per instance: [first chunk, end chunk)
    const int* col_inst;
    const int* col_q;
};

void batch_lp_convert(sbh_context* ctx, int format, int64_t rows, const int* row_inst, const BatchInstance* inst,
                      const double2* A, uint16_t* Alp, double* row_inv);
// format < 0: FP64 product; 0/1: FP16/BF16 copy with FP32 accumulation.
// max_cols: the largest right-hand-side count of any instance (1..16).
void batch_product(sbh_context* ctx, int format, int max_cols, int ntiles, const int2* tiles,
                   const BatchInstance* inst, const int* flag, const double2* A, const uint16_t* Alp,
                   const double* row_inv, const double2* X, double2* Y);
void ragged_dots(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* V, int nvec,
                 const double2* W, double2* part);
void ragged_sum_dots(sbh_context* ctx, const RaggedLayout& lay, const double2* part, int nvec, int j,
                     const int* active, const int* reorth, int pass, double2* H, double2* coef, double* wn0);
void ragged_project(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* V, int nvec,
                    const double2* coef, double2* W, double2* part);
void ragged_norm(sbh_context* ctx, const RaggedLayout& lay, const double2* part, const int* active, const double* wn0,
                 int pass, double* wn, int* reorth, int64_t* reorth_count);
void ragged_flags(sbh_context* ctx, const RaggedLayout& lay, const int* active, const int* reorth, const int* done,
                  int* inst_active, int* inst_reorth, int* inst_pending);
void ragged_givens(sbh_context* ctx, const RaggedLayout& lay, int j, int* active, const double* wn, double2* H,
                   double2* cs, double2* sn, double2* g, const double* tol_abs, double* est, int* jlen, int64_t* iters,
                   double* scale);
void ragged_solve(sbh_context* ctx, const RaggedLayout& lay, const double2* H, const double2* g, const int* jlen,
                  double2* coef);
void ragged_update(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* V, const int* jlen,
                   const double2* coef, double2* X);
void ragged_residual(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double2* B, double2* R,
                     double2* part);
void ragged_restart(sbh_context* ctx, const RaggedLayout& lay, const double2* part, const int* pending_in, int mode,
                    double tol, double inner_tol, int refinement, double* bnorm, double* rnorm, int* done, int* active,
                    int* jlen, double2* g, double* scale, double* tol_abs, int64_t* cycles,
                    unsigned long long* t_done);
void ragged_scale(sbh_context* ctx, const RaggedLayout& lay, const int* flag, const double* scale, const double2* src,
                  double2* dst);
void ragged_fill(sbh_context* ctx, double2* p, int64_t n);
void ragged_stamp(sbh_context* ctx, unsigned long long* out);

// Initial guesses from device-resident vectors.  max_block is the largest
// n * ncol of any instance; src/dst tables hold one device pointer per
// instance (per layer for the gather).
void batch_gather(sbh_context* ctx, int B, int64_t L, int64_t max_block, int layers, const BatchInstance* inst,
                  const double2* const* src, double2* dst);
void batch_scatter(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const double2* x,
                   double2* const* dst);
void batch_gram(sbh_context* ctx, int m, int C, const BatchInstance* inst, const int* col_inst, const int* col_q,
                const double2* P, int64_t L, const double2* rhs, double2* G, double2* h, double* bb);
void batch_combine(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const double2* S, int64_t L,
                   int m, const double2* coef, double2* x);
void batch_learned_raw(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const Source* src,
                       const double* dirs, float* raw);
void batch_learned_phase(sbh_context* ctx, int B, int64_t max_block, const BatchInstance* inst, const Source* src,
                         const double* dirs, const float* Y, double2* x);

// The warm-start network (mlp_kernels.hip): its context, a raw-input
// workspace of at least M samples (rows [M, padded M) zeroed), and the
// forward pass over M device samples whose M x 2 outputs stay valid until the
// network's next call.
sbh_context* mlp_context(const sbh_mlp* net);
float* mlp_raw_workspace(sbh_mlp* net, int64_t M);
const float* mlp_forward_device(sbh_mlp* net, const float* raw, int64_t M);

}  // namespace sbh
