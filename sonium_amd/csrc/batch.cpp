// Batches of independent rigid scatterers: packed device layout, persistent
// assembly, lock-step restarted GMRES / GMRES iterative refinement with the
// per-column Arnoldi bookkeeping on the device, HIP-graph capture of one
// restart cycle, and the launch-shape autotuner.
//
// Per column the recurrences are those of the single-system solver
// (gmres.cpp, after gmres_core, gmres.rs:631-1057): classical Gram–Schmidt
// with a second pass when ||w|| < 0.707 ||w0||, complex Givens rotations
// (gmres_common.rs:91-105), back substitution, and convergence on the
// recomputed FP64 residual.  Differences from the single-system loop:
// columns whose estimate converges mid-cycle stop there and have their
// solution formed and checked at the end of the cycle, when all columns do;
// with refinement (format FP16/BF16) a cycle is one refinement step — an
// inner Arnoldi cycle on the low-precision copy from the current residual,
// x += V y, then the FP64 residual (Carson & Higham 2018, GMRES-IR).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <thread>
#include <type_traits>

#include "batch.hpp"

struct sbh_batch {
    sbh_context* ctx = nullptr;
    int B = 0;
    int64_t total_panels = 0;
    double route_ms = 0.0;
    std::vector<const sbh_mesh*> meshes;
    std::vector<sbh::BatchInstance> inst;
    // geometry and routes (concatenated)
    sbh::DeviceArray<sbh::FarPanel> d_far;
    sbh::DeviceArray<sbh::Source> d_src;
    sbh::DeviceArray<sbh::PanelGeom> d_geom;
    sbh::DeviceArray<sbh::NearPair> d_pairs;
    sbh::DeviceArray<uint64_t> d_leaves;
    sbh::DeviceArray<int64_t> d_row_ptr;
    sbh::DeviceArray<int> d_row_inst;
    std::vector<int64_t> pair_begin;  // per instance, and the total at the end
    // problem
    bool has_problem = false;
    int C = 0;
    int64_t L = 0;
    int64_t a_elems = 0;
    int64_t lp_elems = 0;
    int lp_format = -2;
    sbh::DeviceArray<sbh::BatchInstance> d_inst;
    sbh::DeviceArray<double> d_dirs;
    sbh::DeviceArray<int2> d_items;
    sbh::DeviceArray<sbh::FreqFamily> d_families;
    std::vector<sbh::FreqFamily> families;
    int64_t family_max_n = 0;
    int64_t family_max_pairs = 0;
    int family_class_end[sbh::FAMILY_CLASSES] = {0, 0, 0, 0};
    sbh::DeviceArray<unsigned> d_counter;
    sbh::DeviceArray<int2> d_tiles;
    sbh::DeviceArray<double2> d_A;
    sbh::DeviceArray<uint16_t> d_Alp;
    sbh::DeviceArray<double> d_row_inv;
    sbh::DeviceArray<double2> d_rhs, d_x;
    sbh::DeviceArray<int> d_col_inst, d_col_q;
    sbh::DeviceArray<unsigned long long> d_t0;
    std::vector<int> col_inst, col_q;
    int nitems = 0, ntiles = 0;
    int max_cols = 16;  // largest right-hand-side count of any instance
    std::vector<int2> tile_host;
    // vector-kernel chunks (depend on the tuning)
    sbh::DeviceArray<sbh::VecChunk> d_chunks;
    sbh::DeviceArray<int2> d_chunk_range;
    int nchunks = 0;
    int chunks_for = 0;  // chunk_elems the chunk table was built for
    std::vector<sbh::VecChunk> chunk_host;
    // tuning
    sbh_batch_tuning tuning{16, 4, 4096, 0};
    // initial guesses from device-resident vectors: gathered candidates and
    // their products (cand_m slots of L), per-column Gram data, pointer tables
    sbh::DeviceArray<double2> d_cand, d_cprod, d_coef, d_G, d_h;
    sbh::DeviceArray<double> d_bb;
    sbh::DeviceArray<const double2*> d_ptrs;
    sbh::DeviceArray<int> d_ones;
    int cand_m = 0;
    int cand_format = -2;
    int cu_count = 256;
    // solver workspace
    struct Solver {
        int m = 0, MV = 0;
        sbh::DeviceArray<double2> V, W, R, H, cs, sn, g, coef, part;
        sbh::DeviceArray<double> wn0, wn, est, bnorm, rnorm, scale, tol_abs;
        sbh::DeviceArray<int> active, reorth, done, jlen, inst_active, inst_reorth, inst_pending;
        sbh::DeviceArray<int64_t> iters, cycles, reorth_count;
        sbh::DeviceArray<unsigned long long> t_done;
        hipGraphExec_t exec = nullptr;
        int graph_format = -3;
        double graph_tol = 0.0, graph_inner_tol = 0.0;  // captured as kernel arguments
        ~Solver() {
            if (exec) (void)hipGraphExecDestroy(exec);
        }
    };
    std::unique_ptr<Solver> solver;
};

namespace sbh {
inline sbh_context* ctx_of(const sbh_batch* b) { return b ? b->ctx : nullptr; }
}  // namespace sbh

namespace sbh {

namespace {

inline void use_device(sbh_context* ctx) { SBH_HIP(hipSetDevice(ctx->device)); }

constexpr int kMaxCol = 16;

template <typename T>
void upload(sbh_context* ctx, DeviceArray<T>& d, const std::vector<T>& h) {
    d.allocate(h.size());
    if (!h.empty()) copy_on_stream(ctx, d.ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
}

void build_chunks(sbh_batch* b) {
    const int64_t elems = b->tuning.chunk_elems;
    std::vector<VecChunk> chunks;
    std::vector<int2> range(b->B);
    for (int i = 0; i < b->B; ++i) {
        const BatchInstance& I = b->inst[i];
        const int64_t rows = std::max<int64_t>(1, elems / I.ncol);
        range[i].x = (int)chunks.size();
        for (int64_t r0 = 0; r0 < I.n; r0 += rows) {
            VecChunk c;
            c.inst = i;
            c.r0 = r0;
            c.rows = (int32_t)std::min<int64_t>(rows, I.n - r0);
            chunks.push_back(c);
        }
        range[i].y = (int)chunks.size();
    }
    SBH_REQUIRE(chunks.size() < (size_t)INT32_MAX, "too many vector chunks");
    b->chunks_for = b->tuning.chunk_elems;
    const bool same = chunks.size() == b->chunk_host.size() &&
                      std::equal(chunks.begin(), chunks.end(), b->chunk_host.begin(), [](const VecChunk& x, const VecChunk& y) {
                          return x.inst == y.inst && x.rows == y.rows && x.r0 == y.r0;
                      });
    if (same) return;  // the solver workspace and its captured graph stay valid
    upload(b->ctx, b->d_chunks, chunks);
    upload(b->ctx, b->d_chunk_range, range);
    b->chunk_host = std::move(chunks);
    b->nchunks = (int)b->chunk_host.size();
    b->solver.reset();  // partial-sum buffers are sized by the chunk count
}

std::vector<int> size_order(const sbh_batch* b) {
    // largest instances first so the persistent grid ends balanced
    std::vector<int> order(b->B);
    for (int i = 0; i < b->B; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->inst[x].n > b->inst[y].n; });
    return order;
}

// Frequency families: instances on the same mesh whose wavenumbers, in
// increasing order, step by a constant dk (each within 1e-12 relative of
// k0 + f dk), cut into groups of at most FAMILY_MAX.  Their far entries come
// from one pass over the pair geometry (batch_far_family_kernel); an instance
// without such neighbours keeps the per-instance far loop.
void build_families(sbh_batch* b) {
    b->families.clear();
    for (auto& I : b->inst) I.in_family = 0;
    std::vector<const sbh_mesh*> mesh_order;
    std::map<const sbh_mesh*, std::vector<int>> by_mesh;
    for (int i = 0; i < b->B; ++i) {
        auto& v = by_mesh[b->meshes[i]];
        if (v.empty()) mesh_order.push_back(b->meshes[i]);
        v.push_back(i);
    }
    int64_t max_n = 0, max_pairs = 0;
    for (const sbh_mesh* m : mesh_order) {
        std::vector<int> ids = by_mesh[m];
        std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) { return b->inst[x].k < b->inst[y].k; });
        size_t s = 0;
        while (s < ids.size()) {
            size_t e = s + 1;
            if (e < ids.size()) {
                const double k0 = b->inst[ids[s]].k;
                const double dk = b->inst[ids[s + 1]].k - k0;
                if (dk > 0.0) {
                    e = s + 2;
                    while (e < ids.size() && (int)(e - s) < FAMILY_MAX) {
                        const double kf = b->inst[ids[e]].k;
                        if (std::fabs(kf - (k0 + (double)(e - s) * dk)) > 1e-12 * std::max(1.0, std::fabs(kf))) break;
                        ++e;
                    }
                }
            }
            if (e - s >= 2) {
                FreqFamily fam;
                std::memset(&fam, 0, sizeof(fam));
                const BatchInstance& first = b->inst[ids[s]];
                fam.panel_off = first.panel_off;
                fam.n = first.n;
                fam.pair_begin = b->pair_begin[ids[s]];
                fam.npairs = b->pair_begin[ids[s] + 1] - fam.pair_begin;
                fam.k0 = first.k;
                fam.dk = b->inst[ids[s + 1]].k - first.k;
                fam.count = (int32_t)(e - s);
                for (size_t f = 0; f < e - s; ++f) {
                    fam.member[f] = ids[s + f];
                    fam.kf[f] = b->inst[ids[s + f]].k;
                    b->inst[ids[s + f]].in_family = 1;
                }
                b->families.push_back(fam);
                max_n = std::max<int64_t>(max_n, fam.n);
                max_pairs = std::max<int64_t>(max_pairs, fam.npairs);
            }
            s = e;
        }
    }
    std::stable_sort(b->families.begin(), b->families.end(), [](const FreqFamily& x, const FreqFamily& y) {
        return family_class(x.count) < family_class(y.count);
    });
    for (int c = 0; c < FAMILY_CLASSES; ++c) b->family_class_end[c] = 0;
    for (const FreqFamily& f : b->families) ++b->family_class_end[family_class(f.count)];
    for (int c = 1; c < FAMILY_CLASSES; ++c) b->family_class_end[c] += b->family_class_end[c - 1];
    b->family_max_n = max_n;
    b->family_max_pairs = max_pairs;
    if (!b->families.empty()) upload(b->ctx, b->d_families, b->families);
}

// Assembly items (instance, first row); they depend on the tuning.
void build_items(sbh_batch* b) {
    std::vector<int2> items;
    const int R = b->tuning.rows_per_item;
    for (int i : size_order(b))
        for (int64_t r0 = 0; r0 < b->inst[i].n; r0 += R) items.push_back(make_int2(i, (int)r0));
    SBH_REQUIRE(items.size() < (size_t)INT32_MAX, "too many assembly items");
    upload(b->ctx, b->d_items, items);
    b->nitems = (int)items.size();
}

// Product tiles (instance, first of 128 rows); fixed per problem, since a
// captured solve graph holds their address.
void build_tiles(sbh_batch* b) {
    std::vector<int2> tiles;
    for (int i : size_order(b))
        for (int64_t r0 = 0; r0 < b->inst[i].n; r0 += 128) tiles.push_back(make_int2(i, (int)r0));
    const bool same = tiles.size() == b->tile_host.size() &&
                      std::equal(tiles.begin(), tiles.end(), b->tile_host.begin(),
                                 [](const int2& x, const int2& y) { return x.x == y.x && x.y == y.y; });
    if (same) return;
    upload(b->ctx, b->d_tiles, tiles);
    b->tile_host = std::move(tiles);
    b->ntiles = (int)b->tile_host.size();
    b->solver.reset();  // a captured graph holds the tile table's address
}

RaggedLayout layout(const sbh_batch* b) {
    RaggedLayout l;
    l.B = b->B;
    l.C = b->C;
    l.nchunks = b->nchunks;
    l.MV = b->solver->MV;
    l.m = b->solver->m;
    l.L = b->L;
    l.inst = b->d_inst.ptr;
    l.chunks = b->d_chunks.ptr;
    l.chunk_range = b->d_chunk_range.ptr;
    l.col_inst = b->d_col_inst.ptr;
    l.col_q = b->d_col_q.ptr;
    return l;
}

void ensure_solver(sbh_batch* b, int m) {
    if (b->chunks_for != b->tuning.chunk_elems) build_chunks(b);
    if (b->solver && b->solver->m == m) return;
    b->solver.reset(new sbh_batch::Solver());
    auto& s = *b->solver;
    s.m = m;
    s.MV = m + 2;
    const size_t C = (size_t)b->C, L = (size_t)b->L;
    s.V.allocate((size_t)(m + 1) * L);
    s.W.allocate(L);
    s.R.allocate(L);
    s.H.allocate(C * m * (m + 1));
    s.cs.allocate(C * m);
    s.sn.allocate(C * m);
    s.g.allocate(C * (m + 1));
    s.coef.allocate(C * s.MV);
    s.part.allocate((size_t)b->nchunks * s.MV * kMaxCol);
    for (auto* a : {&s.wn0, &s.wn, &s.est, &s.bnorm, &s.rnorm, &s.scale, &s.tol_abs}) a->allocate(C);
    for (auto* a : {&s.active, &s.reorth, &s.done, &s.jlen}) a->allocate(C);
    for (auto* a : {&s.inst_active, &s.inst_reorth, &s.inst_pending}) a->allocate((size_t)b->B);
    for (auto* a : {&s.iters, &s.cycles, &s.reorth_count}) a->allocate(C);
    s.t_done.allocate(C);
    // finite contents everywhere: frozen columns multiply stale basis vectors
    SBH_HIP(hipMemsetAsync(s.V.ptr, 0, s.V.bytes(), b->ctx->stream));
    SBH_HIP(hipMemsetAsync(s.H.ptr, 0, s.H.bytes(), b->ctx->stream));
    SBH_HIP(hipMemsetAsync(s.reorth.ptr, 0, s.reorth.bytes(), b->ctx->stream));
}

// One restart cycle (m Arnoldi steps, solution update, FP64 residual, start
// of the next cycle) as a fixed launch sequence.
void cycle(sbh_batch* b, int format, double tol, double inner_tol) {
    sbh_context* ctx = b->ctx;
    auto& s = *b->solver;
    const RaggedLayout lay = layout(b);
    const int m = s.m;
    const int refinement = format >= 0 ? 1 : 0;
    const size_t L = (size_t)b->L;
    auto Vj = [&](int j) { return s.V.ptr + (size_t)j * L; };
    for (int j = 0; j < m; ++j) {
        batch_product(ctx, format, b->max_cols, b->ntiles, b->d_tiles.ptr, b->d_inst.ptr, s.inst_active.ptr, b->d_A.ptr,
                      b->d_Alp.ptr, b->d_row_inv.ptr, Vj(j), s.W.ptr);
        ragged_dots(ctx, lay, s.inst_active.ptr, s.V.ptr, j + 1, s.W.ptr, s.part.ptr);
        ragged_sum_dots(ctx, lay, s.part.ptr, j + 1, j, s.active.ptr, s.reorth.ptr, 0, s.H.ptr, s.coef.ptr,
                        s.wn0.ptr);
        ragged_project(ctx, lay, s.inst_active.ptr, s.V.ptr, j + 1, s.coef.ptr, s.W.ptr, s.part.ptr);
        ragged_norm(ctx, lay, s.part.ptr, s.active.ptr, s.wn0.ptr, 0, s.wn.ptr, s.reorth.ptr, s.reorth_count.ptr);
        ragged_flags(ctx, lay, s.active.ptr, s.reorth.ptr, s.done.ptr, s.inst_active.ptr, s.inst_reorth.ptr,
                     s.inst_pending.ptr);
        ragged_dots(ctx, lay, s.inst_reorth.ptr, s.V.ptr, j + 1, s.W.ptr, s.part.ptr);
        ragged_sum_dots(ctx, lay, s.part.ptr, j + 1, j, s.active.ptr, s.reorth.ptr, 1, s.H.ptr, s.coef.ptr,
                        s.wn0.ptr);
        ragged_project(ctx, lay, s.inst_reorth.ptr, s.V.ptr, j + 1, s.coef.ptr, s.W.ptr, s.part.ptr);
        ragged_norm(ctx, lay, s.part.ptr, s.active.ptr, s.wn0.ptr, 1, s.wn.ptr, s.reorth.ptr, s.reorth_count.ptr);
        ragged_givens(ctx, lay, j, s.active.ptr, s.wn.ptr, s.H.ptr, s.cs.ptr, s.sn.ptr, s.g.ptr, s.tol_abs.ptr,
                      s.est.ptr, s.jlen.ptr, s.iters.ptr, s.scale.ptr);
        if (j + 1 < m)
            ragged_scale(ctx, lay, s.inst_active.ptr, s.scale.ptr, s.W.ptr, Vj(j + 1));
        ragged_flags(ctx, lay, s.active.ptr, s.reorth.ptr, s.done.ptr, s.inst_active.ptr, s.inst_reorth.ptr,
                     s.inst_pending.ptr);
    }
    ragged_solve(ctx, lay, s.H.ptr, s.g.ptr, s.jlen.ptr, s.coef.ptr);
    ragged_update(ctx, lay, s.inst_pending.ptr, s.V.ptr, s.jlen.ptr, s.coef.ptr, b->d_x.ptr);
    batch_product(ctx, -1, b->max_cols, b->ntiles, b->d_tiles.ptr, b->d_inst.ptr, s.inst_pending.ptr, b->d_A.ptr, nullptr,
                  nullptr, b->d_x.ptr, s.R.ptr);
    ragged_residual(ctx, lay, s.inst_pending.ptr, b->d_rhs.ptr, s.R.ptr, s.part.ptr);
    ragged_restart(ctx, lay, s.part.ptr, s.inst_pending.ptr, 0, tol, inner_tol, refinement, s.bnorm.ptr,
                   s.rnorm.ptr, s.done.ptr, s.active.ptr, s.jlen.ptr, s.g.ptr, s.scale.ptr, s.tol_abs.ptr,
                   s.cycles.ptr, s.t_done.ptr);
    ragged_scale(ctx, lay, s.inst_pending.ptr, s.scale.ptr, s.R.ptr, Vj(0));
    ragged_flags(ctx, lay, s.active.ptr, s.reorth.ptr, s.done.ptr, s.inst_active.ptr, s.inst_reorth.ptr,
                 s.inst_pending.ptr);
}

void fill_int(sbh_context* ctx, DeviceArray<int>& a, int v) {
    std::vector<int> h(a.count, v);
    SBH_HIP(hipMemcpyAsync(a.ptr, h.data(), a.bytes(), hipMemcpyHostToDevice, ctx->stream));
    SBH_HIP(hipStreamSynchronize(ctx->stream));
}

void check_problem(const sbh_batch* b) {
    SBH_REQUIRE(b, "batch is null");
    SBH_REQUIRE(b->has_problem, "sbh_batch_set_problem has not been called");
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    SBH_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

void run_assembly(sbh_batch* b) {
    BatchAssemblyArgs args;
    args.families = b->d_families.ptr;
    args.nfamilies = (int)b->families.size();
    args.family_max_n = b->family_max_n;
    args.family_max_pairs = b->family_max_pairs;
    for (int c = 0; c < FAMILY_CLASSES; ++c) args.family_class_end[c] = b->family_class_end[c];
    args.inst = b->d_inst.ptr;
    args.items = b->d_items.ptr;
    args.nitems = b->nitems;
    args.rows_per_item = b->tuning.rows_per_item;
    args.counter = b->d_counter.ptr;
    args.far = b->d_far.ptr;
    args.src = b->d_src.ptr;
    args.geom = b->d_geom.ptr;
    args.pairs = b->d_pairs.ptr;
    args.row_ptr = b->d_row_ptr.ptr;
    args.leaves = b->d_leaves.ptr;
    args.dirs = b->d_dirs.ptr;
    args.A = b->d_A.ptr;
    args.rhs = b->d_rhs.ptr;
    const int64_t grid = std::min<int64_t>((int64_t)b->cu_count * b->tuning.blocks_per_cu, b->nitems);
    launch_batch_assembly(b->ctx, args, (int)std::max<int64_t>(grid, 1));
}

}  // namespace
}  // namespace sbh

using namespace sbh;

extern "C" {

sbh_status sbh_batch_create(sbh_context* ctx, int64_t num_instances, const sbh_mesh* const* meshes,
                            int route_threads, sbh_batch** out) {
    return guard_on(ctx_of(ctx), [&] {
        SBH_REQUIRE(ctx && out, "ctx/out is null");
        SBH_REQUIRE(num_instances > 0 && num_instances < (1 << 24), "num_instances out of range");
        SBH_REQUIRE(meshes, "meshes is null");
        *out = nullptr;
        use_device(ctx);
        std::unique_ptr<sbh_batch> b(new sbh_batch());
        b->ctx = ctx;
        b->B = (int)num_instances;
        b->meshes.assign(meshes, meshes + num_instances);
        int64_t panels = 0;
        for (const sbh_mesh* m : b->meshes) {
            SBH_REQUIRE(m, "mesh is null");
            SBH_REQUIRE(m->ctx == ctx, "mesh belongs to another context");
            if (m->general)
                throw Error(SBH_ERR_UNSUPPORTED, "batches take rigid meshes only (velocity BC 0 on every panel)");
            SBH_REQUIRE(m->n > 0 && m->n < (1 << 24), "instance size out of range");
            panels += m->n;
        }
        SBH_REQUIRE(panels < (int64_t)INT32_MAX, "too many panels in the batch");
        b->total_panels = panels;
        // routes: instances are independent, split them over host threads
        auto t0 = std::chrono::steady_clock::now();
        std::vector<RouteTable> rts(b->B);
        int nthreads = route_threads > 0 ? route_threads : (int)std::max(1u, std::thread::hardware_concurrency());
        nthreads = std::min(nthreads, b->B);
        {
            std::vector<std::thread> pool;
            std::vector<std::exception_ptr> errs(nthreads);
            for (int t = 0; t < nthreads; ++t)
                pool.emplace_back([&, t] {
                    try {
                        for (int i = t; i < b->B; i += nthreads) rts[i] = build_routes(b->meshes[i], 1);
                    } catch (...) {
                        errs[t] = std::current_exception();
                    }
                });
            for (auto& th : pool) th.join();
            for (auto& e : errs)
                if (e) std::rethrow_exception(e);
        }
        b->route_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        // concatenate geometry (copied from each mesh's device arrays) and routes
        b->d_far.allocate(panels);
        b->d_src.allocate(panels);
        b->d_geom.allocate(panels);
        std::vector<NearPair> pairs;
        std::vector<uint64_t> leaves;
        std::vector<int64_t> row_ptr((size_t)panels + 1, 0);
        std::vector<int> row_inst((size_t)panels);
        b->inst.resize(b->B);
        b->pair_begin.assign((size_t)b->B + 1, 0);
        int64_t off = 0;
        for (int i = 0; i < b->B; ++i) {
            const sbh_mesh* m = b->meshes[i];
            copy_on_stream(b->ctx, b->d_far.ptr + off, m->d_far.ptr, m->n * sizeof(FarPanel), hipMemcpyDeviceToDevice);
            copy_on_stream(b->ctx, b->d_src.ptr + off, m->d_src.ptr, m->n * sizeof(Source), hipMemcpyDeviceToDevice);
            copy_on_stream(b->ctx, b->d_geom.ptr + off, m->d_geom.ptr, m->n * sizeof(PanelGeom), hipMemcpyDeviceToDevice);
            const int64_t pair_off = (int64_t)pairs.size();
            b->pair_begin[i] = pair_off;
            const int64_t leaf_off = (int64_t)leaves.size();
            for (NearPair p : rts[i].pairs) {
                p.begin += leaf_off;
                pairs.push_back(p);
            }
            leaves.insert(leaves.end(), rts[i].leaves.begin(), rts[i].leaves.end());
            for (int64_t r = 0; r < m->n; ++r) {
                row_ptr[off + r] = pair_off + rts[i].row_ptr[r];
                row_inst[off + r] = i;
            }
            BatchInstance& I = b->inst[i];
            std::memset(&I, 0, sizeof(I));
            I.panel_off = off;
            I.n = m->n;
            off += m->n;
            rts[i] = RouteTable();
        }
        row_ptr[panels] = (int64_t)pairs.size();
        b->pair_begin[b->B] = (int64_t)pairs.size();
        upload(b->ctx, b->d_pairs, pairs);
        upload(b->ctx, b->d_leaves, leaves);
        upload(b->ctx, b->d_row_ptr, row_ptr);
        upload(b->ctx, b->d_row_inst, row_inst);
        b->d_counter.allocate(2);
        b->d_t0.allocate(1);
        zero_on_stream(b->ctx, b->d_t0.ptr, sizeof(unsigned long long));
        int cus = 0;
        SBH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        b->cu_count = std::max(cus, 1);
        context_retain(ctx);
        *out = b.release();
    });
}

void sbh_batch_destroy(sbh_batch* b) {
    if (!b) return;
    sbh_context* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    release_device(ctx, (int64_t)b->d_A.bytes());
    delete b;
    context_release(ctx);
}

sbh_status sbh_batch_set_problem(sbh_batch* b, const sbh_physics* physics, const sbh_c64* beta,
                                 const int32_t* num_rhs, const double* directions, double sign) {
    return guard_on(ctx_of(b), [&] {
        SBH_REQUIRE(b && physics && beta && num_rhs && directions, "null argument");
        use_device(b->ctx);
        SBH_HIP(hipStreamSynchronize(b->ctx->stream));
        int64_t a_off = 0, vec_off = 0, lp_off = 0, dir_off = 0;
        int col = 0;
        b->col_inst.clear();
        b->col_q.clear();
        for (int i = 0; i < b->B; ++i) {
            BatchInstance& I = b->inst[i];
            const sbh_physics& ph = physics[i];
            SBH_REQUIRE(num_rhs[i] >= 1 && num_rhs[i] <= kMaxCol, "num_rhs must be in 1..16");
            SBH_REQUIRE(std::isfinite(ph.wave_number), "wave number is not finite");
            if (ph.harmonic_factor != 1.0)
                throw Error(SBH_ERR_UNSUPPORTED, "only harmonic_factor = +1 (ACOUSTIC_CONVENTION) is supported");
            SBH_REQUIRE(std::fabs(ph.wave_number) * b->meshes[i]->diameter < 5.0e5,
                        "k * mesh diameter exceeds the supported phase range (5e5 rad)");
            I.k = ph.wave_number;
            I.tau = ph.tau;
            I.sign = sign;
            I.beta[0] = beta[i].re;
            I.beta[1] = beta[i].im;
            I.ncol = num_rhs[i];
            I.col_off = col;
            I.dir_off = dir_off;
            I.a_off = a_off;
            I.vec_off = vec_off;
            I.ld = (I.n + 63) / 64 * 64;
            I.lp_off = lp_off;
            for (int q = 0; q < I.ncol; ++q) {
                b->col_inst.push_back(i);
                b->col_q.push_back(q);
            }
            a_off += I.n * I.n;
            vec_off += I.n * I.ncol;
            lp_off += I.n * 2 * I.ld;
            dir_off += I.ncol;
            col += I.ncol;
        }
        b->C = col;
        b->L = vec_off;
        int maxc = 1;
        for (const auto& I : b->inst) maxc = std::max(maxc, (int)I.ncol);
        if (maxc != b->max_cols) b->solver.reset();  // the captured product kernels depend on it
        b->max_cols = maxc;
        b->a_elems = a_off;
        b->lp_elems = lp_off;
        // same-sized tables are overwritten in place: a captured solve graph holds their addresses
        auto put = [&](auto& d, const auto& h) {
            using T = typename std::decay_t<decltype(h)>::value_type;
            if (d.count != h.size()) {
                upload(b->ctx, d, h);
                b->solver.reset();
            } else if (!h.empty()) {
                copy_on_stream(b->ctx, d.ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
            }
        };
        build_families(b);  // sets in_family on the instances before they are uploaded
        put(b->d_inst, b->inst);
        put(b->d_col_inst, b->col_inst);
        put(b->d_col_q, b->col_q);
        if (b->d_dirs.count != (size_t)dir_off * 3) b->d_dirs.allocate((size_t)dir_off * 3);
        copy_on_stream(b->ctx, b->d_dirs.ptr, directions, (size_t)dir_off * 3 * sizeof(double), hipMemcpyHostToDevice);
        if (b->d_A.count != (size_t)a_off) {
            const int64_t bytes = (int64_t)sizeof(double2) * a_off;
            const int64_t old_bytes = (int64_t)b->d_A.bytes();
            b->d_A.release();
            b->solver.reset();
            release_device(b->ctx, old_bytes);
            admit_device(b->ctx, bytes, "the batch's FP64 matrices", true);
            b->d_A.allocate((size_t)a_off);
        }
        if (b->d_rhs.count != (size_t)vec_off) {
            b->d_rhs.allocate((size_t)vec_off);
            b->d_x.allocate((size_t)vec_off);
            b->solver.reset();
        }
        zero_on_stream(b->ctx, b->d_x.ptr, b->d_x.bytes());
        b->lp_format = -2;
        b->cand_m = 0;  // gathered guess candidates belong to the previous problem
        build_items(b);
        build_chunks(b);  // also drops the solver workspace and its graph
        build_tiles(b);
        b->has_problem = true;
    });
}

sbh_status sbh_batch_assemble(sbh_batch* b) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        use_device(b->ctx);
        ragged_stamp(b->ctx, b->d_t0.ptr);
        run_assembly(b);
        b->lp_format = -2;  // a new matrix invalidates the low-precision copy
    });
}

sbh_status sbh_batch_convert(sbh_batch* b, int format) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(format == 0 || format == 1, "format must be FP16 (0) or BF16 (1)");
        use_device(b->ctx);
        if (b->d_Alp.count != (size_t)b->lp_elems) {
            b->d_Alp.allocate((size_t)b->lp_elems);
            b->d_row_inv.allocate((size_t)b->total_panels);
            b->solver.reset();
        }
        batch_lp_convert(b->ctx, format, b->total_panels, b->d_row_inst.ptr, b->d_inst.ptr, b->d_A.ptr, b->d_Alp.ptr,
                         b->d_row_inv.ptr);
        b->lp_format = format;
    });
}

sbh_status sbh_batch_solve(sbh_batch* b, const sbh_batch_solve_config* cfg, sbh_batch_column_status* status) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(cfg, "config is null");
        SBH_REQUIRE(cfg->format >= -1 && cfg->format <= 1, "format must be -1, 0 or 1");
        SBH_REQUIRE(cfg->restart >= 1 && cfg->restart <= 200, "restart must be in 1..200");
        SBH_REQUIRE(cfg->max_cycles >= 1, "max_cycles must be >= 1");
        SBH_REQUIRE(cfg->tolerance > 0.0 && cfg->inner_tolerance > 0.0, "tolerances must be positive");
        if (cfg->format >= 0)
            SBH_REQUIRE(b->lp_format == cfg->format, "sbh_batch_convert with this format has not been called");
        sbh_context* ctx = b->ctx;
        use_device(ctx);
        ensure_solver(b, (int)cfg->restart);
        auto& s = *b->solver;
        const RaggedLayout lay = layout(b);
        // start: ||b|| from the residual of x = 0, then (with a guess) the guess's residual
        const int refinement = cfg->format >= 0 ? 1 : 0;
        if (!cfg->use_guess) SBH_HIP(hipMemsetAsync(b->d_x.ptr, 0, b->d_x.bytes(), ctx->stream));
        for (auto* a : {&s.iters, &s.cycles, &s.reorth_count})
            SBH_HIP(hipMemsetAsync(a->ptr, 0, a->bytes(), ctx->stream));
        SBH_HIP(hipMemsetAsync(s.t_done.ptr, 0, s.t_done.bytes(), ctx->stream));
        SBH_HIP(hipMemsetAsync(s.done.ptr, 0, s.done.bytes(), ctx->stream));
        fill_int(ctx, s.inst_pending, 1);
        ragged_fill(ctx, s.R.ptr, b->L);
        ragged_residual(ctx, lay, s.inst_pending.ptr, b->d_rhs.ptr, s.R.ptr, s.part.ptr);
        ragged_restart(ctx, lay, s.part.ptr, s.inst_pending.ptr, 1, cfg->tolerance, cfg->inner_tolerance,
                       refinement, s.bnorm.ptr, s.rnorm.ptr, s.done.ptr, s.active.ptr, s.jlen.ptr, s.g.ptr,
                       s.scale.ptr, s.tol_abs.ptr, s.cycles.ptr, s.t_done.ptr);
        if (cfg->use_guess) {
            ragged_flags(ctx, lay, s.active.ptr, s.reorth.ptr, s.done.ptr, s.inst_active.ptr, s.inst_reorth.ptr,
                         s.inst_pending.ptr);
            batch_product(ctx, -1, b->max_cols, b->ntiles, b->d_tiles.ptr, b->d_inst.ptr, s.inst_pending.ptr, b->d_A.ptr, nullptr,
                          nullptr, b->d_x.ptr, s.R.ptr);
            ragged_residual(ctx, lay, s.inst_pending.ptr, b->d_rhs.ptr, s.R.ptr, s.part.ptr);
            ragged_restart(ctx, lay, s.part.ptr, s.inst_pending.ptr, 2, cfg->tolerance, cfg->inner_tolerance,
                           refinement, s.bnorm.ptr, s.rnorm.ptr, s.done.ptr, s.active.ptr, s.jlen.ptr, s.g.ptr,
                           s.scale.ptr, s.tol_abs.ptr, s.cycles.ptr, s.t_done.ptr);
        }
        ragged_scale(ctx, lay, s.inst_pending.ptr, s.scale.ptr, s.R.ptr, s.V.ptr);
        ragged_flags(ctx, lay, s.active.ptr, s.reorth.ptr, s.done.ptr, s.inst_active.ptr, s.inst_reorth.ptr,
                     s.inst_pending.ptr);
        if (cfg->use_graph && (!s.exec || s.graph_format != cfg->format || s.graph_tol != cfg->tolerance ||
                               s.graph_inner_tol != cfg->inner_tolerance)) {
            if (s.exec) {
                SBH_HIP(hipGraphExecDestroy(s.exec));
                s.exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            SBH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            try {
                cycle(b, cfg->format, cfg->tolerance, cfg->inner_tolerance);
            } catch (...) {
                (void)hipStreamEndCapture(ctx->stream, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            SBH_HIP(hipStreamEndCapture(ctx->stream, &graph));
            hipError_t e = hipGraphInstantiate(&s.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            SBH_HIP(e);
            s.graph_format = cfg->format;
            s.graph_tol = cfg->tolerance;
            s.graph_inner_tol = cfg->inner_tolerance;
        }
        std::vector<int> pending(b->B);
        for (int64_t c = 0; c < cfg->max_cycles; ++c) {
            SBH_HIP(hipMemcpyAsync(pending.data(), s.inst_pending.ptr, pending.size() * sizeof(int),
                                   hipMemcpyDeviceToHost, ctx->stream));
            SBH_HIP(hipStreamSynchronize(ctx->stream));
            if (std::none_of(pending.begin(), pending.end(), [](int v) { return v != 0; })) break;
            if (cfg->use_graph)
                SBH_HIP(hipGraphLaunch(s.exec, ctx->stream));
            else
                cycle(b, cfg->format, cfg->tolerance, cfg->inner_tolerance);
        }
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        if (status) {
            const size_t C = (size_t)b->C;
            std::vector<int64_t> it(C), cy(C), ro(C);
            std::vector<double> rn(C), bn(C);
            std::vector<int> dn(C);
            std::vector<unsigned long long> td(C), t0(1);
            copy_on_stream(b->ctx, it.data(), s.iters.ptr, C * 8, hipMemcpyDeviceToHost);
            copy_on_stream(b->ctx, cy.data(), s.cycles.ptr, C * 8, hipMemcpyDeviceToHost);
            copy_on_stream(b->ctx, ro.data(), s.reorth_count.ptr, C * 8, hipMemcpyDeviceToHost);
            copy_on_stream(b->ctx, rn.data(), s.rnorm.ptr, C * 8, hipMemcpyDeviceToHost);
            copy_on_stream(b->ctx, bn.data(), s.bnorm.ptr, C * 8, hipMemcpyDeviceToHost);
            copy_on_stream(b->ctx, dn.data(), s.done.ptr, C * 4, hipMemcpyDeviceToHost);
            copy_on_stream(b->ctx, td.data(), s.t_done.ptr, C * 8, hipMemcpyDeviceToHost);
            copy_on_stream(b->ctx, t0.data(), b->d_t0.ptr, 8, hipMemcpyDeviceToHost);
            int rate_khz = 0;
            SBH_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
            for (size_t c = 0; c < C; ++c) {
                sbh_batch_column_status& st = status[c];
                st.iterations = it[c];
                st.cycles = cy[c];
                st.reorthogonalizations = ro[c];
                st.residual = bn[c] > 0.0 ? rn[c] / bn[c] : 0.0;
                st.converged = dn[c] && st.residual < cfg->tolerance ? 1 : 0;
                st.pad = 0;
                st.completion_ms = (dn[c] && rate_khz > 0 && td[c] >= t0[0])
                                       ? (double)(td[c] - t0[0]) / (double)rate_khz
                                       : -1.0;
            }
        }
    });
}

namespace {
sbh_status download_block(sbh_batch* b, int64_t instance, sbh_c64* out, int what) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(out && instance >= 0 && instance < b->B, "bad instance/out");
        use_device(b->ctx);
        SBH_HIP(hipStreamSynchronize(b->ctx->stream));
        const BatchInstance& I = b->inst[instance];
        const double2* src;
        size_t count;
        if (what == 0) {
            src = b->d_x.ptr + I.vec_off;
            count = (size_t)I.n * I.ncol;
        } else if (what == 1) {
            src = b->d_rhs.ptr + I.vec_off;
            count = (size_t)I.n * I.ncol;
        } else {
            src = b->d_A.ptr + I.a_off;
            count = (size_t)I.n * I.n;
        }
        copy_on_stream(b->ctx, out, src, count * sizeof(double2), hipMemcpyDeviceToHost);
    });
}
}  // namespace

sbh_status sbh_batch_set_guess(sbh_batch* b, int64_t instance, const sbh_c64* x0) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(x0 && instance >= 0 && instance < b->B, "bad instance/x0");
        use_device(b->ctx);
        SBH_HIP(hipStreamSynchronize(b->ctx->stream));
        const BatchInstance& I = b->inst[instance];
        const size_t count = (size_t)I.n * I.ncol;
        for (size_t i = 0; i < count; ++i)
            SBH_REQUIRE(std::isfinite(x0[i].re) && std::isfinite(x0[i].im), "initial guess is not finite");
        copy_on_stream(b->ctx, b->d_x.ptr + I.vec_off, x0, count * sizeof(double2), hipMemcpyHostToDevice);
    });
}

sbh_status sbh_batch_apply(sbh_batch* b, int format, const sbh_c64* X, sbh_c64* Y) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(X && Y, "X/Y is null");
        SBH_REQUIRE(format == SBH_BATCH_FP64 || format == 0 || format == 1, "format must be FP64 (-1), FP16 or BF16");
        SBH_REQUIRE(format < 0 || (b->d_Alp.count == (size_t)b->lp_elems && b->lp_format == format),
                    "the batch has no copy in that format (sbh_batch_convert)");
        use_device(b->ctx);
        const size_t L = (size_t)b->L;
        for (size_t i = 0; i < L; ++i)
            SBH_REQUIRE(std::isfinite(X[i].re) && std::isfinite(X[i].im), "X is not finite");
        DeviceArray<double2> dx(L), dy(L);
        DeviceArray<int> active((size_t)b->B);
        fill_int(b->ctx, active, 1);
        copy_on_stream(b->ctx, dx.ptr, X, L * sizeof(double2), hipMemcpyHostToDevice);
        batch_product(b->ctx, format, b->max_cols, b->ntiles, b->d_tiles.ptr, b->d_inst.ptr, active.ptr, b->d_A.ptr,
                      format < 0 ? nullptr : b->d_Alp.ptr, format < 0 ? nullptr : b->d_row_inv.ptr, dx.ptr, dy.ptr);
        copy_on_stream(b->ctx, Y, dy.ptr, L * sizeof(double2), hipMemcpyDeviceToHost);
    });
}

sbh_status sbh_batch_download_lowp(sbh_batch* b, int64_t instance, uint16_t* halves, double* row_inv, int64_t* ld) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(instance >= 0 && instance < b->B, "bad instance");
        SBH_REQUIRE(b->lp_format >= 0 && b->d_Alp.count == (size_t)b->lp_elems, "the batch has no low-precision copy");
        use_device(b->ctx);
        const BatchInstance& I = b->inst[instance];
        if (ld) *ld = I.ld;
        if (halves)
            copy_on_stream(b->ctx, halves, b->d_Alp.ptr + I.lp_off, (size_t)I.n * 2 * I.ld * sizeof(uint16_t),
                           hipMemcpyDeviceToHost);
        if (row_inv)
            copy_on_stream(b->ctx, row_inv, b->d_row_inv.ptr + I.panel_off, (size_t)I.n * sizeof(double),
                           hipMemcpyDeviceToHost);
    });
}

sbh_status sbh_batch_set_guesses(sbh_batch* b, const sbh_c64* X0) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(X0, "X0 is null");
        use_device(b->ctx);
        const size_t L = (size_t)b->L;
        for (size_t i = 0; i < L; ++i)
            SBH_REQUIRE(std::isfinite(X0[i].re) && std::isfinite(X0[i].im), "initial guess is not finite");
        copy_on_stream(b->ctx, b->d_x.ptr, X0, L * sizeof(double2), hipMemcpyHostToDevice);
    });
}

namespace {
int64_t max_block(const sbh_batch* b) {
    int64_t m = 0;
    for (const BatchInstance& I : b->inst) m = std::max<int64_t>(m, I.n * I.ncol);
    return m;
}

// A device pointer of the batch's device (or null where nulls are allowed).
void require_device_pointer(const sbh_batch* b, const void* p, bool allow_null) {
    if (!p) {
        SBH_REQUIRE(allow_null, "a device pointer is null");
        return;
    }
    hipPointerAttribute_t attr{};
    const hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) (void)hipGetLastError();
    SBH_REQUIRE(e == hipSuccess && attr.type == hipMemoryTypeDevice && attr.device == b->ctx->device,
                "a pointer is not device memory of the batch's device");
}

// Uploads a table of device pointers (checked) into b->d_ptrs.
void upload_pointer_table(sbh_batch* b, const void* const* p, size_t count, bool allow_null) {
    for (size_t i = 0; i < count; ++i) require_device_pointer(b, p[i], allow_null);
    if (b->d_ptrs.count < count) b->d_ptrs.allocate(count);
    copy_on_stream(b->ctx, b->d_ptrs.ptr, p, count * sizeof(void*), hipMemcpyHostToDevice);
}
}  // namespace

sbh_status sbh_batch_set_guesses_device(sbh_batch* b, const sbh_c64* const* x0) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(x0, "x0 is null");
        use_device(b->ctx);
        upload_pointer_table(b, reinterpret_cast<const void* const*>(x0), (size_t)b->B, true);
        batch_gather(b->ctx, b->B, b->L, max_block(b), 1, b->d_inst.ptr, b->d_ptrs.ptr, b->d_x.ptr);
    });
}

sbh_status sbh_batch_store_solutions(sbh_batch* b, sbh_c64* const* dst) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(dst, "dst is null");
        use_device(b->ctx);
        upload_pointer_table(b, reinterpret_cast<const void* const*>(dst), (size_t)b->B, true);
        batch_scatter(b->ctx, b->B, max_block(b), b->d_inst.ptr, b->d_x.ptr,
                      const_cast<double2* const*>(reinterpret_cast<const double2* const*>(b->d_ptrs.ptr)));
        SBH_HIP(hipStreamSynchronize(b->ctx->stream));
    });
}

sbh_status sbh_batch_guess_gram(sbh_batch* b, int format, int m, const sbh_c64* const* candidates, sbh_c64* gram,
                                sbh_c64* projections, double* rhs_norm2) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(candidates && gram && projections && rhs_norm2, "null argument");
        SBH_REQUIRE(m >= 1 && m <= 6, "the number of candidates must be in 1..6");
        SBH_REQUIRE(format == SBH_BATCH_FP64 || format == 0 || format == 1, "format must be FP64 (-1), FP16 or BF16");
        SBH_REQUIRE(format < 0 || (b->d_Alp.count == (size_t)b->lp_elems && b->lp_format == format),
                    "the batch has no copy in that format (sbh_batch_convert)");
        use_device(b->ctx);
        sbh_context* ctx = b->ctx;
        const size_t L = (size_t)b->L, C = (size_t)b->C;
        upload_pointer_table(b, reinterpret_cast<const void* const*>(candidates), (size_t)m * b->B, true);
        if (b->d_cand.count < (size_t)m * L) {
            b->d_cand.allocate((size_t)m * L);
            b->d_cprod.allocate((size_t)m * L);
        }
        if (b->d_ones.count != (size_t)b->B) {
            b->d_ones.allocate((size_t)b->B);
            fill_int(ctx, b->d_ones, 1);
        }
        if (b->d_G.count < C * m * m) b->d_G.allocate(C * m * m);
        if (b->d_h.count < C * m) b->d_h.allocate(C * m);
        if (b->d_bb.count < C) b->d_bb.allocate(C);
        batch_gather(ctx, b->B, b->L, max_block(b), m, b->d_inst.ptr, b->d_ptrs.ptr, b->d_cand.ptr);
        for (int a = 0; a < m; ++a)
            batch_product(ctx, format, b->max_cols, b->ntiles, b->d_tiles.ptr, b->d_inst.ptr, b->d_ones.ptr,
                          b->d_A.ptr, format < 0 ? nullptr : b->d_Alp.ptr, format < 0 ? nullptr : b->d_row_inv.ptr,
                          b->d_cand.ptr + a * L, b->d_cprod.ptr + a * L);
        batch_gram(ctx, m, (int)C, b->d_inst.ptr, b->d_col_inst.ptr, b->d_col_q.ptr, b->d_cprod.ptr, b->L,
                   b->d_rhs.ptr, b->d_G.ptr, b->d_h.ptr, b->d_bb.ptr);
        SBH_HIP(hipMemcpyAsync(gram, b->d_G.ptr, C * m * m * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        SBH_HIP(hipMemcpyAsync(projections, b->d_h.ptr, C * m * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
        SBH_HIP(hipMemcpyAsync(rhs_norm2, b->d_bb.ptr, C * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        b->cand_m = m;
        b->cand_format = format;
    });
}

sbh_status sbh_batch_guess_combine(sbh_batch* b, int m, const sbh_c64* coefficients) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(coefficients, "coefficients is null");
        SBH_REQUIRE(m >= 1 && m == b->cand_m && b->d_cand.count >= (size_t)m * b->L,
                    "sbh_batch_guess_gram with this number of candidates has not been called");
        const size_t C = (size_t)b->C;
        for (size_t i = 0; i < C * m; ++i)
            SBH_REQUIRE(std::isfinite(coefficients[i].re) && std::isfinite(coefficients[i].im),
                        "a coefficient is not finite");
        use_device(b->ctx);
        if (b->d_coef.count < C * m) b->d_coef.allocate(C * m);
        copy_on_stream(b->ctx, b->d_coef.ptr, coefficients, C * m * sizeof(double2), hipMemcpyHostToDevice);
        batch_combine(b->ctx, b->B, max_block(b), b->d_inst.ptr, b->d_cand.ptr, b->L, m, b->d_coef.ptr, b->d_x.ptr);
    });
}

sbh_status sbh_batch_learned_guesses(sbh_batch* b, sbh_mlp* net, sbh_c64* out) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(net, "net is null");
        SBH_REQUIRE(mlp_context(net) == b->ctx, "the network belongs to another context");
        SBH_REQUIRE(b->L <= (1 << 24), "the batch has more than 2^24 samples");
        if (out) require_device_pointer(b, out, false);
        use_device(b->ctx);
        const int64_t mb = max_block(b);
        float* raw = mlp_raw_workspace(net, b->L);
        batch_learned_raw(b->ctx, b->B, mb, b->d_inst.ptr, b->d_src.ptr, b->d_dirs.ptr, raw);
        const float* Y = mlp_forward_device(net, raw, b->L);
        batch_learned_phase(b->ctx, b->B, mb, b->d_inst.ptr, b->d_src.ptr, b->d_dirs.ptr, Y,
                            out ? reinterpret_cast<double2*>(out) : b->d_x.ptr);
        SBH_HIP(hipStreamSynchronize(b->ctx->stream));
    });
}

sbh_status sbh_batch_download_all(sbh_batch* b, int what, sbh_c64* out) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        SBH_REQUIRE(out, "out is null");
        SBH_REQUIRE(what == 0 || what == 1, "what must be 0 (solutions) or 1 (right-hand sides)");
        use_device(b->ctx);
        copy_on_stream(b->ctx, out, what == 0 ? b->d_x.ptr : b->d_rhs.ptr, (size_t)b->L * sizeof(double2),
                       hipMemcpyDeviceToHost);
    });
}

sbh_status sbh_batch_download_solution(sbh_batch* b, int64_t instance, sbh_c64* x) {
    return download_block(b, instance, x, 0);
}
sbh_status sbh_batch_download_rhs(sbh_batch* b, int64_t instance, sbh_c64* rhs) {
    return download_block(b, instance, rhs, 1);
}
sbh_status sbh_batch_download_matrix(sbh_batch* b, int64_t instance, sbh_c64* a) {
    return download_block(b, instance, a, 2);
}

sbh_status sbh_batch_info(const sbh_batch* b, int64_t* num_instances, int64_t* num_columns, int64_t* total_panels,
                          int64_t* matrix_entries, double* route_ms) {
    return guard_on(ctx_of(b), [&] {
        SBH_REQUIRE(b, "batch is null");
        if (num_instances) *num_instances = b->B;
        if (num_columns) *num_columns = b->C;
        if (total_panels) *total_panels = b->total_panels;
        if (matrix_entries) *matrix_entries = b->a_elems;
        if (route_ms) *route_ms = b->route_ms;
    });
}

sbh_status sbh_batch_get_tuning(const sbh_batch* b, sbh_batch_tuning* t) {
    return guard_on(ctx_of(b), [&] {
        SBH_REQUIRE(b && t, "null argument");
        *t = b->tuning;
    });
}

sbh_status sbh_batch_set_tuning(sbh_batch* b, const sbh_batch_tuning* t) {
    return guard_on(ctx_of(b), [&] {
        SBH_REQUIRE(b && t, "null argument");
        SBH_REQUIRE(t->rows_per_item >= 1 && t->rows_per_item <= 64, "rows_per_item must be in 1..64");
        SBH_REQUIRE(t->blocks_per_cu >= 1 && t->blocks_per_cu <= 16, "blocks_per_cu must be in 1..16");
        SBH_REQUIRE(t->chunk_elems >= 256 && t->chunk_elems <= 65536, "chunk_elems must be in 256..65536");
        use_device(b->ctx);
        SBH_HIP(hipStreamSynchronize(b->ctx->stream));
        const bool items = t->rows_per_item != b->tuning.rows_per_item;
        b->tuning = *t;
        b->tuning.pad = 0;
        if (b->has_problem) {
            if (items) build_items(b);
            if (b->chunks_for != b->tuning.chunk_elems) build_chunks(b);
        }
    });
}

sbh_status sbh_batch_autotune(sbh_batch* b, double* table, int max_entries, int* num_entries,
                              sbh_batch_tuning* best) {
    return guard_on(ctx_of(b), [&] {
        check_problem(b);
        sbh_context* ctx = b->ctx;
        use_device(ctx);
        SBH_HIP(hipStreamSynchronize(ctx->stream));
        int rows_out = 0;
        auto record = [&](double r, double bp, double ce, double ms) {
            if (table && rows_out < max_entries) {
                table[4 * rows_out + 0] = r;
                table[4 * rows_out + 1] = bp;
                table[4 * rows_out + 2] = ce;
                table[4 * rows_out + 3] = ms;
            }
            ++rows_out;
        };
        hipEvent_t e0, e1;
        SBH_HIP(hipEventCreate(&e0));
        SBH_HIP(hipEventCreate(&e1));
        sbh_batch_tuning t = b->tuning;
        // assembly shape
        double best_ms = INFINITY;
        sbh_batch_tuning tb = t;
        for (int rows : {4, 8, 16, 32})
            for (int bpc : {2, 4, 8}) {
                b->tuning.rows_per_item = rows;
                b->tuning.blocks_per_cu = bpc;
                build_items(b);
                SBH_HIP(hipEventRecord(e0, ctx->stream));
                run_assembly(b);
                SBH_HIP(hipEventRecord(e1, ctx->stream));
                SBH_HIP(hipEventSynchronize(e1));
                const double ms = elapsed_ms(e0, e1);
                record(rows, bpc, 0, ms);
                if (ms < best_ms) {
                    best_ms = ms;
                    tb.rows_per_item = rows;
                    tb.blocks_per_cu = bpc;
                }
            }
        b->tuning.rows_per_item = tb.rows_per_item;
        b->tuning.blocks_per_cu = tb.blocks_per_cu;
        build_items(b);
        // chunk size: one Gram-Schmidt pass (dots + projection) over 16 basis vectors
        best_ms = INFINITY;
        const int nvec = 16;
        DeviceArray<double2> V((size_t)(nvec + 1) * b->L);
        SBH_HIP(hipMemsetAsync(V.ptr, 0, V.bytes(), ctx->stream));
        DeviceArray<int> flag((size_t)b->B);
        fill_int(ctx, flag, 1);
        for (int ce : {2048, 4096, 8192, 16384}) {
            b->tuning.chunk_elems = ce;
            build_chunks(b);
            DeviceArray<double2> part((size_t)b->nchunks * (nvec + 2) * kMaxCol);
            DeviceArray<double2> coef((size_t)b->C * (nvec + 2));
            SBH_HIP(hipMemsetAsync(coef.ptr, 0, coef.bytes(), ctx->stream));
            RaggedLayout lay;
            lay.B = b->B;
            lay.C = b->C;
            lay.nchunks = b->nchunks;
            lay.MV = nvec + 2;
            lay.m = nvec;
            lay.L = b->L;
            lay.inst = b->d_inst.ptr;
            lay.chunks = b->d_chunks.ptr;
            lay.chunk_range = b->d_chunk_range.ptr;
            lay.col_inst = b->d_col_inst.ptr;
            lay.col_q = b->d_col_q.ptr;
            double2* W = V.ptr + (size_t)nvec * b->L;
            SBH_HIP(hipEventRecord(e0, ctx->stream));
            for (int rep = 0; rep < 3; ++rep) {
                ragged_dots(ctx, lay, flag.ptr, V.ptr, nvec, W, part.ptr);
                ragged_project(ctx, lay, flag.ptr, V.ptr, nvec, coef.ptr, W, part.ptr);
            }
            SBH_HIP(hipEventRecord(e1, ctx->stream));
            SBH_HIP(hipEventSynchronize(e1));
            const double ms = elapsed_ms(e0, e1) / 3.0;
            record(0, 0, ce, ms);
            if (ms < best_ms) {
                best_ms = ms;
                tb.chunk_elems = ce;
            }
        }
        SBH_HIP(hipEventDestroy(e0));
        SBH_HIP(hipEventDestroy(e1));
        b->tuning.chunk_elems = tb.chunk_elems;
        build_chunks(b);
        if (num_entries) *num_entries = rows_out;
        if (best) *best = b->tuning;
    });
}

}  // extern "C"
