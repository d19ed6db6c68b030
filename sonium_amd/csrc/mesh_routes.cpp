// Host geometry and frequency-invariant routes for the Burton–Miller assembly.
//
// Compiled with -ffp-contract=off: the near/far decisions made here must be
// the decisions the reference makes, bit for bit, so every expression follows
// the reference's evaluation order (Rust never fuses a*b+c).
//
//  * element geometry: compute_element_geometry,
//    crates/sonium-bem/src/core/mesh/generators.rs:779-870
//  * field-panel Jacobian/normal: compute_parameters,
//    crates/sonium-bem/src/core/integration/regular.rs:636-714
//  * routes: generate_subelements,
//    crates/sonium-bem/src/core/integration/singular.rs:1372-1528
//  * Gauss–Legendre rules: crates/sonium-bem/src/core/integration/gauss.rs:14-103
//
// A pair (i, j) is "near" when the level-0 test of generate_subelements
// (distance from the collocation point to the field panel's local centre over
// sqrt(area)) is below 3; every other off-diagonal pair is a single
// unsubdivided TR13 leaf, which the far kernel evaluates directly.  Near
// candidates are found with a uniform grid whose cell edge exceeds
// 3*sqrt(max area), so no pair with ratio < 3 can be missed.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "internal.hpp"

namespace sbh {
namespace {

constexpr double PI = 3.14159265358979323846;

const double CSI6[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
const double ETA6[6] = {0.0, 0.0, 1.0, 0.0, 0.5, 0.5};

// singular.rs:1563-1584 local_to_global, Tri3 shape functions.
inline void local_to_global(const double* c, double s, double t, double out[3]) {
    const double N[3] = {1.0 - s - t, s, t};
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[j] += N[i] * c[3 * i + j];
}

// singular.rs:1530-1554: first order in 4..=7 whose error estimates fall
// below 5e-4.
inline int gauss_order_for(double disfac) {
    for (int order = 4; order <= 7; ++order) {
        double n = (double)order;
        double base = disfac / (2.0 * n + 1.0);
        if (std::pow(base, 2 * order + 1) < 0.0005 && std::pow(base, 2 * order + 2) < 0.0005 &&
            std::pow(base, 2 * order + 3) < 0.0005)
            return order;
    }
    return 7;
}

struct Cell {
    double xi[3], et[3];
    uint64_t code;
};

// generate_subelements (singular.rs:1372-1528), Tri3 branch, carrying each
// cell's path code.  Returns the number of leaves (<= 110) written to codes
// (and, if verts != nullptr, their vertex coordinates).  *order_ok is false
// when a kept leaf would need a rule other than TR13.
int subdivide(const double src[3], const double* c, double area, uint64_t* codes, double* verts,
              bool* order_ok) {
    const int MAX_NSE = 60, NSE = 4, MAX_SUB = 110;
    const double TOL_F = 3.0;
    Cell sfp[MAX_NSE], sep[MAX_NSE];
    for (int v = 0; v < 3; ++v) {
        sfp[0].xi[v] = CSI6[v];
        sfp[0].et[v] = ETA6[v];
    }
    sfp[0].code = 0;
    int nsfl = 1;
    double faclin = 2.0;
    int count = 0;
    int depth = 0;
    *order_ok = true;
    for (;;) {
        int ndie = 0;
        faclin *= 0.5;
        double arels = area * faclin * faclin;
        int nsel = nsfl;
        std::memcpy(sep, sfp, sizeof(Cell) * nsel);
        for (int idi = 0; idi < nsel; ++idi) {
            const Cell& cell = sep[idi];
            double scent = (0.0 + cell.xi[0] + cell.xi[1] + cell.xi[2]) / 3.0;
            double tcent = (0.0 + cell.et[0] + cell.et[1] + cell.et[2]) / 3.0;
            double p[3];
            local_to_global(c, scent, tcent, p);
            double dx = p[0] - src[0], dy = p[1] - src[1], dz = p[2] - src[2];
            double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            double ratdis = dist / std::sqrt(arels);
            if (ratdis < TOL_F) {
                ndie += 1;
                if (ndie > 15) break;
                nsfl = ndie * NSE;
                int nsf0 = nsfl - NSE;
                double xisp[6], etsp[6];
                for (int j = 0; j < 3; ++j) {
                    int j1 = (j + 1) % 3;
                    xisp[j] = cell.xi[j];
                    xisp[j + 3] = (cell.xi[j] + cell.xi[j1]) / 2.0;
                    etsp[j] = cell.et[j];
                    etsp[j + 3] = (cell.et[j] + cell.et[j1]) / 2.0;
                }
                for (int j = 0; j < 3; ++j) {
                    Cell& ch = sfp[nsf0 + j];
                    int j1 = j + 3;
                    int j2 = (j == 0) ? 5 : (j - 1 + 3);
                    ch.xi[0] = xisp[j]; ch.xi[1] = xisp[j1]; ch.xi[2] = xisp[j2];
                    ch.et[0] = etsp[j]; ch.et[1] = etsp[j1]; ch.et[2] = etsp[j2];
                    ch.code = cell.code | ((uint64_t)j << (2 * depth));
                }
                Cell& cc = sfp[nsf0 + NSE - 1];
                cc.xi[0] = xisp[3]; cc.xi[1] = xisp[4]; cc.xi[2] = xisp[5];
                cc.et[0] = etsp[3]; cc.et[1] = etsp[4]; cc.et[2] = etsp[5];
                cc.code = cell.code | ((uint64_t)3 << (2 * depth));
            } else {
                if (gauss_order_for(0.5 / ratdis) != 4) *order_ok = false;
                codes[count] = cell.code | ((uint64_t)depth << LEAF_DEPTH_SHIFT);
                if (verts) {
                    for (int v = 0; v < 3; ++v) {
                        verts[6 * count + 2 * v] = cell.xi[v];
                        verts[6 * count + 2 * v + 1] = cell.et[v];
                    }
                }
                ++count;
                if (count >= MAX_SUB) return count;
            }
        }
        if (ndie == 0) break;
        depth += 1;
        if (depth >= 29) {
            // The path code holds 29 levels; deeper trees cannot occur for a
            // collocation point off the panel (the split stops once cells
            // are 3 sqrt(area) away), so treat this as malformed input.
            *order_ok = false;
            return count;
        }
    }
    return count;
}

// Level-0 ratio of generate_subelements for the whole element.
inline double level0_ratio(const double src[3], const double* c, double area) {
    double scent = (0.0 + CSI6[0] + CSI6[1] + CSI6[2]) / 3.0;
    double tcent = (0.0 + ETA6[0] + ETA6[1] + ETA6[2]) / 3.0;
    double p[3];
    local_to_global(c, scent, tcent, p);
    double dx = p[0] - src[0], dy = p[1] - src[1], dz = p[2] - src[2];
    double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
    return dist / std::sqrt(area * 1.0 * 1.0);
}

}  // namespace

// gauss.rs:14-103: tables for 1, 2, 4, 6, 8; Newton-generated otherwise.
void gauss_legendre_rule(int order, double* x, double* w) {
    static const double X1[] = {0.0}, W1[] = {2.0};
    static const double X2[] = {-0.5773502691896257, 0.5773502691896257}, W2[] = {1.0, 1.0};
    static const double X4[] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563,
                                0.8611363115940526};
    static const double W4[] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461,
                                0.3478548451374538};
    static const double X6[] = {-0.9324695142031521, -0.6612093864662645, -0.2386191860831969,
                                0.2386191860831969,  0.6612093864662645,  0.9324695142031521};
    static const double W6[] = {0.1713244923791704, 0.3607615730481386, 0.4679139345726910,
                                0.4679139345726910, 0.3607615730481386, 0.1713244923791704};
    static const double X8[] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                                -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                                0.7966664774136267,  0.9602898564975363};
    static const double W8[] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                                0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                                0.2223810344533745, 0.1012285362903763};
    const double* tx = nullptr;
    const double* tw = nullptr;
    switch (order) {
        case 1: tx = X1; tw = W1; break;
        case 2: tx = X2; tw = W2; break;
        case 4: tx = X4; tw = W4; break;
        case 6: tx = X6; tw = W6; break;
        case 8: tx = X8; tw = W8; break;
        default: break;
    }
    if (tx) {
        std::memcpy(x, tx, sizeof(double) * order);
        std::memcpy(w, tw, sizeof(double) * order);
        return;
    }
    // gauss.rs:38-83
    int half = (order + 1) / 2;
    double order_f = (double)order;
    for (int index = 0; index < half; ++index) {
        double root = std::cos(PI * ((double)index + 0.75) / (order_f + 0.5));
        for (int it = 0; it < 32; ++it) {
            double p_prev = 1.0, p = root;
            for (int degree = 2; degree <= order; ++degree) {
                double p_next =
                    ((double)(2 * degree - 1) * root * p - (double)(degree - 1) * p_prev) / (double)degree;
                p_prev = p;
                p = p_next;
            }
            double derivative = order_f * (root * p - p_prev) / (root * root - 1.0);
            double next = root - p / derivative;
            if (std::fabs(next - root) <= 4.0 * 2.220446049250313e-16 * std::max(std::fabs(next), 1.0)) {
                root = next;
                break;
            }
            root = next;
        }
        double p_prev = 1.0, p = root;
        for (int degree = 2; degree <= order; ++degree) {
            double p_next =
                ((double)(2 * degree - 1) * root * p - (double)(degree - 1) * p_prev) / (double)degree;
            p_prev = p;
            p = p_next;
        }
        double derivative = order_f * (root * p - p_prev) / (root * root - 1.0);
        double weight = 2.0 / ((1.0 - root * root) * derivative * derivative);
        x[index] = -root;
        x[order - 1 - index] = root;
        w[index] = weight;
        w[order - 1 - index] = weight;
    }
}

int subelements_for_pair(const double* src, const double* coords, double area, double* leaves_out) {
    uint64_t codes[110];
    bool ok = true;
    return subdivide(src, coords, area, codes, leaves_out, &ok);
}

void build_mesh_geometry(sbh_mesh* m) {
    const int64_t n = m->n;
    m->centers.assign(3 * n, 0.0);
    m->normals.assign(3 * n, 0.0);
    m->areas.assign(n, 0.0);
    m->geom.assign(n, PanelGeom{});
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t v = 0; v < m->num_nodes; ++v)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], m->nodes[3 * v + d]);
            hi[d] = std::max(hi[d], m->nodes[3 * v + d]);
        }
    m->diameter = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                            (hi[2] - lo[2]) * (hi[2] - lo[2]));
    for (int64_t e = 0; e < n; ++e) {
        const double* p[3];
        for (int v = 0; v < 3; ++v) p[v] = &m->nodes[3 * m->tris[3 * e + v]];
        double* cen = &m->centers[3 * e];
        double* nrm = &m->normals[3 * e];
        // generators.rs:790-800 centre
        for (int j = 0; j < 3; ++j) cen[j] = 0.0;
        for (int v = 0; v < 3; ++v)
            for (int j = 0; j < 3; ++j) cen[j] += p[v][j];
        for (int j = 0; j < 3; ++j) cen[j] /= 3.0;
        // generators.rs:803-828 normal and area
        double v1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        double v2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                        v1[0] * v2[1] - v1[1] * v2[0]};
        double len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        m->areas[e] = len / 2.0;
        if (len > 1e-15) {
            nrm[0] = cr[0] / len;
            nrm[1] = cr[1] / len;
            nrm[2] = cr[2] / len;
        }
        // generators.rs:857-866 outward flip; project meshes keep the vertex
        // order (compute_element_normal, element.rs:368-392, used by load_project)
        double ndc = nrm[0] * cen[0] + nrm[1] * cen[1] + nrm[2] * cen[2];
        if (!m->vertex_order_normals && ndc < 0.0) {
            nrm[0] = -nrm[0];
            nrm[1] = -nrm[1];
            nrm[2] = -nrm[2];
        }
        PanelGeom& g = m->geom[e];
        for (int v = 0; v < 3; ++v)
            for (int j = 0; j < 3; ++j) g.c[3 * v + j] = p[v][j];
        // regular.rs:687-704: dx/ds = v1 - v0, dx/dt = v2 - v0, n = ds x dt, J = |n|
        double ds[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        double dt[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        double na[3] = {ds[1] * dt[2] - ds[2] * dt[1], ds[2] * dt[0] - ds[0] * dt[2],
                        ds[0] * dt[1] - ds[1] * dt[0]};
        double J = std::sqrt(na[0] * na[0] + na[1] * na[1] + na[2] * na[2]);
        g.jac = J;
        if (J > 1e-15) {
            g.nf[0] = na[0] / J;
            g.nf[1] = na[1] / J;
            g.nf[2] = na[2] / J;
        }
        g.area = m->areas[e];
        // singular.rs:1615-1640 estimate_element_size
        double area2 = 0.5 * std::sqrt(na[0] * na[0] + na[1] * na[1] + na[2] * na[2]);
        g.size = std::max(std::sqrt(area2), 1e-10);
    }
}

RouteTable build_routes(const sbh_mesh* m, int threads) {
    const int64_t n = m->n;
    RouteTable rt;
    rt.row_ptr.assign(n + 1, 0);
    if (n == 0) return rt;
    // Level-0 local centres (the points generate_subelements tests first).
    std::vector<double> lc(3 * n);
    double max_area = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double* c = m->geom[j].c;
        double scent = (0.0 + CSI6[0] + CSI6[1] + CSI6[2]) / 3.0;
        double tcent = (0.0 + ETA6[0] + ETA6[1] + ETA6[2]) / 3.0;
        local_to_global(c, scent, tcent, &lc[3 * j]);
        max_area = std::max(max_area, m->areas[j]);
    }
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t j = 0; j < n; ++j)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], std::min(lc[3 * j + d], m->centers[3 * j + d]));
            hi[d] = std::max(hi[d], std::max(lc[3 * j + d], m->centers[3 * j + d]));
        }
    // Cell edge strictly above the largest near radius 3*sqrt(area_j).
    double h = 3.0 * std::sqrt(max_area) * (1.0 + 1e-6) + 1e-300;
    int64_t dims[3];
    for (;;) {
        int64_t total = 1;
        for (int d = 0; d < 3; ++d) {
            dims[d] = std::max<int64_t>(1, (int64_t)std::floor((hi[d] - lo[d]) / h) + 1);
            total *= dims[d];
        }
        if (total <= 8 * n + 1024) break;
        h *= 1.5;
    }
    auto cell_of = [&](const double* p, int64_t* ijk) {
        for (int d = 0; d < 3; ++d) {
            int64_t v = (int64_t)std::floor((p[d] - lo[d]) / h);
            ijk[d] = std::min(std::max<int64_t>(v, 0), dims[d] - 1);
        }
    };
    const int64_t ncell = dims[0] * dims[1] * dims[2];
    std::vector<int64_t> cstart(ncell + 1, 0);
    std::vector<int64_t> cidx(n);
    std::vector<int64_t> cell_id(n);
    for (int64_t j = 0; j < n; ++j) {
        int64_t ijk[3];
        cell_of(&lc[3 * j], ijk);
        cell_id[j] = (ijk[2] * dims[1] + ijk[1]) * dims[0] + ijk[0];
        cstart[cell_id[j] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    {
        std::vector<int64_t> fill(cstart.begin(), cstart.end() - 1);
        for (int64_t j = 0; j < n; ++j) cidx[fill[cell_id[j]]++] = j;
    }
    // Per-row results, built in parallel, concatenated in row order.
    struct RowOut {
        std::vector<int32_t> cols;
        std::vector<int32_t> counts;
        std::vector<uint64_t> codes;
    };
    std::vector<RowOut> rows(n);
    std::atomic<int64_t> next(0);
    std::atomic<bool> bad(false);
    auto work = [&]() {
        std::vector<int64_t> cand;
        uint64_t codes[110];
        for (;;) {
            int64_t i = next.fetch_add(16);
            if (i >= n) break;
            int64_t iend = std::min(n, i + 16);
            for (; i < iend; ++i) {
                const double* src = &m->centers[3 * i];
                int64_t ijk[3];
                cell_of(src, ijk);
                cand.clear();
                for (int64_t dz = -1; dz <= 1; ++dz)
                    for (int64_t dy = -1; dy <= 1; ++dy)
                        for (int64_t dx = -1; dx <= 1; ++dx) {
                            int64_t x = ijk[0] + dx, y = ijk[1] + dy, z = ijk[2] + dz;
                            if (x < 0 || y < 0 || z < 0 || x >= dims[0] || y >= dims[1] || z >= dims[2])
                                continue;
                            int64_t c = (z * dims[1] + y) * dims[0] + x;
                            for (int64_t q = cstart[c]; q < cstart[c + 1]; ++q) cand.push_back(cidx[q]);
                        }
                std::sort(cand.begin(), cand.end());
                RowOut& ro = rows[i];
                for (int64_t j : cand) {
                    if (j == i) continue;
                    const PanelGeom& g = m->geom[j];
                    if (!(level0_ratio(src, g.c, g.area) < 3.0)) continue;
                    bool ok = true;
                    int cnt = subdivide(src, g.c, g.area, codes, nullptr, &ok);
                    if (!ok) bad = true;
                    ro.cols.push_back((int32_t)j);
                    ro.counts.push_back(cnt);
                    ro.codes.insert(ro.codes.end(), codes, codes + cnt);
                }
            }
        }
    };
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = (int)std::min<int64_t>(nt, std::max<int64_t>(1, n / 64));
    if (nt <= 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t) pool.emplace_back(work);
        for (auto& th : pool) th.join();
    }
    if (bad)
        throw Error(SBH_ERR_UNSUPPORTED,
                    "near-pair subdivision produced a leaf that needs a Gauss order above 4 or "
                    "more than 29 levels (degenerate geometry)");
    int64_t npairs = 0, nleaves = 0;
    for (int64_t i = 0; i < n; ++i) {
        npairs += (int64_t)rows[i].cols.size();
        nleaves += (int64_t)rows[i].codes.size();
    }
    rt.pairs.reserve(npairs);
    rt.leaves.reserve(nleaves);
    for (int64_t i = 0; i < n; ++i) {
        const RowOut& ro = rows[i];
        for (size_t q = 0; q < ro.cols.size(); ++q) {
            NearPair p;
            p.i = (int32_t)i;
            p.j = ro.cols[q];
            p.count = ro.counts[q];
            p.pad = 0;
            p.begin = (int64_t)rt.leaves.size();
            int64_t off = 0;
            for (size_t r = 0; r < q; ++r) off += ro.counts[r];
            rt.leaves.insert(rt.leaves.end(), ro.codes.begin() + off, ro.codes.begin() + off + ro.counts[q]);
            rt.pairs.push_back(p);
        }
        rt.row_ptr[i + 1] = (int64_t)rt.pairs.size();
    }
    return rt;
}

}  // namespace sbh
