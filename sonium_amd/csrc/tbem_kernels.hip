// Dense Burton–Miller assembly on gfx950 (FP64).
//
// Four kernels fill the N x N row-major complex matrix of
// build_tbem_system_with_beta_parallel (crates/sonium-bem/src/core/assembly/
// tbem.rs:1284-1466) for one wavenumber:
//
//   far_kernel   every off-diagonal pair with the 13-point rule on the whole
//                field panel (the unsubdivided leaf of generate_subelements,
//                regular.rs:124-321).  One lane per field panel (column),
//                rows streamed through the workgroup: the lane keeps its
//                panel in registers, the collocation point and normal are
//                uniform (scalar loads), and 64 lanes store 1 KiB of a row.
//                This is the FP64-VALU-bound kernel: 13 kernel evaluations
//                per pair, each one rsqrt, one sincos (Cody–Waite reduction +
//                Taylor polynomials) and ~75 FMA-class operations.
//   near_kernel  overwrites the near pairs (level-0 ratio < 3) with the sum
//                over their subdivision leaves; one wave per pair, lanes over
//                (leaf, point), evaluated in the reference's operation order.
//   self_kernel  overwrites the diagonal: Stokes edge integral of the
//                hypersingular kernel plus Duffy-collapsed Gauss–Legendre on
//                the 3 x nsec2 sub-triangles (singular.rs:297-578), the
//                QuadratureParams::for_ka tier (singular.rs:51-85) chosen per
//                panel from k*size, and the free terms (tbem.rs:1330-1358).
//   bc_rhs_kernel  only when some field panel carries a nonzero BC value:
//                the sum over such columns of tbem.rs:1419-1435.
//
// plane_wave_kernel adds the incident-field right-hand side
// (incident.rs:807-834 with bem_solver.rs:5005-5034).

#include <cmath>

#include "internal.hpp"

namespace sbh {
namespace {

constexpr double FOUR_PI = 12.566370614359172;  // 4.0 * PI, exact product

// TR13 (gauss.rs:526-540); weights are used times 0.5 (gauss.rs:117-130).
__constant__ double kTrS[13] = {0.333333333333333, 0.260345966079040, 0.260345966079040,
                                0.479308067841920, 0.065130102902216, 0.065130102902216,
                                0.869739794195568, 0.638444188569810, 0.048690315425316,
                                0.638444188569810, 0.312865496004874, 0.048690315425316,
                                0.312865496004874};
__constant__ double kTrT[13] = {0.333333333333333, 0.260345966079040, 0.479308067841920,
                                0.260345966079040, 0.065130102902216, 0.869739794195568,
                                0.065130102902216, 0.048690315425316, 0.638444188569810,
                                0.312865496004874, 0.638444188569810, 0.312865496004874,
                                0.048690315425316};
__constant__ double kTrW[13] = {-0.149570044467682, 0.175615257433208, 0.175615257433208,
                                0.175615257433208, 0.053347235608838, 0.053347235608838,
                                0.053347235608838, 0.077113760890257, 0.077113760890257,
                                0.077113760890257, 0.077113760890257, 0.077113760890257,
                                0.077113760890257};

// TR7 (gauss.rs:516-524) for field evaluation; weights are used times 0.5.
__constant__ double kTr7S[7] = {0.333333333333333, 0.797426985353087, 0.101286507323456, 0.101286507323456,
                                0.470142064105115, 0.059715871789770, 0.470142064105115};
__constant__ double kTr7T[7] = {0.333333333333333, 0.101286507323456, 0.797426985353087, 0.101286507323456,
                                0.059715871789770, 0.470142064105115, 0.470142064105115};
__constant__ double kTr7W[7] = {0.225, 0.125939180544827, 0.125939180544827, 0.125939180544827,
                                0.132394152788506, 0.132394152788506, 0.132394152788506};

// Local vertex and mid-side coordinates (singular.rs:119-124 CSI6/ETA6).
__constant__ double kCsi6[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
__constant__ double kEta6[6] = {0.0, 0.0, 1.0, 0.0, 0.5, 0.5};

// Gauss–Legendre nodes/weights for orders 1..8 (row = order), uploaded once
// per device from the host restatement of gauss.rs:14-103.
__constant__ double kGlX[9][8];
__constant__ double kGlW[9][8];

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// a * b + c with the uniform addend c read from an SGPR pair.  Written with
// __builtin_fma the compiler copies each Horner constant into a VGPR pair
// before the FMA; the three-address form with an SGPR source saves those
// moves and registers (the far kernel runs about 12% faster on 20480 panels).
// Pure arithmetic: no memory operand, so nothing is left in flight.
__device__ __forceinline__ double fma_vvs(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// sin and cos of x for |x| < 2^20.  Cody & Waite (1980) two-constant
// reduction by pi with FMA, x = q pi + r, |r| <= pi/2; then
// sin r = r + r^3 P(r^2) and cos r = 1 - r^2/2 + r^4 Q(r^2) with P, Q of
// degree 7 interpolating at Chebyshev nodes in r^2 (gen_sincos_coefficients.py;
// error below 3.6e-17 for sin and 1.6e-17 for cos), and both signs flipped
// for odd q.
__device__ __forceinline__ void sincos_red(double x, double* sp, double* cp) {
    constexpr double ONE_OVER_PI = 0.3183098861837907;
    constexpr double PI_HI = 3.141592653589793;
    constexpr double PI_LO = 1.2246467991473532e-16;
    const double q = __builtin_rint(x * ONE_OVER_PI);
    double r = __builtin_fma(-q, PI_HI, x);
    r = __builtin_fma(-q, PI_LO, r);
    const double r2 = r * r;
    constexpr double P0 = -0.16666666666666666, P1 = 0.008333333333333316, P2 = -0.00019841269841254974,
                     P3 = 2.7557319219163234e-06, P4 = -2.5052107616996182e-08, P5 = 1.6058977312464087e-10,
                     P6 = -7.643970296798572e-13, P7 = 2.7314447669863995e-15;
    constexpr double Q0 = 0.041666666666666664, Q1 = -0.001388888888888888, Q2 = 2.480158730157985e-05,
                     Q3 = -2.7557319221569563e-07, Q4 = 2.0876756602875463e-09, Q5 = -1.1470712272749466e-11,
                     Q6 = 4.7778779252458453e-14, Q7 = -1.5218638369243573e-16;
    double ps = __builtin_fma(r2, P7, P6);
    ps = fma_vvs(r2, ps, P5);
    ps = fma_vvs(r2, ps, P4);
    ps = fma_vvs(r2, ps, P3);
    ps = fma_vvs(r2, ps, P2);
    ps = fma_vvs(r2, ps, P1);
    ps = fma_vvs(r2, ps, P0);
    const double sr = __builtin_fma(r * r2, ps, r);
    double pc = __builtin_fma(r2, Q7, Q6);
    pc = fma_vvs(r2, pc, Q5);
    pc = fma_vvs(r2, pc, Q4);
    pc = fma_vvs(r2, pc, Q3);
    pc = fma_vvs(r2, pc, Q2);
    pc = fma_vvs(r2, pc, Q1);
    pc = fma_vvs(r2, pc, Q0);
    const double cr = __builtin_fma(r2 * r2, pc, __builtin_fma(-0.5, r2, 1.0));
    // (-1)^q: bit 0 of q moved to the sign bit of both results
    const unsigned long long sign = (unsigned long long)((unsigned)(int)q & 1u) << 63;
    *sp = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, sr) ^ sign);
    *cp = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, cr) ^ sign);
}

// 1/sqrt(d2) to about 1 ulp: the hardware estimate (24 correct bits) and one
// third-order (Householder) step, y1 = y0 (1 + e/2 + 3e^2/8) with
// e = 1 - d2 y0^2, whose truncation error (5/16) e^3 is below 1e-22.
__device__ __forceinline__ double rsqrt_nr(double d2) {
    const double y = __builtin_amdgcn_rsq(d2);
    const double e = __builtin_fma(-d2 * y, y, 1.0);
    return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}

struct Acc {
    double2 g, h, ht, e;
};

__device__ __forceinline__ Acc acc_zero() {
    Acc a;
    a.g = a.h = a.ht = a.e = make_double2(0.0, 0.0);
    return a;
}

// tbem.rs:1404-1417: matrix coefficient of a field panel for its BC.
__device__ __forceinline__ double2 coefficient(const Acc& r, const PanelBc& bc, double2 gamma,
                                               double2 tau, double2 beta, double2 iwr) {
    double2 c = make_double2(0.0, 0.0);
    if (bc.type == 0) {
        c = cadd(cmul(cmul(r.h, gamma), tau), cmul(r.e, beta));
    } else if (bc.type == 1) {
        double2 t = cadd(cmul(cmul(r.g, gamma), tau), cmul(r.ht, beta));
        c = make_double2(-t.x, -t.y);
    }
    if (bc.type == 0 && bc.has_adm) {
        double2 t = cadd(cmul(cmul(r.g, gamma), tau), cmul(r.ht, beta));
        c = csub(c, cmul(cmul(t, make_double2(bc.adm[0], bc.adm[1])), iwr));
    }
    return c;
}

// tbem.rs:1419-1435: BC right-hand-side contribution of a field panel.
__device__ __forceinline__ double2 bc_contribution(const Acc& r, const PanelBc& bc, double2 gamma,
                                                   double2 tau, double2 beta, double2 iwr) {
    double2 v = make_double2(bc.value[0], bc.value[1]);
    if (bc.type == 0) return cmul(cmul(cadd(cmul(cmul(r.g, gamma), tau), cmul(r.ht, beta)), v), iwr);
    if (bc.type == 1) return cmul(cadd(cmul(cmul(r.h, gamma), tau), cmul(r.e, beta)), v);
    return make_double2(0.0, 0.0);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void wave_sum_acc(Acc& a) {
    a.g.x = wave_sum(a.g.x); a.g.y = wave_sum(a.g.y);
    a.h.x = wave_sum(a.h.x); a.h.y = wave_sum(a.h.y);
    a.ht.x = wave_sum(a.ht.x); a.ht.y = wave_sum(a.ht.y);
    a.e.x = wave_sum(a.e.x); a.e.y = wave_sum(a.e.y);
}

// ---------------------------------------------------------------------------
// One quadrature point of the regular kernel, regular.rs:214-281.  DERIV also
// accumulates the analytic wavenumber derivative of every channel into dr
// (regular.rs:229-278), from the same point, as the reference's shared
// quadrature loop does.  The point evaluation uses 1/r from rsqrt_nr and FMA
// contraction (the reference divides by r nine times); the quadrature rules,
// subdivisions and channel formulas are the reference's.  The surrounding
// code keeps the reference's operation order (no contraction).
#pragma clang fp contract(off)

template <bool DERIV>
__device__ inline void ref_point_t(const double pos[3], double wga, const double src[3],
                                   const double srcn[3], const double nf[3], double wavruim, double k2,
                                   double k, double hf, Acc& r, Acc& dr) {
#pragma clang fp contract(on)
    const double d0 = pos[0] - src[0], d1 = pos[1] - src[1], d2 = pos[2] - src[2];
    const double dsq = d0 * d0 + d1 * d1 + d2 * d2;
    if (!(dsq > 1e-30)) return;  // |d| <= 1e-15: normalize() returns 0, then `dis < 1e-15` skips
    const double inv = rsqrt_nr(dsq);
    const double dis = dsq * inv;
    const double u0 = d0 * inv, u1 = d1 * inv, u2 = d2 * inv;
    const double re1 = wavruim * dis;
    const double re2 = (wga * (1.0 / FOUR_PI)) * inv;
    double s, c;
    sincos_red(re1, &s, &c);
    const double2 zg = make_double2(c * re2, s * re2);
    const double2 z1 = make_double2(-inv, wavruim);
    const double2 zb = cmul(zg, z1);
    const double re1_h = u0 * nf[0] + u1 * nf[1] + u2 * nf[2];
    const double re2_h = -(u0 * srcn[0] + u1 * srcn[1] + u2 * srcn[2]);
    const double rq = re1_h * re2_h;
    const double nxny = srcn[0] * nf[0] + srcn[1] * nf[1] + srcn[2] * nf[2];
    const double inv2 = inv * inv;
    const double t3 = 3.0 * rq + nxny;
    const double2 zef = make_double2((3.0 * inv2 - k2) * rq + nxny * inv2, -wavruim * inv * t3);
    r.g = cadd(r.g, zg);
    r.h = cadd(r.h, cscale(zb, re1_h));
    r.ht = cadd(r.ht, cscale(zb, re2_h));
    r.e = cadd(r.e, cmul(zg, zef));
    if (DERIV) {
        const double2 dzg = cmul(zg, make_double2(0.0, hf * dis));
        const double2 dzb = cadd(cmul(dzg, z1), cmul(zg, make_double2(0.0, hf)));
        const double2 dzef = make_double2(-2.0 * k * rq, -hf * inv * t3);
        dr.g = cadd(dr.g, dzg);
        dr.h = cadd(dr.h, cscale(dzb, re1_h));
        dr.ht = cadd(dr.ht, cscale(dzb, re2_h));
        dr.e = cadd(dr.e, cadd(cmul(dzg, zef), cmul(zg, dzef)));
    }
}

__device__ inline void ref_point(const double pos[3], double wga, const double src[3], const double srcn[3],
                                 const double nf[3], double wavruim, double k2, Acc& r) {
    ref_point_t<false>(pos, wga, src, srcn, nf, wavruim, k2, 0.0, 0.0, r, r);
}

// Shape-function position of local (s, t) (regular.rs:682-690).
__device__ inline void shape_pos(const double* c, double s, double t, double p[3]) {
    double N0 = 1.0 - s - t;
    p[0] = 0.0 + N0 * c[0] + s * c[3] + t * c[6];
    p[1] = 0.0 + N0 * c[1] + s * c[4] + t * c[7];
    p[2] = 0.0 + N0 * c[2] + s * c[5] + t * c[8];
}

// Vertices of the leaf with path `code` (midpoint arithmetic is exact).
__device__ inline void decode_leaf(uint64_t code, double xi[3], double et[3]) {
    xi[0] = 0.0; xi[1] = 1.0; xi[2] = 0.0;
    et[0] = 0.0; et[1] = 0.0; et[2] = 1.0;
    int depth = (int)(code >> LEAF_DEPTH_SHIFT);
    for (int l = 0; l < depth; ++l) {
        int ch = (int)((code >> (2 * l)) & 3u);
        double mx[3] = {(xi[0] + xi[1]) / 2.0, (xi[1] + xi[2]) / 2.0, (xi[2] + xi[0]) / 2.0};
        double me[3] = {(et[0] + et[1]) / 2.0, (et[1] + et[2]) / 2.0, (et[2] + et[0]) / 2.0};
        if (ch == 3) {
            xi[0] = mx[0]; xi[1] = mx[1]; xi[2] = mx[2];
            et[0] = me[0]; et[1] = me[1]; et[2] = me[2];
        } else {
            int cm = (ch + 2) % 3;
            double x0 = xi[ch], e0 = et[ch];
            xi[0] = x0; xi[1] = mx[ch]; xi[2] = mx[cm];
            et[0] = e0; et[1] = me[ch]; et[2] = me[cm];
        }
    }
}

// Point q of a subdivided leaf (regular.rs:164-196): affine map of the TR13
// point into the leaf, weight times |det|, then compute_parameters.
template <bool DERIV>
__device__ inline void near_point_t(const PanelGeom& G, uint64_t code, int q, const double src[3],
                                    const double srcn[3], double wavruim, double k2, double k, double hf, Acc& r,
                                    Acc& dr) {
    double xi[3], et[3];
    decode_leaf(code, xi, et);
    double cs = kTrS[q], ec = kTrT[q], wg = kTrW[q] * 0.5;
    double l0 = 1.0 - cs - ec;
    double xio = xi[0] * l0 + xi[1] * cs + xi[2] * ec;
    double eto = et[0] * l0 + et[1] * cs + et[2] * ec;
    double dx1 = xi[1] - xi[0], dy1 = et[1] - et[0];
    double dx2 = xi[2] - xi[0], dy2 = et[2] - et[0];
    double det = fabs(dx1 * dy2 - dx2 * dy1);
    double weih2 = wg * det;
    double pos[3];
    shape_pos(G.c, xio, eto, pos);
    ref_point_t<DERIV>(pos, weih2 * G.jac, src, srcn, G.nf, wavruim, k2, k, hf, r, dr);
}

__device__ inline void near_point(const PanelGeom& G, uint64_t code, int q, const double src[3],
                                  const double srcn[3], double wavruim, double k2, Acc& r) {
    near_point_t<false>(G, code, q, src, srcn, wavruim, k2, 0.0, 0.0, r, r);
}

// Full TR13 on the unsubdivided panel in reference order (bc_rhs_kernel).
template <bool DERIV>
__device__ inline void full_point_t(const PanelGeom& G, int q, const double src[3], const double srcn[3],
                                    double wavruim, double k2, double k, double hf, Acc& r, Acc& dr) {
    double pos[3];
    shape_pos(G.c, kTrS[q], kTrT[q], pos);
    ref_point_t<DERIV>(pos, (kTrW[q] * 0.5) * G.jac, src, srcn, G.nf, wavruim, k2, k, hf, r, dr);
}

__device__ inline void full_point(const PanelGeom& G, int q, const double src[3], const double srcn[3],
                                  double wavruim, double k2, Acc& r) {
    full_point_t<false>(G, q, src, srcn, wavruim, k2, 0.0, 0.0, r, r);
}

struct SelfTier {
    int ngpo1, ngausin, nsec1, nsec2;
};

// singular.rs:51-85 QuadratureParams::for_ka
__device__ inline SelfTier tier_for(double ka) {
    if (ka < 0.3) return {3, 4, 4, 2};
    if (ka < 1.0) return {4, 5, 6, 2};
    if (ka < 2.0) return {5, 6, 8, 3};
    return {6, 7, 10, 4};
}

__device__ inline int self_samples(const SelfTier& t) {
    return 3 * t.nsec1 * t.ngpo1 + 3 * t.nsec2 * t.ngausin * t.ngausin;
}

// Sample s of the self-panel rule (singular.rs:322-410 edge part, :412-578
// bulk part).  Edge samples add to E only; bulk samples to G, H, Ht and
// E += k^2 G (n_x . n_y).  The sub-triangle loop repeats the second
// sub-triangle for every isec >= 1, as the reference does.  DERIV also adds
// the wavenumber derivatives (singular.rs:380-386, 509-545) to dr.
template <bool DERIV>
__device__ inline void self_sample_t(const PanelGeom& G, const SelfTier& t, int s, const double src[3],
                                     const double srcn[3], double wavruim, double k2, double k, double hf, Acc& r,
                                     Acc& dr) {
    const double* c = G.c;
    const int per_edge = t.nsec1 * t.ngpo1;
    const int nedge = 3 * per_edge;
    if (s < nedge) {
        int ieg = s / per_edge;
        int rem = s - ieg * per_edge;
        int isec = rem / t.ngpo1;
        int ig = rem - isec * t.ngpo1;
        int ig1 = (ieg + 1) % 3;
        double dpoi[3];
        double leneg = 0.0;
        for (int i = 0; i < 3; ++i) {
            dpoi[i] = c[3 * ig1 + i] - c[3 * ieg + i];
            leneg += dpoi[i] * dpoi[i];
        }
        leneg = sqrt(leneg);
        double dpoo[3] = {dpoi[0] / leneg, dpoi[1] / leneg, dpoi[2] / leneg};
        double leneg_scaled = leneg / (2.0 * (double)t.nsec1);
        double delsec = 2.0 / (double)t.nsec1;
        double secmid = -1.0 - delsec / 2.0;
        for (int k = 0; k <= isec; ++k) secmid += delsec;
        double sga = secmid + kGlX[t.ngpo1][ig] / (double)t.nsec1;
        double wga = kGlW[t.ngpo1][ig] * leneg_scaled;
        double diff[3];
        for (int i = 0; i < 3; ++i) {
            double crd = c[3 * ieg + i] + dpoi[i] * (sga + 1.0) / 2.0;
            diff[i] = crd - src[i];
        }
        double dis = sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
        if (!(dis > 1e-15)) return;
        double u[3] = {diff[0] / dis, diff[1] / dis, diff[2] / dis};
        double re1 = wavruim * dis;
        double re2 = FOUR_PI * dis;
        double sn, cs;
        sincos_red(re1, &sn, &cs);
        double2 zg = make_double2(cs / re2, sn / re2);
        double2 z1 = make_double2(-1.0 / dis, wavruim);
        double2 zgf = cmul(zg, z1);
        double2 zd0 = cscale(zgf, u[0]), zd1 = cscale(zgf, u[1]), zd2 = cscale(zgf, u[2]);
        double2 w0 = csub(cscale(zd1, dpoo[2]), cscale(zd2, dpoo[1]));
        double2 w1 = csub(cscale(zd2, dpoo[0]), cscale(zd0, dpoo[2]));
        double2 w2 = csub(cscale(zd0, dpoo[1]), cscale(zd1, dpoo[0]));
        double2 dot = cadd(cadd(cscale(w0, srcn[0]), cscale(w1, srcn[1])), cscale(w2, srcn[2]));
        r.e = cadd(r.e, cscale(dot, wga));
        if (DERIV) {
            double2 dzg = cmul(zg, make_double2(0.0, hf * dis));
            double2 zgd = cadd(cmul(dzg, z1), cmul(zg, make_double2(0.0, hf)));
            double2 e0 = cscale(zgd, u[0]), e1 = cscale(zgd, u[1]), e2 = cscale(zgd, u[2]);
            double2 v0 = csub(cscale(e1, dpoo[2]), cscale(e2, dpoo[1]));
            double2 v1 = csub(cscale(e2, dpoo[0]), cscale(e0, dpoo[2]));
            double2 v2 = csub(cscale(e0, dpoo[1]), cscale(e1, dpoo[0]));
            double2 dd = cadd(cadd(cscale(v0, srcn[0]), cscale(v1, srcn[1])), cscale(v2, srcn[2]));
            dr.e = cadd(dr.e, cscale(dd, wga));
        }
        return;
    }
    s -= nedge;
    const int ng2 = t.ngausin * t.ngausin;
    const int per_edge_b = t.nsec2 * ng2;
    int ieg = s / per_edge_b;
    int rem = s - ieg * per_edge_b;
    int isec = rem / ng2;
    int rem2 = rem - isec * ng2;
    int i = rem2 / t.ngausin;
    int j = rem2 - i * t.ngausin;
    int ig1 = (ieg + 1) % 3;
    int ig2 = ieg + 3;
    const double* CSI6 = kCsi6;
    const double* ETA6 = kEta6;
    double aresub = 1.0 / 24.0 / (double)t.nsec2;
    double ss0 = 1.0 / 3.0, ts0 = 1.0 / 3.0, ss1, ss2, ts1, ts2;
    if (isec == 0) {
        ss1 = CSI6[ieg]; ss2 = CSI6[ig2]; ts1 = ETA6[ieg]; ts2 = ETA6[ig2];
    } else {
        ss1 = CSI6[ig2]; ss2 = CSI6[ig1]; ts1 = ETA6[ig2]; ts2 = ETA6[ig1];
    }
    double wei = kGlW[t.ngausin][i] * kGlW[t.ngausin][j];
    double sga = kGlX[t.ngausin][i], tga = kGlX[t.ngausin][j];
    double sgg = 0.5 * (1.0 - sga) * ss0 + 0.25 * (1.0 + sga) * ((1.0 - tga) * ss1 + (1.0 + tga) * ss2);
    double tgg = 0.5 * (1.0 - sga) * ts0 + 0.25 * (1.0 + sga) * ((1.0 - tga) * ts1 + (1.0 + tga) * ts2);
    double pos[3];
    shape_pos(c, sgg, tgg, pos);
    double wga = wei * (1.0 + sga) * aresub * G.jac;
    double d0 = pos[0] - src[0], d1 = pos[1] - src[1], d2 = pos[2] - src[2];
    double dis = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    if (!(dis > 1e-15)) return;
    double u0 = d0 / dis, u1 = d1 / dis, u2 = d2 / dis;
    double re1 = wavruim * dis;
    double re2 = wga / (FOUR_PI * dis);
    double sn, cs;
    sincos_red(re1, &sn, &cs);
    double2 zg = make_double2(cs * re2, sn * re2);
    double2 z1 = make_double2(-1.0 / dis, wavruim);
    double2 zb = cmul(zg, z1);
    double re1_h = u0 * G.nf[0] + u1 * G.nf[1] + u2 * G.nf[2];
    double re2_h = -(u0 * srcn[0] + u1 * srcn[1] + u2 * srcn[2]);
    r.g = cadd(r.g, zg);
    r.h = cadd(r.h, cscale(zb, re1_h));
    r.ht = cadd(r.ht, cscale(zb, re2_h));
    double nd = srcn[0] * G.nf[0] + srcn[1] * G.nf[1] + srcn[2] * G.nf[2];
    r.e = cadd(r.e, cscale(cscale(zg, k2), nd));
    if (DERIV) {
        double2 dzg = cmul(zg, make_double2(0.0, hf * dis));
        double2 dzb = cadd(cmul(dzg, z1), cmul(zg, make_double2(0.0, hf)));
        dr.g = cadd(dr.g, dzg);
        dr.h = cadd(dr.h, cscale(dzb, re1_h));
        dr.ht = cadd(dr.ht, cscale(dzb, re2_h));
        dr.e = cadd(dr.e, cscale(cadd(cscale(dzg, k2), cscale(zg, 2.0 * k)), nd));
    }
}

__device__ inline void self_sample(const PanelGeom& G, const SelfTier& t, int s, const double src[3],
                                   const double srcn[3], double wavruim, double k2, Acc& r) {
    self_sample_t<false>(G, t, s, src, srcn, wavruim, k2, 0.0, 0.0, r, r);
}

// Parameters of a derivative assembly d/dp (tbem.rs:1858-2019): dk/dp = ws,
// beta and its total derivative, i rho omega and its derivative, and
// dY/dp = adm_scale * Y.
struct DerivParams {
    double k, wavruim, hf, tau_r, sign, ws, adm_scale;
    double2 beta, beta_d, iwr, iwr_d;
};

// d_cb and d_hb of one pair (tbem.rs:1972-1977); b = values, d = derivatives.
__device__ inline void deriv_blocks(const Acc& b, const Acc& d, const DerivParams& P, double2* d_cb, double2* d_hb) {
    const double2 gamma = make_double2(1.0, 0.0);
    const double2 tau = make_double2(P.tau_r, 0.0);
    *d_cb = cadd(cadd(cscale(cmul(cmul(d.h, gamma), tau), P.ws), cscale(cmul(d.e, P.beta), P.ws)),
                 cmul(b.e, P.beta_d));
    *d_hb = cadd(cadd(cscale(cmul(cmul(d.g, gamma), tau), P.ws), cscale(cmul(d.ht, P.beta), P.ws)),
                 cmul(b.ht, P.beta_d));
}

__device__ inline double2 h_block(const Acc& b, const DerivParams& P) {
    return cadd(cmul(cmul(b.g, make_double2(1.0, 0.0)), make_double2(P.tau_r, 0.0)), cmul(b.ht, P.beta));
}

// Matrix derivative of one pair for the field panel's BC (tbem.rs:1978-1995).
__device__ inline double2 deriv_coefficient(const Acc& b, const Acc& d, const PanelBc& bc, const DerivParams& P) {
    double2 d_cb, d_hb;
    deriv_blocks(b, d, P, &d_cb, &d_hb);
    if (bc.type == 0) {
        double2 v = cscale(d_cb, P.sign);
        if (bc.has_adm) {
            const double2 y = make_double2(bc.adm[0], bc.adm[1]);
            const double2 h_b = h_block(b, P);
            v = csub(v, cmul(cscale(y, P.sign), cadd(cmul(d_hb, P.iwr), cmul(h_b, P.iwr_d))));
            v = csub(v, cmul(cmul(cscale(cscale(y, P.adm_scale), P.sign), h_b), P.iwr));
        }
        return v;
    }
    if (bc.type == 1) return cscale(d_hb, -P.sign);
    return make_double2(0.0, 0.0);
}

// Right-hand-side derivative of a nonzero-BC field panel (tbem.rs:1998-2013).
__device__ inline double2 deriv_bc_contribution(const Acc& b, const Acc& d, const PanelBc& bc, const DerivParams& P) {
    double2 d_cb, d_hb;
    deriv_blocks(b, d, P, &d_cb, &d_hb);
    const double2 sv = cscale(make_double2(bc.value[0], bc.value[1]), P.sign);
    if (bc.type == 0) return cmul(sv, cadd(cmul(d_hb, P.iwr), cmul(h_block(b, P), P.iwr_d)));
    if (bc.type == 1) return cmul(sv, d_cb);
    return make_double2(0.0, 0.0);
}

// Source free terms of the derivative (tbem.rs:1896-1915).
__device__ inline void deriv_free_terms(const PanelBc& bc, const DerivParams& P, double2* diag, double2* rhs) {
    const double2 tau = make_double2(P.tau_r, 0.0);
    const double2 btd = cadd(cmul(P.iwr_d, P.beta), cmul(P.iwr, P.beta_d));
    double2 dg = make_double2(0.0, 0.0), rh = make_double2(0.0, 0.0);
    if (bc.type == 0) {
        if (bc.has_adm) {
            const double2 y = make_double2(bc.adm[0], bc.adm[1]);
            dg = csub(dg, cscale(cmul(cmul(cscale(y, P.sign), btd), tau), 0.5));
            const double2 yd = cscale(y, P.adm_scale);
            dg = csub(dg, cscale(cmul(cmul(cmul(cscale(yd, P.sign), P.iwr), P.beta), tau), 0.5));
        }
        const double2 avg = make_double2(bc.value[0], bc.value[1]);
        rh = cadd(rh, cscale(cmul(cmul(cscale(avg, P.sign), btd), tau), 0.5));
    } else if (bc.type == 1) {
        dg = cadd(dg, cscale(cmul(cscale(P.beta_d, P.sign), tau), 0.5));
    }
    *diag = dg;
    *rhs = rh;
}

#pragma clang fp contract(on)

// Per-panel constants of the TR13 points y_q = v0 + s_q e1 + t_q e2: the Gram
// entries of the two edges.  Callers compute them once per field panel.
struct FarPoints {
    double e11, e12x2, e22;  // e1.e1, 2 e1.e2, e2.e2
};

__device__ __forceinline__ FarPoints far_points(const FarPanel& P) {
    FarPoints f;
    f.e11 = P.e1[0] * P.e1[0] + P.e1[1] * P.e1[1] + P.e1[2] * P.e1[2];
    f.e12x2 = 2.0 * (P.e1[0] * P.e2[0] + P.e1[1] * P.e2[1] + P.e1[2] * P.e2[2]);
    f.e22 = P.e2[0] * P.e2[0] + P.e2[1] * P.e2[1] + P.e2[2] * P.e2[2];
    return f;
}

// 13-point sums of one (collocation point, field panel) pair.  MODE 0 needs only
// H and E; MODE 1 also G and Ht.  With rel = v0 - x:
//   |y_q - x|^2 = |rel|^2 + s_q (2 rel.e1 + s_q e1.e1 + 2 t_q e1.e2) + t_q (2 rel.e2 + t_q e2.e2)
//   (y_q - x).n_y = rel.n_y                        (e1 and e2 lie in the panel plane)
//   (y_q - x).n_x = rel.n_x + s_q e1.n_x + t_q e2.n_x
// so the per-point geometry is five FMAs for the distance and one product for
// each normal projection, after six dot products per pair.  Off the diagonal
// (which the self-panel rule overwrites) the point-to-point distance is at
// least a fraction of the panel size, so the expansion loses no accuracy.
//
// The channel sums are grouped by powers of k, and the panel factor 0.5 J/(4 pi)
// is applied once per pair.  With zg = w_q e^{ikr}/r, pn = r^.n_y,
// qn = -r^.n_x, rq = pn qn and t = 3 rq + n_x.n_y:
//   H  = sum zg (ik - 1/r) pn                       = -S1 + ik S2,
//   Ht = sum zg (ik - 1/r) qn                       = -Q1 + ik Q2,
//   E  = sum zg [(3/r^2 - k^2) rq + n_x.n_y/r^2 - ik t/r] = T1 - k^2 T2 - ik T3,
// with S1 = sum zg pn/r, S2 = sum zg pn, Q1 = sum zg qn/r, Q2 = sum zg qn,
// T1 = sum zg t/r^2, T2 = sum zg rq, T3 = sum zg t/r: per point one product
// and two FMAs per sum instead of the complex kernel factors.
template <int MODE>
__device__ __forceinline__ void far_sums(const FarPanel& P, const FarPoints& F, const Source& src, double k,
                                         double k2, double& hr, double& hi, double& er, double& ei, double& gr,
                                         double& gi, double& tr, double& ti) {
    const double rel0 = P.v0[0] - src.c[0];
    const double rel1 = P.v0[1] - src.c[1];
    const double rel2 = P.v0[2] - src.c[2];
    const double nxny = src.n[0] * P.nf[0] + src.n[1] * P.nf[1] + src.n[2] * P.nf[2];
    const double a = rel0 * rel0 + rel1 * rel1 + rel2 * rel2;
    const double b1 = 2.0 * (rel0 * P.e1[0] + rel1 * P.e1[1] + rel2 * P.e1[2]);
    const double b2 = 2.0 * (rel0 * P.e2[0] + rel1 * P.e2[1] + rel2 * P.e2[2]);
    const double pr = rel0 * P.nf[0] + rel1 * P.nf[1] + rel2 * P.nf[2];
    const double g0 = -(rel0 * src.n[0] + rel1 * src.n[1] + rel2 * src.n[2]);
    const double g1 = -(P.e1[0] * src.n[0] + P.e1[1] * src.n[1] + P.e1[2] * src.n[2]);
    const double g2 = -(P.e2[0] * src.n[0] + P.e2[1] * src.n[1] + P.e2[2] * src.n[2]);
    double s1r = 0.0, s1i = 0.0, s2r = 0.0, s2i = 0.0;
    double t1r = 0.0, t1i = 0.0, t2r = 0.0, t2i = 0.0, t3r = 0.0, t3i = 0.0;
    double q1r = 0.0, q1i = 0.0, q2r = 0.0, q2i = 0.0;
    gr = gi = 0.0;
    // TR13 constants are uniform: scalar loads from constant memory.
#pragma clang loop unroll_count(2)
    for (int q = 0; q < 13; ++q) {
        const double sq = kTrS[q], tq = kTrT[q];
        const double u1 = __builtin_fma(tq, F.e12x2, __builtin_fma(sq, F.e11, b1));
        const double u2 = __builtin_fma(tq, F.e22, b2);
        // The diagonal is overwritten by the self-panel rule; keep it finite.
        const double d2 = fmax(__builtin_fma(sq, u1, __builtin_fma(tq, u2, a)), 1e-300);
        const double inv = rsqrt_nr(d2);
        const double dis = d2 * inv;
        const double pn = pr * inv;                                                  // r^.n_y
        const double qn = __builtin_fma(sq, g1, __builtin_fma(tq, g2, g0)) * inv;    // -r^.n_x
        double sn, cs;
        sincos_red(k * dis, &sn, &cs);
        const double re2 = kTrW[q] * inv;
        const double zgr = cs * re2, zgi = sn * re2;
        const double pi = pn * inv;
        s1r = __builtin_fma(zgr, pi, s1r);
        s1i = __builtin_fma(zgi, pi, s1i);
        s2r = __builtin_fma(zgr, pn, s2r);
        s2i = __builtin_fma(zgi, pn, s2i);
        const double rq = pn * qn;
        const double t = __builtin_fma(3.0, rq, nxny);
        const double ti_ = t * inv;
        const double ti2 = ti_ * inv;
        t1r = __builtin_fma(zgr, ti2, t1r);
        t1i = __builtin_fma(zgi, ti2, t1i);
        t2r = __builtin_fma(zgr, rq, t2r);
        t2i = __builtin_fma(zgi, rq, t2i);
        t3r = __builtin_fma(zgr, ti_, t3r);
        t3i = __builtin_fma(zgi, ti_, t3i);
        if (MODE == 1) {
            gr += zgr;
            gi += zgi;
            const double qi = qn * inv;
            q1r = __builtin_fma(zgr, qi, q1r);
            q1i = __builtin_fma(zgi, qi, q1i);
            q2r = __builtin_fma(zgr, qn, q2r);
            q2i = __builtin_fma(zgi, qn, q2i);
        }
    }
    const double w = P.wj;
    hr = -w * __builtin_fma(k, s2i, s1r);
    hi = w * __builtin_fma(k, s2r, -s1i);
    er = w * __builtin_fma(k, t3i, __builtin_fma(-k2, t2r, t1r));
    ei = w * __builtin_fma(-k, t3r, __builtin_fma(-k2, t2i, t1i));
    if (MODE == 1) {
        gr *= w;
        gi *= w;
        tr = -w * __builtin_fma(k, q2i, q1r);
        ti = w * __builtin_fma(k, q2r, -q1i);
    }
}

// Rigid/vibrating coefficient (H * gamma) * tau + E * beta with gamma = 1.
__device__ __forceinline__ double2 velocity_coefficient(double hr, double hi, double er, double ei, double tau_r,
                                                        double2 beta) {
    return make_double2(hr * tau_r + (er * beta.x - ei * beta.y), hi * tau_r + (er * beta.y + ei * beta.x));
}

// ---------------------------------------------------------------------------
// Far kernel.  MODE 0: every field panel has a velocity BC without
// admittance (rigid or vibrating body), so only H and E are needed and
// A_ij = H*gamma*tau + E*beta.  MODE 1: general per-column coefficients.
//
// Row maxima for the low-precision copy (rowmax non-null, rows_per_block <=
// FAR_MAX_ROWS): the block also writes, per row i, the maximum of
// max(|Re A_ij|, |Im A_ij|) over its 256 columns at rowmax[i * nbx +
// blockIdx.x], leaving out the near pairs and the diagonal that the later
// kernels overwrite (they fold their entries into rowmax[i * nbx + nbx - 1]
// as bit patterns with an atomic max; the column-0 blocks clear that slot).
constexpr int FAR_MAX_ROWS = 16;

template <int MODE>
__global__ __launch_bounds__(256) void far_kernel(const FarPanel* __restrict__ panels,
                                                  const Source* __restrict__ sources,
                                                  const PanelBc* __restrict__ bcs, int64_t n,
                                                  int64_t lda, int rows_per_block, double k,
                                                  double tau_r, double2 beta, double2 iwr,
                                                  double2* __restrict__ A, const int64_t* __restrict__ row_ptr,
                                                  const NearPair* __restrict__ pairs, double* __restrict__ rowmax,
                                                  int64_t nbx) {
    __shared__ double smax[FAR_MAX_ROWS][256];
    __shared__ unsigned nmask[FAR_MAX_ROWS][8];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = j < n;
    const int64_t jj = active ? j : n - 1;
    const FarPanel P = panels[jj];
    const FarPoints F = far_points(P);
    PanelBc bc;
    if (MODE == 1) bc = bcs[jj];
    const double k2 = k * k;
    const int64_t i0 = (int64_t)blockIdx.y * rows_per_block;
    const int64_t i1 = min(n, i0 + rows_per_block);
    const int64_t c0 = (int64_t)blockIdx.x * 256;
    // This block's rows own one contiguous range of the row-sorted near-pair
    // list.  Its first 512 pairs are loaded into registers here, so the loads
    // complete behind the row loop; the near mask is built after it.
    int64_t e0 = 0, e1 = 0;
    int near_r[2] = {-1, -1}, near_c[2] = {0, 0};
    if (rowmax) {
        if (threadIdx.x < FAR_MAX_ROWS * 8) nmask[threadIdx.x >> 3][threadIdx.x & 7] = 0u;
        e0 = row_ptr[i0];
        e1 = row_ptr[i1];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int64_t e = e0 + threadIdx.x + 256 * u;
            if (e < e1) {
                near_r[u] = pairs[e].i - (int)i0;
                near_c[u] = (int)((int64_t)pairs[e].j - c0);
            }
        }
    }
    for (int64_t i = i0; i < i1; ++i) {
        const Source src = sources[i];
        double hr, hi, er, ei, gr, gi, tr, ti;
        far_sums<MODE>(P, F, src, k, k2, hr, hi, er, ei, gr, gi, tr, ti);
        double2 out;
        if (MODE == 0) {
            out = velocity_coefficient(hr, hi, er, ei, tau_r, beta);
        } else {
            Acc a;
            a.g = make_double2(gr, gi);
            a.h = make_double2(hr, hi);
            a.ht = make_double2(tr, ti);
            a.e = make_double2(er, ei);
            out = coefficient(a, bc, make_double2(1.0, 0.0), make_double2(tau_r, 0.0), beta, iwr);
        }
        if (active) A[i * lda + j] = out;
        if (rowmax) smax[i - i0][threadIdx.x] = active ? fmax(fabs(out.x), fabs(out.y)) : 0.0;
    }
    if (rowmax) {
        __syncthreads();  // the mask was cleared by every wave before its row loop
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (near_r[u] >= 0 && near_c[u] >= 0 && near_c[u] < 256)
                atomicOr(&nmask[near_r[u]][near_c[u] >> 5], 1u << (near_c[u] & 31));
        for (int64_t e = e0 + 512 + threadIdx.x; e < e1; e += 256) {
            const NearPair np = pairs[e];
            const int64_t c = (int64_t)np.j - c0;
            if (c >= 0 && c < 256) atomicOr(&nmask[np.i - i0][c >> 5], 1u << (c & 31));
        }
        __syncthreads();
        // per row: the maximum over the columns that are neither near pairs nor the diagonal
        const int r = threadIdx.x >> 4, seg = threadIdx.x & 15;
        double m = 0.0;
        if (i0 + r < i1) {
            const unsigned bits = nmask[r][seg >> 1] >> (16 * (seg & 1));
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int col = seg * 16 + c;
                const bool skip = ((bits >> c) & 1u) != 0u || c0 + col == i0 + r;
                m = fmax(m, skip ? 0.0 : smax[r][col]);
            }
        }
        __syncthreads();
        smax[r][seg] = m;
        __syncthreads();
        if (threadIdx.x < FAR_MAX_ROWS && i0 + threadIdx.x < i1) {
            double mm = 0.0;
#pragma unroll
            for (int c = 0; c < 16; ++c) mm = fmax(mm, smax[threadIdx.x][c]);
            const int64_t i = i0 + threadIdx.x;
            rowmax[i * nbx + blockIdx.x] = mm;
            if (blockIdx.x == 0) rowmax[i * nbx + nbx - 1] = 0.0;
        }
    }
}

// Folds max(|Re v|, |Im v|) of a near or diagonal entry into the row's slot
// (non-negative doubles order like their bit patterns; NaN entries are
// skipped, as fmax skips them).
__device__ __forceinline__ void fold_row_max(double* slot, double2 v) {
    const double m = fmax(fabs(v.x), fabs(v.y));
    if (m == m) atomicMax(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)__double_as_longlong(m));
}

// Coefficient of one near pair, summed over its leaves by one wave (lanes
// over (leaf, point)); every lane returns the total.
__device__ __forceinline__ double2 near_entry(const NearPair& np, const uint64_t* __restrict__ leaves,
                                              const PanelGeom& G, const Source& src, const PanelBc& bc,
                                              double k, double wavruim, double tau_r, double2 beta, double2 iwr,
                                              int lane) {
    const double k2 = k * k;
    Acc a = acc_zero();
    const int total = np.count * 13;
    for (int t = lane; t < total; t += 64) {
        int leaf = t / 13;
        int q = t - leaf * 13;
        near_point(G, leaves[np.begin + leaf], q, src.c, src.n, wavruim, k2, a);
    }
    wave_sum_acc(a);
    return coefficient(a, bc, make_double2(1.0, 0.0), make_double2(tau_r, 0.0), beta, iwr);
}

// Diagonal entry and BC free-term right-hand side of panel G (self integral
// summed by one wave, tbem.rs:1330-1358); every lane returns the totals.
// G refers to global memory: the sample loop indexes its nodes by edge.
__device__ __forceinline__ void self_entry(const PanelGeom& G, const Source& src, const PanelBc& bc, double k,
                                           double wavruim, double tau_r, double2 beta, double2 iwr, int lane,
                                           double2* diag_out, double2* rhs_out) {
    const double k2 = k * k;
    const SelfTier t = tier_for(k * G.size);
    const int ns = self_samples(t);
    Acc a = acc_zero();
    for (int s = lane; s < ns; s += 64) self_sample(G, t, s, src.c, src.n, wavruim, k2, a);
    wave_sum_acc(a);
    const double2 gamma = make_double2(1.0, 0.0);
    const double2 tau = make_double2(tau_r, 0.0);
    double2 diag = make_double2(0.0, 0.0);
    double2 local_rhs = make_double2(0.0, 0.0);
    double2 avg = make_double2(bc.value[0], bc.value[1]);
    // tbem.rs:1344-1358
    if (bc.type == 0) {
        diag = csub(diag, cscale(gamma, 0.5));
        if (bc.has_adm) {
            double2 adm = make_double2(bc.adm[0], bc.adm[1]);
            diag = csub(diag, cscale(cmul(cmul(cmul(adm, iwr), beta), tau), 0.5));
        }
        local_rhs = cadd(local_rhs, cscale(cmul(cmul(cmul(avg, iwr), beta), tau), 0.5));
    } else if (bc.type == 1) {
        diag = cadd(diag, cscale(cmul(beta, tau), 0.5));
        local_rhs = csub(local_rhs, cscale(cmul(avg, gamma), 0.5));
    }
    *diag_out = cadd(diag, coefficient(a, bc, gamma, tau, beta, iwr));
    *rhs_out = local_rhs;
}

__device__ __forceinline__ PanelBc rigid_bc() {
    PanelBc bc;
    bc.type = 0;
    bc.has_adm = 0;
    bc.value[0] = bc.value[1] = 0.0;
    bc.adm[0] = bc.adm[1] = 0.0;
    return bc;
}

// One wave per near pair; lanes over (leaf, point).
__global__ __launch_bounds__(256) void near_kernel(const NearPair* __restrict__ pairs, int64_t npairs,
                                                   const uint64_t* __restrict__ leaves,
                                                   const PanelGeom* __restrict__ geom,
                                                   const Source* __restrict__ sources,
                                                   const PanelBc* __restrict__ bcs, int general,
                                                   int64_t lda, double k, double wavruim,
                                                   double tau_r, double2 beta, double2 iwr,
                                                   double2* __restrict__ A, double* __restrict__ rowmax,
                                                   int64_t nbx) {
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= npairs) return;
    const NearPair np = pairs[p];
    const PanelBc bc = general ? bcs[np.j] : rigid_bc();
    const PanelGeom G = geom[np.j];  // constant indexing only: lives in registers
    const Source S = sources[np.i];
    const double2 v = near_entry(np, leaves, G, S, bc, k, wavruim, tau_r, beta, iwr, lane);
    if (lane == 0) {
        A[(int64_t)np.i * lda + np.j] = v;
        if (rowmax) fold_row_max(rowmax + (int64_t)np.i * nbx + nbx - 1, v);
    }
}

// One wave per panel: self integrals, free terms, BC free-term RHS.
__global__ __launch_bounds__(256) void self_kernel(const PanelGeom* __restrict__ geom,
                                                   const Source* __restrict__ sources,
                                                   const PanelBc* __restrict__ bcs, int general,
                                                   int64_t n, int64_t lda, double k, double wavruim,
                                                   double tau_r, double2 beta, double2 iwr,
                                                   double2* __restrict__ A, double2* __restrict__ rhs,
                                                   double* __restrict__ rowmax, int64_t nbx) {
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const PanelBc bc = general ? bcs[i] : rigid_bc();
    double2 diag, local_rhs;
    self_entry(geom[i], sources[i], bc, k, wavruim, tau_r, beta, iwr, lane, &diag, &local_rhs);
    if (lane == 0) {
        A[i * lda + i] = diag;
        rhs[i] = local_rhs;
        if (rowmax) fold_row_max(rowmax + i * nbx + nbx - 1, diag);
    }
}

// ---------------------------------------------------------------------------
// Parameter derivative of the stabilized system (tbem.rs:1858-2019 and the
// CUDA workspace's assemble_wavenumber_derivative, tbem.rs:4399-4591): the same
// routes and launch structure as the value assembly, each pair producing its
// values b and wavenumber derivatives d from one pass over its points.

// Far sums with derivatives (fast arithmetic, harmonic factor +1), grouped by
// powers of k as in far_sums.  With zg = w_q e^{ikr}/r, pn = r^.n_y,
// qn = -r^.n_x, rq = pn qn, t = 3 rq + n_x.n_y, g = qn r (so that pn r = pr,
// the pair constant (v0 - x).n_y of a flat panel), dG/dk = i r G and
// d/dk[G (ik - 1/r)] = -k r G:
//   E   = T1 - k^2 T2 - ik T3                 (T1 = sum zg t/r^2, T2 = sum zg rq, T3 = sum zg t/r)
//   dH  = -k pr G                             (G = sum zg)
//   dE  = sum [i r zg ef + zg def] = k (T2 + n_x.n_y G) - i k^2 U,  U = sum zg r rq,
// where ef = t/r^2 - k^2 rq - ik t/r and def = -2k rq - i t/r (the i t/r
// terms cancel); MODE 1 also G, Ht = -Q1 + ik Q2 (Q1 = sum zg qn/r, Q2 = sum zg qn),
// dG = i D (D = sum zg r) and dHt = -k Q3 (Q3 = sum zg g).  MODE 0 (velocity
// panels without admittance) needs E, dH, dE.  The panel factor is applied
// once per pair.
template <int MODE>
// Same geometry as far_sums (distances from the edge Gram entries, normal
// projections from per-pair dot products).
__device__ __forceinline__ void far_sums_d(const FarPanel& P, const FarPoints& F, const Source& src, double k,
                                           double k2, Acc& b, Acc& d) {
    const double rel0 = P.v0[0] - src.c[0];
    const double rel1 = P.v0[1] - src.c[1];
    const double rel2 = P.v0[2] - src.c[2];
    const double nxny = src.n[0] * P.nf[0] + src.n[1] * P.nf[1] + src.n[2] * P.nf[2];
    const double a = rel0 * rel0 + rel1 * rel1 + rel2 * rel2;
    const double b1 = 2.0 * (rel0 * P.e1[0] + rel1 * P.e1[1] + rel2 * P.e1[2]);
    const double b2 = 2.0 * (rel0 * P.e2[0] + rel1 * P.e2[1] + rel2 * P.e2[2]);
    const double pr = rel0 * P.nf[0] + rel1 * P.nf[1] + rel2 * P.nf[2];
    const double g0 = -(rel0 * src.n[0] + rel1 * src.n[1] + rel2 * src.n[2]);
    const double g1 = -(P.e1[0] * src.n[0] + P.e1[1] * src.n[1] + P.e1[2] * src.n[2]);
    const double g2 = -(P.e2[0] * src.n[0] + P.e2[1] * src.n[1] + P.e2[2] * src.n[2]);
    double gsr = 0.0, gsi = 0.0, t1r = 0.0, t1i = 0.0, t2r = 0.0, t2i = 0.0, t3r = 0.0, t3i = 0.0, ur = 0.0, ui = 0.0;
    double q1r = 0.0, q1i = 0.0, q2r = 0.0, q2i = 0.0, q3r = 0.0, q3i = 0.0, dsr = 0.0, dsi = 0.0;
#pragma unroll 2
    for (int q = 0; q < 13; ++q) {
        const double sq = kTrS[q], tq = kTrT[q];
        const double u1 = __builtin_fma(tq, F.e12x2, __builtin_fma(sq, F.e11, b1));
        const double u2 = __builtin_fma(tq, F.e22, b2);
        const double d2 = fmax(__builtin_fma(sq, u1, __builtin_fma(tq, u2, a)), 1e-300);  // diagonal: overwritten later
        const double inv = rsqrt_nr(d2);
        const double dis = d2 * inv;
        const double pn = pr * inv;
        const double gq = __builtin_fma(sq, g1, __builtin_fma(tq, g2, g0));  // qn r
        const double qn = gq * inv;
        double sn, cs;
        sincos_red(k * dis, &sn, &cs);
        const double re2 = kTrW[q] * inv;
        const double zgr = cs * re2, zgi = sn * re2;
        gsr += zgr;
        gsi += zgi;
        const double rq = pn * qn;
        const double t = __builtin_fma(3.0, rq, nxny);
        const double ti1 = t * inv;
        const double ti2 = ti1 * inv;
        t1r = __builtin_fma(zgr, ti2, t1r);
        t1i = __builtin_fma(zgi, ti2, t1i);
        t2r = __builtin_fma(zgr, rq, t2r);
        t2i = __builtin_fma(zgi, rq, t2i);
        t3r = __builtin_fma(zgr, ti1, t3r);
        t3i = __builtin_fma(zgi, ti1, t3i);
        const double drq = dis * rq;
        ur = __builtin_fma(zgr, drq, ur);
        ui = __builtin_fma(zgi, drq, ui);
        if (MODE == 1) {
            const double qi = qn * inv;
            q1r = __builtin_fma(zgr, qi, q1r);
            q1i = __builtin_fma(zgi, qi, q1i);
            q2r = __builtin_fma(zgr, qn, q2r);
            q2i = __builtin_fma(zgi, qn, q2i);
            q3r = __builtin_fma(zgr, gq, q3r);
            q3i = __builtin_fma(zgi, gq, q3i);
            dsr = __builtin_fma(zgr, dis, dsr);
            dsi = __builtin_fma(zgi, dis, dsi);
        }
    }
    const double w = P.wj;
    b = acc_zero();
    d = acc_zero();
    b.e = make_double2(w * __builtin_fma(k, t3i, __builtin_fma(-k2, t2r, t1r)),
                       w * __builtin_fma(-k, t3r, __builtin_fma(-k2, t2i, t1i)));
    const double kpw = -k * pr * w;
    d.h = make_double2(kpw * gsr, kpw * gsi);
    d.e = make_double2(w * __builtin_fma(k, __builtin_fma(nxny, gsr, t2r), k2 * ui),
                       w * __builtin_fma(k, __builtin_fma(nxny, gsi, t2i), -k2 * ur));
    if (MODE == 1) {
        b.g = make_double2(w * gsr, w * gsi);
        b.ht = make_double2(-w * __builtin_fma(k, q2i, q1r), w * __builtin_fma(k, q2r, -q1i));
        d.g = make_double2(-w * dsi, w * dsr);
        d.ht = make_double2(-k * w * q3r, -k * w * q3i);
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void far_deriv_kernel(const FarPanel* __restrict__ panels,
                                                        const Source* __restrict__ sources,
                                                        const PanelBc* __restrict__ bcs, int64_t n, int64_t lda,
                                                        int rows_per_block, DerivParams P,
                                                        double2* __restrict__ A) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = j < n;
    const int64_t jj = active ? j : n - 1;
    const FarPanel F = panels[jj];
    const FarPoints FP = far_points(F);
    const PanelBc bc = MODE == 1 ? bcs[jj] : rigid_bc();
    const double k2 = P.k * P.k;
    const int64_t i0 = (int64_t)blockIdx.y * rows_per_block;
    const int64_t i1 = min(n, i0 + rows_per_block);
    for (int64_t i = i0; i < i1; ++i) {
        const Source src = sources[i];
        Acc b, d;
        far_sums_d<MODE>(F, FP, src, P.k, k2, b, d);
        const double2 out = deriv_coefficient(b, d, bc, P);
        if (active) A[i * lda + j] = out;
    }
}

// One wave per near pair; lanes over (leaf, point); reference operation order.
__global__ __launch_bounds__(256) void near_deriv_kernel(const NearPair* __restrict__ pairs, int64_t npairs,
                                                         const uint64_t* __restrict__ leaves,
                                                         const PanelGeom* __restrict__ geom,
                                                         const Source* __restrict__ sources,
                                                         const PanelBc* __restrict__ bcs, int general, int64_t lda,
                                                         DerivParams P, double2* __restrict__ A) {
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= npairs) return;
    const NearPair np = pairs[p];
    const PanelBc bc = general ? bcs[np.j] : rigid_bc();
    const PanelGeom G = geom[np.j];
    const Source S = sources[np.i];
    const double k2 = P.k * P.k;
    Acc b = acc_zero(), d = acc_zero();
    const int total = np.count * 13;
    for (int t = lane; t < total; t += 64) {
        const int leaf = t / 13;
        near_point_t<true>(G, leaves[np.begin + leaf], t - leaf * 13, S.c, S.n, P.wavruim, k2, P.k, P.hf, b, d);
    }
    wave_sum_acc(b);
    wave_sum_acc(d);
    if (lane == 0) A[(int64_t)np.i * lda + np.j] = deriv_coefficient(b, d, bc, P);
}

// One wave per panel: self-integral derivatives, free terms, BC free-term RHS.
__global__ __launch_bounds__(256) void self_deriv_kernel(const PanelGeom* __restrict__ geom,
                                                         const Source* __restrict__ sources,
                                                         const PanelBc* __restrict__ bcs, int general, int64_t n,
                                                         int64_t lda, DerivParams P, double2* __restrict__ A,
                                                         double2* __restrict__ rhs) {
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const PanelBc bc = general ? bcs[i] : rigid_bc();
    const PanelGeom& G = geom[i];
    const Source S = sources[i];
    const double k2 = P.k * P.k;
    const SelfTier t = tier_for(P.k * G.size);
    const int ns = self_samples(t);
    Acc b = acc_zero(), d = acc_zero();
    for (int s = lane; s < ns; s += 64) self_sample_t<true>(G, t, s, S.c, S.n, P.wavruim, k2, P.k, P.hf, b, d);
    wave_sum_acc(b);
    wave_sum_acc(d);
    if (lane == 0) {
        double2 diag, local_rhs;
        deriv_free_terms(bc, P, &diag, &local_rhs);
        A[i * lda + i] = cadd(diag, deriv_coefficient(b, d, bc, P));
        rhs[i] = local_rhs;
    }
}

// Block per row: right-hand-side derivative over the nonzero-BC columns in a
// fixed order, as bc_rhs_kernel.
__global__ __launch_bounds__(256) void bc_rhs_deriv_kernel(const int32_t* __restrict__ cols, int64_t ncols,
                                                           const NearPair* __restrict__ pairs,
                                                           const int64_t* __restrict__ row_ptr,
                                                           const uint64_t* __restrict__ leaves,
                                                           const PanelGeom* __restrict__ geom,
                                                           const Source* __restrict__ sources,
                                                           const PanelBc* __restrict__ bcs, DerivParams P,
                                                           double2* __restrict__ rhs) {
    __shared__ double part[4][2];
    const int64_t i = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const Source src = sources[i];
    const double k2 = P.k * P.k;
    double2 sum = make_double2(0.0, 0.0);
    const int64_t pb = row_ptr[i], pe = row_ptr[i + 1];
    for (int64_t c = wave; c < ncols; c += 4) {
        const int32_t j = cols[c];
        const PanelGeom G = geom[j];
        Acc b = acc_zero(), d = acc_zero();
        if (j == i) {
            const SelfTier t = tier_for(P.k * G.size);
            const int ns = self_samples(t);
            for (int s = lane; s < ns; s += 64)
                self_sample_t<true>(G, t, s, src.c, src.n, P.wavruim, k2, P.k, P.hf, b, d);
        } else {
            int64_t lo = pb, hi = pe;
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                if (pairs[mid].j < j) lo = mid + 1; else hi = mid;
            }
            if (lo < pe && pairs[lo].j == j) {
                const NearPair np = pairs[lo];
                const int total = np.count * 13;
                for (int t = lane; t < total; t += 64) {
                    const int leaf = t / 13;
                    near_point_t<true>(G, leaves[np.begin + leaf], t - leaf * 13, src.c, src.n, P.wavruim, k2, P.k,
                                       P.hf, b, d);
                }
            } else if (lane < 13) {
                full_point_t<true>(G, lane, src.c, src.n, P.wavruim, k2, P.k, P.hf, b, d);
            }
        }
        wave_sum_acc(b);
        wave_sum_acc(d);
        sum = cadd(sum, deriv_bc_contribution(b, d, bcs[j], P));
    }
    if (lane == 0) {
        part[wave][0] = sum.x;
        part[wave][1] = sum.y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 t = rhs[i];
        for (int w = 0; w < 4; ++w) t = cadd(t, make_double2(part[w][0], part[w][1]));
        rhs[i] = t;
    }
}

// Block per row: sum over nonzero-BC columns of tbem.rs:1419-1435, in a
// fixed order (wave w takes columns w, w+4, ...; waves summed 0..3).
__global__ __launch_bounds__(256) void bc_rhs_kernel(const int32_t* __restrict__ cols, int64_t ncols,
                                                     const NearPair* __restrict__ pairs,
                                                     const int64_t* __restrict__ row_ptr,
                                                     const uint64_t* __restrict__ leaves,
                                                     const PanelGeom* __restrict__ geom,
                                                     const Source* __restrict__ sources,
                                                     const PanelBc* __restrict__ bcs, double k,
                                                     double wavruim, double tau_r, double2 beta,
                                                     double2 iwr, double2* __restrict__ rhs) {
    __shared__ double part[4][2];
    const int64_t i = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const Source src = sources[i];
    const double k2 = k * k;
    const double2 gamma = make_double2(1.0, 0.0);
    const double2 tau = make_double2(tau_r, 0.0);
    double2 sum = make_double2(0.0, 0.0);
    const int64_t pb = row_ptr[i], pe = row_ptr[i + 1];
    for (int64_t c = wave; c < ncols; c += 4) {
        const int32_t j = cols[c];
        const PanelGeom G = geom[j];
        Acc a = acc_zero();
        if (j == i) {
            const SelfTier t = tier_for(k * G.size);
            const int ns = self_samples(t);
            for (int s = lane; s < ns; s += 64) self_sample(G, t, s, src.c, src.n, wavruim, k2, a);
        } else {
            int64_t lo = pb, hi = pe;  // pairs of row i are sorted by column
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                if (pairs[mid].j < j) lo = mid + 1; else hi = mid;
            }
            if (lo < pe && pairs[lo].j == j) {
                const NearPair np = pairs[lo];
                const int total = np.count * 13;
                for (int t = lane; t < total; t += 64) {
                    int leaf = t / 13;
                    near_point(G, leaves[np.begin + leaf], t - leaf * 13, src.c, src.n, wavruim, k2, a);
                }
            } else if (lane < 13) {
                full_point(G, lane, src.c, src.n, wavruim, k2, a);
            }
        }
        wave_sum_acc(a);
        sum = cadd(sum, bc_contribution(a, bcs[j], gamma, tau, beta, iwr));
    }
    if (lane == 0) {
        part[wave][0] = sum.x;
        part[wave][1] = sum.y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 t = rhs[i];
        for (int w = 0; w < 4; ++w) t = cadd(t, make_double2(part[w][0], part[w][1]));
        rhs[i] = t;
    }
}

// incident.rs:216-330, 807-834; bem_solver.rs:5005-5034 (reference order).
#pragma clang fp contract(off)
__device__ __forceinline__ double2 plane_wave_value(const Source& s, const double* d, double2 amp, double k,
                                                    double hf, double tau_r, double2 beta, double sign) {
    double kdotx = hf * k * (d[0] * s.c[0] + d[1] * s.c[1] + d[2] * s.c[2]);
    double kdotn = hf * k * (d[0] * s.n[0] + d[1] * s.n[1] + d[2] * s.n[2]);
    double sn, cs;
    sincos_red(kdotx, &sn, &cs);
    double2 p = cmul(amp, make_double2(cs, sn));
    double2 dpdn = cmul(make_double2(0.0, kdotn), p);
    double2 gp = cmul(make_double2(1.0, 0.0), p);
    double2 bt = cmul(beta, make_double2(tau_r, 0.0));
    double2 inc = cadd(gp, cmul(bt, dpdn));
    inc = make_double2(-inc.x, -inc.y);
    return cmul(inc, make_double2(sign, 0.0));
}

__global__ void plane_wave_kernel(const Source* __restrict__ sources, int64_t n,
                                  const double* __restrict__ dirs, const double2* __restrict__ amps,
                                  int64_t nrhs, double k, double hf, double tau_r, double2 beta,
                                  double sign, int accumulate, double2* __restrict__ rhs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nrhs) return;
    const int64_t i = t / nrhs;
    const int64_t c = t - i * nrhs;
    const double2 add = plane_wave_value(sources[i], dirs + 3 * c, amps[c], k, hf, tau_r, beta, sign);
    rhs[t] = accumulate ? cadd(rhs[t], add) : add;
}

#pragma clang fp contract(on)

// Far entries of a frequency family (instances on one mesh with wavenumbers
// k_f = k0 + f dk): one lane per column, FAMILY_ROWS rows per workgroup,
// blockIdx.z the family.  Per TR13 point the k-independent geometry
// (distance, 1/r, normal projections, weight) is formed once, the phase
// exp(i k0 r) and the step exp(i dk r) come from two sincos evaluations, and
// each member then needs only the phase advance (one complex product) and
// its channel sums.  The members' own k_f enter the channel formulas; the
// phase recurrence differs from a direct sincos(k_f r) by at most about
// F ulp of the phase (F <= 16).  MODE 0 (velocity panels without
// admittance): A_ij = H tau + E beta, the far_sums/velocity_coefficient
// formulas of the single-frequency far kernel.
constexpr int FAMILY_ROWS = 16;

template <int F>
__global__ __launch_bounds__(256) void batch_far_family_kernel(const FreqFamily* __restrict__ families,
                                                               const BatchInstance* __restrict__ inst,
                                                               const FarPanel* __restrict__ far,
                                                               const Source* __restrict__ sources,
                                                               double2* __restrict__ A) {
    const FreqFamily& fam = families[blockIdx.z];
    const int64_t n = fam.n;
    const int64_t i0 = (int64_t)blockIdx.y * FAMILY_ROWS;
    if ((int64_t)blockIdx.x * 256 >= n || i0 >= n) return;
    const int count = fam.count;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = j < n;
    const FarPanel P = far[fam.panel_off + (active ? j : n - 1)];
    const FarPoints FP = far_points(P);
    const Source* src = sources + fam.panel_off;
    const double k0 = fam.k0, dk = fam.dk;
    const int64_t i1 = min(n, i0 + (int64_t)FAMILY_ROWS);
    for (int64_t i = i0; i < i1; ++i) {
        const Source S = src[i];
        const double rel0 = P.v0[0] - S.c[0];
        const double rel1 = P.v0[1] - S.c[1];
        const double rel2 = P.v0[2] - S.c[2];
        const double nxny = S.n[0] * P.nf[0] + S.n[1] * P.nf[1] + S.n[2] * P.nf[2];
        const double a = rel0 * rel0 + rel1 * rel1 + rel2 * rel2;
        const double b1 = 2.0 * (rel0 * P.e1[0] + rel1 * P.e1[1] + rel2 * P.e1[2]);
        const double b2 = 2.0 * (rel0 * P.e2[0] + rel1 * P.e2[1] + rel2 * P.e2[2]);
        const double pr = rel0 * P.nf[0] + rel1 * P.nf[1] + rel2 * P.nf[2];
        const double g0 = -(rel0 * S.n[0] + rel1 * S.n[1] + rel2 * S.n[2]);
        const double g1 = -(P.e1[0] * S.n[0] + P.e1[1] * S.n[1] + P.e1[2] * S.n[2]);
        const double g2 = -(P.e2[0] * S.n[0] + P.e2[1] * S.n[1] + P.e2[2] * S.n[2]);
        double hr[F], hi[F], er[F], ei[F];
#pragma unroll
        for (int f = 0; f < F; ++f) hr[f] = hi[f] = er[f] = ei[f] = 0.0;
#pragma clang loop unroll(disable)
        for (int q = 0; q < 13; ++q) {
            const double sq = kTrS[q], tq = kTrT[q];
            const double u1 = __builtin_fma(tq, FP.e12x2, __builtin_fma(sq, FP.e11, b1));
            const double u2 = __builtin_fma(tq, FP.e22, b2);
            const double d2 = fmax(__builtin_fma(sq, u1, __builtin_fma(tq, u2, a)), 1e-300);
            const double inv = rsqrt_nr(d2);
            const double dis = d2 * inv;
            const double pn = pr * inv;                                                // r^.n_y
            const double qn = __builtin_fma(sq, g1, __builtin_fma(tq, g2, g0)) * inv;  // -r^.n_x
            const double re2 = (kTrW[q] * P.wj) * inv;
            const double inv2 = inv * inv;
            const double rq = pn * qn;
            // ef = (3/r^2 - k^2) rq + n.n/r^2  -  i k (3 rq + n.n)/r
            const double ef0 = (3.0 * inv2) * rq + nxny * inv2;
            const double ef1 = -inv * (3.0 * rq + nxny);
            // member f, phase p = exp(i k_f r): H += p re2 pn (-1/r + i k_f),
            // E += p re2 (ef0 - k_f^2 rq + i k_f ef1); the k_f-free factors once here
            const double wb = re2 * pn, wa = -wb * inv;
            const double E0 = re2 * ef0, R = re2 * rq, E1 = re2 * ef1;
            double phr, phi, str, sti;
            sincos_red(k0 * dis, &phi, &phr);
            sincos_red(dk * dis, &sti, &str);
            // all F members, also past count (kf = 0 there; those sums are dropped)
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const double kf = fam.kf[f];
                const double u = __builtin_fma(-kf * kf, R, E0);
                const double v = kf * E1;
                const double t = kf * wb;
                hr[f] = __builtin_fma(-phi, t, __builtin_fma(phr, wa, hr[f]));
                hi[f] = __builtin_fma(phr, t, __builtin_fma(phi, wa, hi[f]));
                er[f] = __builtin_fma(-phi, v, __builtin_fma(phr, u, er[f]));
                ei[f] = __builtin_fma(phi, u, __builtin_fma(phr, v, ei[f]));
                const double nr = __builtin_fma(phr, str, -phi * sti);
                phi = __builtin_fma(phr, sti, phi * str);
                phr = nr;
            }
        }
        if (active) {
#pragma unroll
            for (int f = 0; f < F; ++f) {
                if (f < count) {
                    const BatchInstance& I = inst[fam.member[f]];
                    A[I.a_off + i * n + j] =
                        velocity_coefficient(hr[f], hi[f], er[f], ei[f], I.tau, make_double2(I.beta[0], I.beta[1]));
                }
            }
        }
    }
}

// Near pairs of a frequency family: the subdivided-leaf quadrature of each
// pair (regular.rs:164-281, the rule of near_entry) with the point geometry
// formed once for all members.  Sixteen lanes take one pair (a wave four),
// lanes over (leaf, point); per point the phase exp(i k0 r) and the step
// exp(i dk r) come from two sincos evaluations and each member advances the
// phase by one complex product.  Rigid panels need the H and E channels
// only; the 16-lane sums use row DPP, and lane f writes member f's entry.
constexpr int NEAR_FAMILY_LANES = 16;

template <int CTRL>
__device__ __forceinline__ double dpp_read(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

// Sum over the 16 lanes of a DPP row; every lane of the row gets a total.
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_read<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_read<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_read<0x141>(v);  // row_half_mirror: lane i <- 7 - i within each half row
    v += dpp_read<0x140>(v);  // row_mirror: lane i <- 15 - i
    return v;
}

template <int F>
__global__ __launch_bounds__(256) void batch_near_family_kernel(const FreqFamily* __restrict__ families,
                                                                const BatchInstance* __restrict__ inst,
                                                                const NearPair* __restrict__ pairs,
                                                                const uint64_t* __restrict__ leaves,
                                                                const PanelGeom* __restrict__ geom,
                                                                const Source* __restrict__ sources,
                                                                double2* __restrict__ A) {
#pragma clang fp contract(on)
    const FreqFamily& fam = families[blockIdx.y];
    const int sub = threadIdx.x & (NEAR_FAMILY_LANES - 1);
    const int count = fam.count;
    const double k0 = fam.k0, dk = fam.dk;
    const int64_t n = fam.n;
    const int64_t step = (int64_t)gridDim.x * (256 / NEAR_FAMILY_LANES);
    for (int64_t p = (int64_t)blockIdx.x * (256 / NEAR_FAMILY_LANES) + (threadIdx.x / NEAR_FAMILY_LANES);
         p < fam.npairs; p += step) {
        const NearPair np = pairs[fam.pair_begin + p];
        const PanelGeom G = geom[fam.panel_off + np.j];
        const Source S = sources[fam.panel_off + np.i];
        const double nxny = S.n[0] * G.nf[0] + S.n[1] * G.nf[1] + S.n[2] * G.nf[2];
        double hr[F], hi[F], er[F], ei[F];
#pragma unroll
        for (int f = 0; f < F; ++f) hr[f] = hi[f] = er[f] = ei[f] = 0.0;
        const int total = np.count * 13;
        for (int t = sub; t < total; t += NEAR_FAMILY_LANES) {
            const int leaf = t / 13;
            const int q = t - leaf * 13;
            double xi[3], et[3];
            decode_leaf(leaves[np.begin + leaf], xi, et);
            const double cs = kTrS[q], ec = kTrT[q];
            const double l0 = 1.0 - cs - ec;
            const double xio = xi[0] * l0 + xi[1] * cs + xi[2] * ec;
            const double eto = et[0] * l0 + et[1] * cs + et[2] * ec;
            const double det = fabs((xi[1] - xi[0]) * (et[2] - et[0]) - (xi[2] - xi[0]) * (et[1] - et[0]));
            const double wga = ((kTrW[q] * 0.5) * det) * G.jac;
            double pos[3];
            shape_pos(G.c, xio, eto, pos);
            const double d0 = pos[0] - S.c[0], d1 = pos[1] - S.c[1], d2 = pos[2] - S.c[2];
            const double dsq = d0 * d0 + d1 * d1 + d2 * d2;
            if (!(dsq > 1e-30)) continue;  // the point evaluation skips |d| <= 1e-15
            const double inv = rsqrt_nr(dsq);
            const double dis = dsq * inv;
            const double u0 = d0 * inv, u1 = d1 * inv, u2 = d2 * inv;
            const double re2 = (wga * (1.0 / FOUR_PI)) * inv;
            const double pn = u0 * G.nf[0] + u1 * G.nf[1] + u2 * G.nf[2];   // r^.n_y
            const double qn = -(u0 * S.n[0] + u1 * S.n[1] + u2 * S.n[2]);  // -r^.n_x
            const double rq = pn * qn;
            const double inv2 = inv * inv;
            const double ef0 = (3.0 * inv2) * rq + nxny * inv2;
            const double ef1 = -inv * (3.0 * rq + nxny);
            // the channel factors of batch_far_family_kernel
            const double wb = re2 * pn, wa = -wb * inv;
            const double E0 = re2 * ef0, R = re2 * rq, E1 = re2 * ef1;
            double phr, phi, str, sti;
            sincos_red(k0 * dis, &phi, &phr);
            sincos_red(dk * dis, &sti, &str);
            // all F members, also past count (kf = 0 there; those sums are dropped)
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const double kf = fam.kf[f];
                const double u = E0 - (kf * kf) * R;
                const double v = kf * E1;
                const double t = kf * wb;
                hr[f] += phr * wa - phi * t;
                hi[f] += phi * wa + phr * t;
                er[f] += phr * u - phi * v;
                ei[f] += phr * v + phi * u;
                const double nr = phr * str - phi * sti;
                phi = phr * sti + phi * str;
                phr = nr;
            }
        }
#pragma unroll
        for (int f = 0; f < F; ++f) {
            if (f < count) {
                const double h_r = row16_sum(hr[f]), h_i = row16_sum(hi[f]);
                const double e_r = row16_sum(er[f]), e_i = row16_sum(ei[f]);
                if (sub == f) {
                    const BatchInstance& I = inst[fam.member[f]];
                    const double tau = I.tau;
                    A[I.a_off + (int64_t)np.i * n + np.j] =
                        make_double2(h_r * tau + (e_r * I.beta[0] - e_i * I.beta[1]),
                                     h_i * tau + (e_r * I.beta[1] + e_i * I.beta[0]));
                }
            }
        }
    }
}

// Batch of rigid instances, two persistent launches over the same items
// (instance, first row); the host orders items by decreasing instance size
// and each workgroup takes items from a counter until none are left.
// PHASE 0 writes the far entries of the item's rows (lane per column) and
// their plane-wave right-hand sides; PHASE 1, launched after it, overwrites
// the near pairs and the diagonal of those rows, which is the overwrite order
// of the single-system launch.  The phases are separate kernels because the
// near/self code needs about twice the registers of the far loop.
template <int PHASE>
__global__ __launch_bounds__(256) void batch_assemble_kernel(BatchAssemblyArgs args) {
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(args.counter + PHASE, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= args.nitems) break;
        const int2 item = args.items[it];
        const BatchInstance& I = args.inst[item.x];
        const int64_t n = I.n;
        const int64_t r0 = item.y;
        const int64_t r1 = min(n, r0 + (int64_t)args.rows_per_item);
        const Source* src = args.src + I.panel_off;
        double2* A = args.A + I.a_off;
        const double k = I.k;
        const double2 beta = make_double2(I.beta[0], I.beta[1]);
        if (PHASE == 0) {
            const FarPanel* far = args.far + I.panel_off;
            const double k2 = k * k;
            // members of a frequency family got their far entries from batch_far_family_kernel
            for (int64_t jc = 0; jc < (I.in_family ? 0 : n); jc += 256) {
                const int64_t j = jc + threadIdx.x;
                const bool active = j < n;
                const FarPanel P = far[active ? j : n - 1];
                const FarPoints F = far_points(P);
                for (int64_t i = r0; i < r1; ++i) {
                    double hr, hi, er, ei, gr, gi, tr, ti;
                    far_sums<0>(P, F, src[i], k, k2, hr, hi, er, ei, gr, gi, tr, ti);
                    if (active) A[i * n + j] = velocity_coefficient(hr, hi, er, ei, I.tau, beta);
                }
            }
            const int64_t nr = (r1 - r0) * I.ncol;
            for (int64_t t = threadIdx.x; t < nr; t += 256) {
                const int64_t i = r0 + t / I.ncol;
                const int c = (int)(t % I.ncol);
                args.rhs[I.vec_off + i * I.ncol + c] = plane_wave_value(
                    src[i], args.dirs + 3 * (I.dir_off + c), make_double2(1.0, 0.0), k, 1.0, I.tau, beta, I.sign);
            }
        } else {
            const PanelGeom* geom = args.geom + I.panel_off;
            const PanelBc bc = rigid_bc();
            const double2 iwr = make_double2(0.0, 0.0);  // no admittance on rigid panels
            // members of a frequency family got their near entries from batch_near_family_kernel
            const int64_t pb = args.row_ptr[I.panel_off + r0];
            const int64_t pe = I.in_family ? pb : args.row_ptr[I.panel_off + r1];
            for (int64_t p = pb + wave; p < pe; p += 4) {
                const NearPair np = args.pairs[p];
                const PanelGeom G = geom[np.j];
                const Source S = src[np.i];
                const double2 v = near_entry(np, args.leaves, G, S, bc, k, k, I.tau, beta, iwr, lane);
                if (lane == 0) A[(int64_t)np.i * n + np.j] = v;
            }
            for (int64_t i = r0 + wave; i < r1; i += 4) {
                double2 diag, unused;
                self_entry(geom[i], src[i], bc, k, k, I.tau, beta, iwr, lane, &diag, &unused);
                if (lane == 0) A[i * n + i] = diag;
            }
        }
    }
}

__global__ void reset_counters_kernel(unsigned* counter) { counter[threadIdx.x] = 0u; }

// ---------------------------------------------------------------------------
// Matrix-free exact product y = sign * A x (apply_tbem_stabilized_exact,
// crates/sonium-bem/src/core/assembly/hmatrix.rs:2363-2425) with the entries of
// the assembly kernels, evaluated on the fly in O(N) memory:
//   exact_far_kernel   far[i] = sum over j != i of the 13-point entry times x_j
//                      (lane per column, EXACT_ROWS rows per workgroup, row sums
//                      reduced across the workgroup in a fixed order);
//   exact_near_kernel  per near pair: (subdivided entry - 13-point entry) x_j;
//   exact_self_kernel  per panel: diagonal entry times x_i;
//   exact_combine_kernel  y_i = sign (far_i + near pairs of row i in CSR
//                      order + self_i).
constexpr int EXACT_ROWS = 4;

template <int MODE>
__device__ __forceinline__ double2 far_entry(const FarPanel& P, const FarPoints& F, const Source& S,
                                             const PanelBc& bc, double k, double k2, double tau_r, double2 beta,
                                             double2 iwr) {
    double hr, hi, er, ei, gr, gi, tr, ti;
    far_sums<MODE>(P, F, S, k, k2, hr, hi, er, ei, gr, gi, tr, ti);
    if (MODE == 0) return velocity_coefficient(hr, hi, er, ei, tau_r, beta);
    Acc a;
    a.g = make_double2(gr, gi);
    a.h = make_double2(hr, hi);
    a.ht = make_double2(tr, ti);
    a.e = make_double2(er, ei);
    return coefficient(a, bc, make_double2(1.0, 0.0), make_double2(tau_r, 0.0), beta, iwr);
}

template <int MODE>
__global__ __launch_bounds__(256) void exact_far_kernel(const FarPanel* __restrict__ panels,
                                                        const Source* __restrict__ sources,
                                                        const PanelBc* __restrict__ bcs, int64_t n, double k,
                                                        double tau_r, double2 beta, double2 iwr,
                                                        const double2* __restrict__ x, double2* __restrict__ out) {
    __shared__ double2 red[4][EXACT_ROWS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * EXACT_ROWS;
    const double k2 = k * k;
    // rows one after another, each over every column chunk in order (the
    // panels are L2-resident), so one row's point loop is live at a time
#pragma clang loop unroll(disable)
    for (int r = 0; r < EXACT_ROWS; ++r) {
        const int64_t i = min(i0 + r, n - 1);  // rows past n are computed and dropped
        const Source S = sources[i];
        double2 acc = make_double2(0.0, 0.0);
        for (int64_t jc = 0; jc < n; jc += 256) {
            const int64_t j = jc + threadIdx.x;
            const bool active = j < n;
            const int64_t jj = active ? j : n - 1;
            const FarPanel P = panels[jj];
            const FarPoints F = far_points(P);
            PanelBc bc;
            if (MODE == 1) bc = bcs[jj];
            const double2 xj = x[jj];
            const double2 v = far_entry<MODE>(P, F, S, bc, k, k2, tau_r, beta, iwr);
            if (active && j != i) {
                acc.x += v.x * xj.x - v.y * xj.y;
                acc.y += v.x * xj.y + v.y * xj.x;
            }
        }
        acc.x = wave_sum(acc.x);
        acc.y = wave_sum(acc.y);
        if (lane == 0) red[wave][r] = acc;
    }
    __syncthreads();
    if (threadIdx.x < EXACT_ROWS && i0 + threadIdx.x < n) {
        double2 t = red[0][threadIdx.x];
        for (int w = 1; w < 4; ++w) t = cadd(t, red[w][threadIdx.x]);
        out[i0 + threadIdx.x] = t;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void exact_near_kernel(const NearPair* __restrict__ pairs, int64_t npairs,
                                                         const uint64_t* __restrict__ leaves,
                                                         const PanelGeom* __restrict__ geom,
                                                         const FarPanel* __restrict__ panels,
                                                         const Source* __restrict__ sources,
                                                         const PanelBc* __restrict__ bcs, double k, double wavruim,
                                                         double tau_r, double2 beta, double2 iwr,
                                                         const double2* __restrict__ x, double2* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= npairs) return;
    const NearPair np = pairs[p];
    const PanelBc bc = MODE == 1 ? bcs[np.j] : rigid_bc();
    const PanelGeom G = geom[np.j];
    const Source S = sources[np.i];
    const double2 v = near_entry(np, leaves, G, S, bc, k, wavruim, tau_r, beta, iwr, lane);
    if (lane == 0) {
        const FarPanel P = panels[np.j];
        const double2 f = far_entry<MODE>(P, far_points(P), S, bc, k, k * k, tau_r, beta, iwr);
        out[p] = cmul(csub(v, f), x[np.j]);
    }
}

__global__ __launch_bounds__(256) void exact_self_kernel(const PanelGeom* __restrict__ geom,
                                                         const Source* __restrict__ sources,
                                                         const PanelBc* __restrict__ bcs, int general, int64_t n,
                                                         double k, double wavruim, double tau_r, double2 beta,
                                                         double2 iwr, const double2* __restrict__ x,
                                                         double2* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const PanelBc bc = general ? bcs[i] : rigid_bc();
    double2 diag, unused;
    self_entry(geom[i], sources[i], bc, k, wavruim, tau_r, beta, iwr, lane, &diag, &unused);
    if (lane == 0) out[i] = cmul(diag, x[i]);
}

__global__ void exact_combine_kernel(const double2* __restrict__ far, const double2* __restrict__ near,
                                     const int64_t* __restrict__ row_ptr, const double2* __restrict__ self,
                                     int64_t n, double sign, double2* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 s = far[i];
    for (int64_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) s = cadd(s, near[p]);
    s = cadd(s, self[i]);
    y[i] = make_double2(sign * s.x, sign * s.y);
}

// ---------------------------------------------------------------------------
// Pressure at evaluation points (compute_scattered_field / compute_total_field,
// crates/sonium-bem/src/core/postprocess/pressure.rs:1388-1480, 1566-1620,
// 1987-2044), constant basis on Tri3 panels:
//   p(x) = sum_j sum_q [p_j G (-1/r + ik) (r.n_j)/r - v_j i rho omega G] J_j w_q
// with r = y_q - x, G = e^{ikr}/(4 pi r) and the 7-point rule.  A workgroup
// takes FIELD_ROWS points and one slice of the panels (lane per panel); each
// quadrature point's position and weight are formed once per panel and used
// for all the workgroup's points, whose coordinates are wave-uniform.  The
// slices' partial sums are added per point in slice order.
constexpr int FIELD_ROWS = 8;

template <int HAS_V>
__global__ __launch_bounds__(256) void field_kernel(const FarPanel* __restrict__ panels, int64_t n,
                                                    const double* __restrict__ points, int64_t m,
                                                    const double2* __restrict__ p, const double2* __restrict__ v,
                                                    double k, double pf, int64_t panels_per_slice,
                                                    double2* __restrict__ part) {
    __shared__ double2 red[4][FIELD_ROWS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * FIELD_ROWS;
    const int64_t jbeg = (int64_t)blockIdx.y * panels_per_slice;
    const int64_t jend = min(n, jbeg + panels_per_slice);
    double xs[FIELD_ROWS][3];
#pragma unroll
    for (int r = 0; r < FIELD_ROWS; ++r) {
        const double* x = points + 3 * min(i0 + r, m - 1);  // rows past m are computed and dropped
        xs[r][0] = x[0];
        xs[r][1] = x[1];
        xs[r][2] = x[2];
    }
    double2 acc[FIELD_ROWS];
#pragma unroll
    for (int r = 0; r < FIELD_ROWS; ++r) acc[r] = make_double2(0.0, 0.0);
    for (int64_t jc = jbeg; jc < jend; jc += 256) {
        const int64_t j = jc + threadIdx.x;
        const bool active = j < jend;
        const FarPanel P = panels[active ? j : n - 1];
        const double2 pj = p[active ? j : n - 1];
        const double2 vj = HAS_V ? v[active ? j : n - 1] : make_double2(0.0, 0.0);
        const double scale = (active && P.wj > 0.0) ? P.wj : 0.0;  // w J / (4 pi), times 0.5 for the rule
        // (y_q - x).n_y = (v0 - x).n_y for every point y_q of the flat panel
        double pr[FIELD_ROWS];
#pragma unroll
        for (int r = 0; r < FIELD_ROWS; ++r)
            pr[r] = (P.v0[0] - xs[r][0]) * P.nf[0] + (P.v0[1] - xs[r][1]) * P.nf[1] + (P.v0[2] - xs[r][2]) * P.nf[2];
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const double sq = kTr7S[q], tq = kTr7T[q];
            const double y0 = __builtin_fma(sq, P.e1[0], __builtin_fma(tq, P.e2[0], P.v0[0]));
            const double y1 = __builtin_fma(sq, P.e1[1], __builtin_fma(tq, P.e2[1], P.v0[1]));
            const double y2 = __builtin_fma(sq, P.e1[2], __builtin_fma(tq, P.e2[2], P.v0[2]));
            const double wq = kTr7W[q] * scale;
#pragma unroll
            for (int r = 0; r < FIELD_ROWS; ++r) {
                const double dx = y0 - xs[r][0], dy = y1 - xs[r][1], dz = y2 - xs[r][2];
                const double d2 = dx * dx + dy * dy + dz * dz;
                // a point within 1e-15 of the quadrature point contributes nothing
                const double inv = rsqrt_nr(fmax(d2, 1e-30));
                double sn, cs;
                sincos_red(k * (d2 * inv), &sn, &cs);
                const double wt = d2 > 1e-30 ? wq * inv : 0.0;
                const double zgr = cs * wt, zgi = sn * wt;
                const double drdn = pr[r] * inv;
                // p_j * zg * (-1/r + ik) * drdn
                const double ar = (-zgr * inv - zgi * k) * drdn, ai = (zgr * k - zgi * inv) * drdn;
                acc[r].x = __builtin_fma(pj.x, ar, __builtin_fma(-pj.y, ai, acc[r].x));
                acc[r].y = __builtin_fma(pj.x, ai, __builtin_fma(pj.y, ar, acc[r].y));
                if (HAS_V) {
                    // - v_j * (i pf) * zg = v_j * pf * (zgi - i zgr)
                    const double br = pf * zgi, bi = -pf * zgr;
                    acc[r].x = __builtin_fma(vj.x, br, __builtin_fma(-vj.y, bi, acc[r].x));
                    acc[r].y = __builtin_fma(vj.x, bi, __builtin_fma(vj.y, br, acc[r].y));
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < FIELD_ROWS; ++r) {
        acc[r].x = wave_sum(acc[r].x);
        acc[r].y = wave_sum(acc[r].y);
        if (lane == 0) red[wave][r] = acc[r];
    }
    __syncthreads();
    if (threadIdx.x < FIELD_ROWS && i0 + threadIdx.x < m) {
        double2 t = red[0][threadIdx.x];
        for (int w = 1; w < 4; ++w) t = cadd(t, red[w][threadIdx.x]);
        part[(int64_t)blockIdx.y * m + i0 + threadIdx.x] = t;
    }
}

// Incident pressure and normal derivative of one source at x with normal nx
// (incident.rs:216-420); n == nullptr skips the derivative.  Reference
// operation order, as plane_wave_value.
#pragma clang fp contract(off)
__device__ inline void incident_value(const IncidentSource& s, const double* x, const double* nx, double k, double hf,
                                      double2* p, double2* dpdn) {
    const double2 amp = make_double2(s.amp[0], s.amp[1]);
    if (s.kind == 0) {
        double sn, cs;
        sincos_red(hf * k * (s.v[0] * x[0] + s.v[1] * x[1] + s.v[2] * x[2]), &sn, &cs);
        *p = cmul(amp, make_double2(cs, sn));
        if (nx) *dpdn = cmul(make_double2(0.0, hf * k * (s.v[0] * nx[0] + s.v[1] * nx[1] + s.v[2] * nx[2])), *p);
        return;
    }
    const double dx = x[0] - s.v[0], dy = x[1] - s.v[1], dz = x[2] - s.v[2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    *p = make_double2(0.0, 0.0);
    if (nx) *dpdn = make_double2(0.0, 0.0);
    if (!(r > 1e-10)) return;
    double sn, cs;
    sincos_red(hf * k * r, &sn, &cs);
    const double w = 1.0 / (12.566370614359172 * r);
    const double2 g = make_double2(cs * w, sn * w);
    *p = cmul(amp, g);
    if (nx) {
        const double2 dgdr = cmul(make_double2(-1.0 / r, hf * k), g);
        const double drdn = (dx * nx[0] + dy * nx[1] + dz * nx[2]) / r;
        *dpdn = cmul(amp, cscale(dgdr, drdn));
    }
}

// rhs_i (+)= sign * -(gamma p_inc + beta tau dp_inc/dn) over the sources in order.
__global__ void source_rhs_kernel(const Source* __restrict__ centers, int64_t n, const IncidentSource* __restrict__ srcs,
                                  int64_t nsrc, double k, double hf, double tau_r, double2 beta, double sign,
                                  int accumulate, double2* __restrict__ rhs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Source c = centers[i];
    double2 p = make_double2(0.0, 0.0), q = make_double2(0.0, 0.0);
    for (int64_t w = 0; w < nsrc; ++w) {
        double2 pw, qw;
        incident_value(srcs[w], c.c, c.n, k, hf, &pw, &qw);
        p = cadd(p, pw);
        q = cadd(q, qw);
    }
    // incident.rs:825-831 and bem_solver.rs:5005-5034, as plane_wave_value
    double2 inc = cadd(cmul(make_double2(1.0, 0.0), p), cmul(cmul(beta, make_double2(tau_r, 0.0)), q));
    inc = make_double2(-inc.x, -inc.y);
    inc = cmul(inc, make_double2(sign, 0.0));
    rhs[i] = accumulate ? cadd(rhs[i], inc) : inc;
}
#pragma clang fp contract(on)

// scattered[i] = sum over slices of the partials; incident[i] = plane waves at
// x_i (incident.rs:216-260); total[i] = incident + scattered (FieldPoint::new,
// pressure.rs:32-41).  incident and total may be null.
__global__ void field_finish_kernel(const double2* __restrict__ part, int slices, const double* __restrict__ points,
                                    int64_t m, const IncidentSource* __restrict__ srcs, int64_t nsrc, double k,
                                    double hf, double2* __restrict__ scattered, double2* __restrict__ incident,
                                    double2* __restrict__ total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double2 s = make_double2(0.0, 0.0);
    for (int z = 0; z < slices; ++z) s = cadd(s, part[(int64_t)z * m + i]);
    scattered[i] = s;
    if (!incident && !total) return;
    double2 inc = make_double2(0.0, 0.0);
    for (int64_t w = 0; w < nsrc; ++w) {
        double2 p, unused;
        incident_value(srcs[w], points + 3 * i, nullptr, k, hf, &p, &unused);
        inc = cadd(inc, p);
    }
    if (incident) incident[i] = inc;
    if (total) total[i] = cadd(inc, s);
}

// Surface Cauchy data of each panel from its BC when no velocity is given
// (constant_cauchy_data, pressure.rs:1481-1520): velocity panels (sol, value),
// admittance panels (sol, value + adm sol), pressure panels
// (value - p_inc(c), (sol - dp_inc/dn(c)) / (i pf)) with the plane waves at the
// collocation point.
__global__ void field_cauchy_kernel(const PanelBc* __restrict__ bcs, const Source* __restrict__ src,
                                    const double2* __restrict__ sol, int64_t n, const IncidentSource* __restrict__ srcs,
                                    int64_t nsrc, double k, double hf, double pf, double2* __restrict__ pe,
                                    double2* __restrict__ ve) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const PanelBc bc = bcs[j];
    const double2 s = sol[j];
    const double2 val = make_double2(bc.value[0], bc.value[1]);
    if (bc.type == 1) {
        const Source c = src[j];
        double2 pin = make_double2(0.0, 0.0), qin = make_double2(0.0, 0.0);
        for (int64_t w = 0; w < nsrc; ++w) {
            double2 p, q;
            incident_value(srcs[w], c.c, c.n, k, hf, &p, &q);
            pin = cadd(pin, p);
            qin = cadd(qin, q);
        }
        pe[j] = csub(val, pin);
        const double2 q = csub(s, qin);
        // q / (i pf) as a complex division, zero when pf vanishes
        ve[j] = fabs(pf) > 1e-30 ? make_double2(q.y * pf / (pf * pf), -q.x * pf / (pf * pf)) : make_double2(0.0, 0.0);
    } else {
        pe[j] = s;
        ve[j] = bc.has_adm ? cadd(val, cmul(make_double2(bc.adm[0], bc.adm[1]), s)) : val;
    }
}

__global__ void sincos_probe_kernel(const double* x, int64_t n, double* s, double* c) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) sincos_red(x[t], &s[t], &c[t]);
}

}  // namespace

void upload_quadrature_tables(sbh_context* ctx) {
    double X[9][8] = {}, W[9][8] = {};
    for (int o = 1; o <= 8; ++o) gauss_legendre_rule(o, X[o], W[o]);
    SBH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kGlX), X, sizeof(X)));
    SBH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(kGlW), W, sizeof(W)));
    (void)ctx;
}

void launch_assembly(sbh_workspace* ws, const sbh_physics& ph, sbh_c64 beta_c, sbh_system* sys) {
    injection_check(INJECT_LAUNCH);
    sbh_context* ctx = ws->ctx;
    const sbh_mesh* m = ws->mesh;
    const int64_t n = m->n;
    const double k = ph.wave_number;
    const double wavruim = ph.harmonic_factor * ph.wave_number;
    const double2 beta = make_double2(beta_c.re, beta_c.im);
    const double2 iwr = make_double2(0.0, ph.pressure_factor);
    if (n == 0) return;
    // The sin/cos argument reduction is valid for |k r| < 2^20.
    SBH_REQUIRE(std::fabs(k) * m->diameter < 5.0e5,
                "k * mesh diameter exceeds the supported phase range (5e5 rad)");
    if (ph.harmonic_factor != 1.0)
        throw Error(SBH_ERR_UNSUPPORTED, "only harmonic_factor = +1 (ACOUSTIC_CONVENTION) is supported");
    const int general = m->general ? 1 : 0;
    int rows_per_block = 16;
    while ((n + rows_per_block - 1) / rows_per_block > 65535) rows_per_block *= 2;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)((n + rows_per_block - 1) / rows_per_block));
    // row maxima for a low-precision copy of this system (only while one
    // exists), valid once every kernel below has run
    sys->rowmax_valid = false;
    const int64_t nbx = sys->rowmax_stride;
    const bool row_max = sys->lowp_copies.load() > 0 && rows_per_block <= FAR_MAX_ROWS && nbx == (int64_t)grid.x + 1 && sys->d_rowmax.ptr &&
                         ws->d_row_ptr.ptr && !sys->matrix_exposed;
    double* rm = row_max ? sys->d_rowmax.ptr : nullptr;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_ASSEMBLE_FAR, &ev);
    if (general)
        hipLaunchKernelGGL(far_kernel<1>, grid, dim3(256), 0, ctx->stream, m->d_far.ptr, m->d_src.ptr,
                           m->d_bc.ptr, n, sys->lda, rows_per_block, k, ph.tau, beta, iwr, sys->d_A.ptr,
                           ws->d_row_ptr.ptr, ws->d_pairs.ptr, rm, nbx);
    else
        hipLaunchKernelGGL(far_kernel<0>, grid, dim3(256), 0, ctx->stream, m->d_far.ptr, m->d_src.ptr,
                           m->d_bc.ptr, n, sys->lda, rows_per_block, k, ph.tau, beta, iwr, sys->d_A.ptr,
                           ws->d_row_ptr.ptr, ws->d_pairs.ptr, rm, nbx);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_ASSEMBLE_FAR, ev);
    if (ws->near_pairs > 0) {
        ctx->begin_timed(SBH_KERNEL_ASSEMBLE_NEAR, &ev);
        hipLaunchKernelGGL(near_kernel, dim3((unsigned)((ws->near_pairs + 3) / 4)), dim3(256), 0,
                           ctx->stream, ws->d_pairs.ptr, ws->near_pairs, ws->d_leaves.ptr,
                           m->d_geom.ptr, m->d_src.ptr, m->d_bc.ptr, general, sys->lda, k, wavruim,
                           ph.tau, beta, iwr, sys->d_A.ptr, rm, nbx);
        SBH_HIP(hipGetLastError());
        ctx->end_timed(SBH_KERNEL_ASSEMBLE_NEAR, ev);
    }
    ctx->begin_timed(SBH_KERNEL_ASSEMBLE_SELF, &ev);
    hipLaunchKernelGGL(self_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream,
                       m->d_geom.ptr, m->d_src.ptr, m->d_bc.ptr, general, n, sys->lda, k, wavruim,
                       ph.tau, beta, iwr, sys->d_A.ptr, sys->d_rhs.ptr, rm, nbx);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_ASSEMBLE_SELF, ev);
    sys->rowmax_valid = row_max;
    const int64_t ncols = (int64_t)m->nonzero_bc.size();
    if (ncols > 0) {
        ctx->begin_timed(SBH_KERNEL_RHS, &ev);
        hipLaunchKernelGGL(bc_rhs_kernel, dim3((unsigned)n), dim3(256), 0, ctx->stream,
                           m->d_nonzero_bc.ptr, ncols, ws->d_pairs.ptr, ws->d_row_ptr.ptr,
                           ws->d_leaves.ptr, m->d_geom.ptr, m->d_src.ptr, m->d_bc.ptr, k, wavruim,
                           ph.tau, beta, iwr, sys->d_rhs.ptr);
        SBH_HIP(hipGetLastError());
        ctx->end_timed(SBH_KERNEL_RHS, ev);
    }
}

void launch_source_rhs(sbh_context* ctx, const sbh_mesh* m, const sbh_physics& ph, sbh_c64 beta, double sign,
                       const IncidentSource* d_sources, int64_t nsources, double2* rhs, bool accumulate) {
    const int64_t n = m->n;
    if (n == 0) return;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_RHS, &ev);
    hipLaunchKernelGGL(source_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, m->d_src.ptr,
                       n, d_sources, nsources, ph.wave_number, ph.harmonic_factor, ph.tau,
                       make_double2(beta.re, beta.im), sign, accumulate ? 1 : 0, rhs);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_RHS, ev);
}

void launch_derivative_assembly(sbh_workspace* ws, const sbh_physics& ph, sbh_c64 beta, sbh_c64 beta_d,
                                double sign, double wavenumber_scale, double pf_derivative, double admittance_scale,
                                sbh_system* sys) {
    sbh_context* ctx = ws->ctx;
    const sbh_mesh* m = ws->mesh;
    const int64_t n = m->n;
    if (n == 0) return;
    SBH_REQUIRE(std::fabs(ph.wave_number) * m->diameter < 5.0e5,
                "k * mesh diameter exceeds the supported phase range (5e5 rad)");
    if (ph.harmonic_factor != 1.0)
        throw Error(SBH_ERR_UNSUPPORTED, "only harmonic_factor = +1 (ACOUSTIC_CONVENTION) is supported");
    DerivParams P;
    P.k = ph.wave_number;
    P.wavruim = ph.harmonic_factor * ph.wave_number;
    P.hf = ph.harmonic_factor;
    P.tau_r = ph.tau;
    P.sign = sign;
    P.ws = wavenumber_scale;
    P.adm_scale = admittance_scale;
    P.beta = make_double2(beta.re, beta.im);
    P.beta_d = make_double2(beta_d.re, beta_d.im);
    P.iwr = make_double2(0.0, ph.pressure_factor);
    P.iwr_d = make_double2(0.0, pf_derivative);
    const int general = m->general ? 1 : 0;
    sys->rowmax_valid = false;  // the matrix now holds a derivative
    int rows_per_block = 16;
    while ((n + rows_per_block - 1) / rows_per_block > 65535) rows_per_block *= 2;
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)((n + rows_per_block - 1) / rows_per_block));
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_ASSEMBLE_FAR, &ev);
    if (general)
        hipLaunchKernelGGL(far_deriv_kernel<1>, grid, dim3(256), 0, ctx->stream, m->d_far.ptr, m->d_src.ptr,
                           m->d_bc.ptr, n, sys->lda, rows_per_block, P, sys->d_A.ptr);
    else
        hipLaunchKernelGGL(far_deriv_kernel<0>, grid, dim3(256), 0, ctx->stream, m->d_far.ptr, m->d_src.ptr,
                           m->d_bc.ptr, n, sys->lda, rows_per_block, P, sys->d_A.ptr);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_ASSEMBLE_FAR, ev);
    if (ws->near_pairs > 0) {
        ctx->begin_timed(SBH_KERNEL_ASSEMBLE_NEAR, &ev);
        hipLaunchKernelGGL(near_deriv_kernel, dim3((unsigned)((ws->near_pairs + 3) / 4)), dim3(256), 0, ctx->stream,
                           ws->d_pairs.ptr, ws->near_pairs, ws->d_leaves.ptr, m->d_geom.ptr, m->d_src.ptr,
                           m->d_bc.ptr, general, sys->lda, P, sys->d_A.ptr);
        SBH_HIP(hipGetLastError());
        ctx->end_timed(SBH_KERNEL_ASSEMBLE_NEAR, ev);
    }
    ctx->begin_timed(SBH_KERNEL_ASSEMBLE_SELF, &ev);
    hipLaunchKernelGGL(self_deriv_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, m->d_geom.ptr,
                       m->d_src.ptr, m->d_bc.ptr, general, n, sys->lda, P, sys->d_A.ptr, sys->d_rhs.ptr);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_ASSEMBLE_SELF, ev);
    const int64_t ncols = (int64_t)m->nonzero_bc.size();
    if (ncols > 0) {
        ctx->begin_timed(SBH_KERNEL_RHS, &ev);
        hipLaunchKernelGGL(bc_rhs_deriv_kernel, dim3((unsigned)n), dim3(256), 0, ctx->stream,
                           m->d_nonzero_bc.ptr, ncols, ws->d_pairs.ptr, ws->d_row_ptr.ptr, ws->d_leaves.ptr,
                           m->d_geom.ptr, m->d_src.ptr, m->d_bc.ptr, P, sys->d_rhs.ptr);
        SBH_HIP(hipGetLastError());
        ctx->end_timed(SBH_KERNEL_RHS, ev);
    }
}

void launch_incident_rhs(sbh_context* ctx, const sbh_mesh* m, const sbh_physics& ph, sbh_c64 beta,
                         double sign, const double* d_dirs, const double2* d_amps, int64_t num_rhs,
                         double2* d_rhs, bool accumulate) {
    const int64_t total = m->n * num_rhs;
    if (total == 0) return;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_RHS, &ev);
    hipLaunchKernelGGL(plane_wave_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       ctx->stream, m->d_src.ptr, m->n, d_dirs, d_amps, num_rhs, ph.wave_number,
                       ph.harmonic_factor, ph.tau, make_double2(beta.re, beta.im), sign,
                       accumulate ? 1 : 0, d_rhs);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_RHS, ev);
}

void launch_batch_assembly(sbh_context* ctx, const BatchAssemblyArgs& args, int grid) {
    if (args.nitems == 0) return;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_ASSEMBLE_FAR, &ev);
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(2), 0, ctx->stream, args.counter);
    SBH_HIP(hipGetLastError());
    if (args.nfamilies > 0) {
        const int64_t nb = (args.family_max_n + 255) / 256;
        const int64_t rb = (args.family_max_n + FAMILY_ROWS - 1) / FAMILY_ROWS;
        SBH_REQUIRE(nb < 65536 && rb < 65536 && args.nfamilies < 65536, "frequency-family grid out of range");
        for (int c = 0; c < FAMILY_CLASSES; ++c) {
            const int fb = c ? args.family_class_end[c - 1] : 0, fe = args.family_class_end[c];
            if (fe == fb) continue;
            const dim3 grid((unsigned)nb, (unsigned)rb, (unsigned)(fe - fb));
            const FreqFamily* fams = args.families + fb;
            if (c == 0)
                hipLaunchKernelGGL(batch_far_family_kernel<16>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.far, args.src, args.A);
            else if (c == 1)
                hipLaunchKernelGGL(batch_far_family_kernel<8>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.far, args.src, args.A);
            else if (c == 2)
                hipLaunchKernelGGL(batch_far_family_kernel<4>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.far, args.src, args.A);
            else
                hipLaunchKernelGGL(batch_far_family_kernel<2>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.far, args.src, args.A);
            SBH_HIP(hipGetLastError());
        }
    }
    hipLaunchKernelGGL(batch_assemble_kernel<0>, dim3((unsigned)grid), dim3(256), 0, ctx->stream, args);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_ASSEMBLE_FAR, ev);
    ctx->begin_timed(SBH_KERNEL_ASSEMBLE_NEAR, &ev);
    if (args.nfamilies > 0 && args.family_max_pairs > 0) {
        const int64_t per_block = 256 / NEAR_FAMILY_LANES;
        const int64_t nb = std::min<int64_t>((args.family_max_pairs + per_block - 1) / per_block, 16384);
        for (int c = 0; c < FAMILY_CLASSES; ++c) {
            const int fb = c ? args.family_class_end[c - 1] : 0, fe = args.family_class_end[c];
            if (fe == fb) continue;
            const dim3 grid((unsigned)nb, (unsigned)(fe - fb));
            const FreqFamily* fams = args.families + fb;
            if (c == 0)
                hipLaunchKernelGGL(batch_near_family_kernel<16>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.pairs, args.leaves, args.geom, args.src, args.A);
            else if (c == 1)
                hipLaunchKernelGGL(batch_near_family_kernel<8>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.pairs, args.leaves, args.geom, args.src, args.A);
            else if (c == 2)
                hipLaunchKernelGGL(batch_near_family_kernel<4>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.pairs, args.leaves, args.geom, args.src, args.A);
            else
                hipLaunchKernelGGL(batch_near_family_kernel<2>, grid, dim3(256), 0, ctx->stream, fams, args.inst,
                                   args.pairs, args.leaves, args.geom, args.src, args.A);
            SBH_HIP(hipGetLastError());
        }
    }
    hipLaunchKernelGGL(batch_assemble_kernel<1>, dim3((unsigned)grid), dim3(256), 0, ctx->stream, args);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_ASSEMBLE_NEAR, ev);
}

void launch_exact_apply(const sbh_operator* op, const double2* x, double2* y) {
    const sbh_workspace* ws = op->ws;
    const sbh_mesh* m = ws->mesh;
    sbh_context* ctx = op->ctx;
    const int64_t n = m->n;
    const sbh_physics& ph = op->physics;
    const double k = ph.wave_number;
    const double wavruim = ph.harmonic_factor * ph.wave_number;
    const double2 beta = op->beta;
    const double2 iwr = make_double2(0.0, ph.pressure_factor);
    const int general = m->general ? 1 : 0;
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_MATVEC, &ev);
    const dim3 far_grid((unsigned)((n + EXACT_ROWS - 1) / EXACT_ROWS));
    if (general)
        hipLaunchKernelGGL(exact_far_kernel<1>, far_grid, dim3(256), 0, ctx->stream, m->d_far.ptr, m->d_src.ptr,
                           m->d_bc.ptr, n, k, ph.tau, beta, iwr, x, op->d_far.ptr);
    else
        hipLaunchKernelGGL(exact_far_kernel<0>, far_grid, dim3(256), 0, ctx->stream, m->d_far.ptr, m->d_src.ptr,
                           m->d_bc.ptr, n, k, ph.tau, beta, iwr, x, op->d_far.ptr);
    SBH_HIP(hipGetLastError());
    if (ws->near_pairs > 0) {
        const dim3 grid((unsigned)((ws->near_pairs + 3) / 4));
        if (general)
            hipLaunchKernelGGL(exact_near_kernel<1>, grid, dim3(256), 0, ctx->stream, ws->d_pairs.ptr, ws->near_pairs,
                               ws->d_leaves.ptr, m->d_geom.ptr, m->d_far.ptr, m->d_src.ptr, m->d_bc.ptr, k, wavruim,
                               ph.tau, beta, iwr, x, op->d_near.ptr);
        else
            hipLaunchKernelGGL(exact_near_kernel<0>, grid, dim3(256), 0, ctx->stream, ws->d_pairs.ptr, ws->near_pairs,
                               ws->d_leaves.ptr, m->d_geom.ptr, m->d_far.ptr, m->d_src.ptr, m->d_bc.ptr, k, wavruim,
                               ph.tau, beta, iwr, x, op->d_near.ptr);
        SBH_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(exact_self_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, m->d_geom.ptr,
                       m->d_src.ptr, m->d_bc.ptr, general, n, k, wavruim, ph.tau, beta, iwr, x, op->d_self.ptr);
    SBH_HIP(hipGetLastError());
    hipLaunchKernelGGL(exact_combine_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       op->d_far.ptr, op->d_near.ptr, ws->d_row_ptr.ptr, op->d_self.ptr, n, op->sign, y);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_MATVEC, ev);
}

void launch_field(sbh_context* ctx, const sbh_mesh* m, const sbh_physics& ph, const double2* sol, const double2* v,
                  const double* d_points, int64_t npts, const IncidentSource* d_sources, int64_t nsources,
                  double2* scattered, double2* incident, double2* total) {
    const int64_t n = m->n;
    if (npts == 0) return;
    const double k = ph.wave_number, hf = ph.harmonic_factor, pf = ph.pressure_factor;
    const double2* p = sol;
    DeviceArray<double2> pe, ve;
    if (!v && m->general) {
        pe.allocate((size_t)n);
        ve.allocate((size_t)n);
        hipLaunchKernelGGL(field_cauchy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           m->d_bc.ptr, m->d_src.ptr, sol, n, d_sources, nsources, k, hf, pf, pe.ptr, ve.ptr);
        SBH_HIP(hipGetLastError());
        p = pe.ptr;
        v = ve.ptr;
    }
    const int64_t row_blocks = (npts + FIELD_ROWS - 1) / FIELD_ROWS;
    SBH_REQUIRE(row_blocks < ((int64_t)1 << 31), "too many evaluation points");
    // Enough workgroups to fill the chip: split the panels when the points alone give few.
    const int64_t chunks = (n + 255) / 256;
    const int64_t slices = std::max<int64_t>(1, std::min<int64_t>({chunks, 2048 / row_blocks, 65535}));
    const int64_t per_slice = (chunks + slices - 1) / slices * 256;
    DeviceArray<double2> part((size_t)(slices * npts));
    const dim3 grid((unsigned)row_blocks, (unsigned)slices);
    hipEvent_t ev;
    ctx->begin_timed(SBH_KERNEL_FIELD, &ev);
    if (v)
        hipLaunchKernelGGL(field_kernel<1>, grid, dim3(256), 0, ctx->stream, m->d_far.ptr, n, d_points, npts, p, v,
                           hf * k, pf, per_slice, part.ptr);
    else
        hipLaunchKernelGGL(field_kernel<0>, grid, dim3(256), 0, ctx->stream, m->d_far.ptr, n, d_points, npts, p, v,
                           hf * k, pf, per_slice, part.ptr);
    SBH_HIP(hipGetLastError());
    ctx->end_timed(SBH_KERNEL_FIELD, ev);
    hipLaunchKernelGGL(field_finish_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, ctx->stream, part.ptr,
                       (int)slices, d_points, npts, d_sources, nsources, k, hf, scattered, incident, total);
    SBH_HIP(hipGetLastError());
    SBH_HIP(hipStreamSynchronize(ctx->stream));  // the temporaries are freed on return
}

__global__ void rsqrt_probe_kernel(const double* x, int64_t n, double* y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = rsqrt_nr(x[t]);
}

void launch_rsqrt_probe(sbh_context* ctx, const double* x, int64_t n, double* y) {
    if (n == 0) return;
    hipLaunchKernelGGL(rsqrt_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, x, n, y);
    SBH_HIP(hipGetLastError());
}

void launch_sincos_probe(sbh_context* ctx, const double* x, int64_t n, double* s, double* c) {
    if (n == 0) return;
    hipLaunchKernelGGL(sincos_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, x, n, s, c);
    SBH_HIP(hipGetLastError());
}

}  // namespace sbh
