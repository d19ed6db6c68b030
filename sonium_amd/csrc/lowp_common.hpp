// Complex FP16/BF16 storage helpers shared by the single-system and batched
// low-precision kernels (device code, included inside namespace sbh::{anon}).
//
// Row scaling: row i is stored as A_ij * 2^(T - e_i) with 2^e_i <= max_j
// max(|Re A_ij|, |Im A_ij|) < 2^(e_i + 1); T = 14 for FP16 (the largest scaled
// entry stays below 2^15) and 0 for BF16 (FP32 exponent range).  row_inv
// holds 2^(e_i - T), which undoes the scaling exactly.
#pragma once

constexpr int BLOCK = 256;

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

template <int FMT>
struct Half;
// from_float rounds an FP32 value (nearest even).  The empty asm pins f as an
// FP32 register value: without it the compiler folds (half)(float)double into
// one software double->half rounding, which is both slower than the hardware
// conversions and a different rounding (no FP32 intermediate).
template <>
struct Half<0> {
    static constexpr int TOP = 14;
    __device__ static u16 from_float(float f) {
        asm("" : "+v"(f));
        _Float16 h = (_Float16)f;
        return __builtin_bit_cast(u16, h);
    }
    __device__ static float to_float(u16 v) { return (float)__builtin_bit_cast(_Float16, v); }
};
template <>
struct Half<1> {
    static constexpr int TOP = 0;
    __device__ static u16 from_float(float f) {
        asm("" : "+v"(f));
        __bf16 h = (__bf16)f;
        return __builtin_bit_cast(u16, h);
    }
    __device__ static float to_float(u16 v) { return (float)__builtin_bit_cast(__bf16, v); }
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// One block of NT threads converts one row: row maximum, power-of-two scale,
// then every entry as {re, im} halves packed in one 32-bit word (row padded
// with zeros to ld entries).  A row of at most NT * CONV_HELD entries is read
// once and held in registers between the maximum and the conversion (thread t
// holds entries t, t + NT, ...); a longer row is read twice, CONV_LOADS loads
// in flight per thread per pass.
constexpr int CONV_HELD = 12;
constexpr int CONV_LOADS = 8;

template <int NT>
__device__ __forceinline__ double block_max(double m) {
    __shared__ double sh[NT / 64];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    m = sh[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) m = fmax(m, sh[w]);
    return m;
}

template <int FMT>
__device__ __forceinline__ unsigned pack_scaled(double2 v, double scale) {
    const unsigned re = Half<FMT>::from_float((float)(v.x * scale));
    const unsigned im = Half<FMT>::from_float((float)(v.y * scale));
    return re | (im << 16);
}

template <int FMT>
__device__ __forceinline__ double row_scale(double m, double* __restrict__ inv_out) {
    int e = 0;
    if (m > 0.0 && isfinite(m)) e = ilogb(m);
    if (threadIdx.x == 0) *inv_out = ldexp(1.0, e - Half<FMT>::TOP);
    return ldexp(1.0, Half<FMT>::TOP - e);
}

template <int FMT, int NT>
__device__ __forceinline__ void convert_row(const double2* __restrict__ row, int64_t n, int64_t ld,
                                            u16* __restrict__ o, double* __restrict__ inv_out) {
    unsigned* __restrict__ o32 = reinterpret_cast<unsigned*>(o);
    const int t = threadIdx.x;
    if (n <= (int64_t)NT * CONV_HELD) {
        double2 v[CONV_HELD];
        double m = 0.0;
#pragma unroll
        for (int u = 0; u < CONV_HELD; ++u) {
            const int64_t j = t + (int64_t)NT * u;
            v[u] = j < n ? row[j] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < CONV_HELD; ++u) m = fmax(m, fmax(fabs(v[u].x), fabs(v[u].y)));
        const double scale = row_scale<FMT>(block_max<NT>(m), inv_out);
#pragma unroll
        for (int u = 0; u < CONV_HELD; ++u) {
            const int64_t j = t + (int64_t)NT * u;
            if (j < n) o32[j] = pack_scaled<FMT>(v[u], scale);
        }
        for (int64_t j = n + t; j < ld; j += NT) o32[j] = 0u;
        return;
    }
    double m = 0.0;
    int64_t j0 = t;
    for (; j0 + (int64_t)NT * (CONV_LOADS - 1) < n; j0 += (int64_t)NT * CONV_LOADS) {
        double2 v[CONV_LOADS];
#pragma unroll
        for (int u = 0; u < CONV_LOADS; ++u) v[u] = row[j0 + (int64_t)NT * u];
#pragma unroll
        for (int u = 0; u < CONV_LOADS; ++u) m = fmax(m, fmax(fabs(v[u].x), fabs(v[u].y)));
    }
    for (; j0 < n; j0 += NT) {
        const double2 v = row[j0];
        m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
    }
    const double scale = row_scale<FMT>(block_max<NT>(m), inv_out);
    j0 = t;
    for (; j0 + (int64_t)NT * (CONV_LOADS - 1) < n; j0 += (int64_t)NT * CONV_LOADS) {
        double2 v[CONV_LOADS];
#pragma unroll
        for (int u = 0; u < CONV_LOADS; ++u) v[u] = row[j0 + (int64_t)NT * u];
#pragma unroll
        for (int u = 0; u < CONV_LOADS; ++u) o32[j0 + (int64_t)NT * u] = pack_scaled<FMT>(v[u], scale);
    }
    for (; j0 < ld; j0 += NT) o32[j0] = j0 < n ? pack_scaled<FMT>(row[j0], scale) : 0u;
}

// v_mfma_f32_32x32x16_{f16,bf16}: A 32x16 (lane l: row l%32, k 8*(l/32)..+7),
// B 16x32 (lane l: column l%32, the same k), C 32x32 FP32 in 16 registers.
template <int FMT>
__device__ __forceinline__ f32x16 mfma32(u32x4 a, u32x4 b, f32x16 c) {
    if constexpr (FMT == 0)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

// v_dot2_f32_{f16,bf16}: c + a.lo * b.lo + a.hi * b.hi over two packed halves.
template <int FMT>
__device__ __forceinline__ float dot2(unsigned a, unsigned b, float c) {
    if constexpr (FMT == 0)
        return __builtin_amdgcn_fdot2(__builtin_bit_cast(h2, a), __builtin_bit_cast(h2, b), c, false);
    else
        return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf2, a), __builtin_bit_cast(bf2, b), c, false);
}

// Column-batched reductions over n x k row-major blocks: the sum of the S
// partials sh[q * k + c] (q < S) of every column c lands in sh[c].  It is a
// pairwise tree of fixed shape (width S, then ceil(width / 2), ...), so the
// order is fixed and no thread adds more than one value per level.  The caller
// has written sh; every thread of the block calls this.
__device__ __forceinline__ double2 add2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double add2(double a, double b) { return a + b; }
template <typename V>
__device__ __forceinline__ void column_tree(V* sh, int S, int k, int sub, int c) {
    for (int width = S; width > 1;) {
        const int half = (width + 1) / 2;
        __syncthreads();
        if (sub < width - half) sh[sub * k + c] = add2(sh[sub * k + c], sh[(sub + half) * k + c]);
        width = half;
    }
    __syncthreads();
}

// column_tree over nv slices of BLOCK values each (slice v at sh + v * BLOCK),
// one barrier per level for all of them; per slice the additions are those of
// column_tree.
template <typename V>
__device__ __forceinline__ void column_tree_multi(V* sh, int S, int k, int sub, int c, int nv) {
    for (int width = S; width > 1;) {
        const int half = (width + 1) / 2;
        __syncthreads();
        if (sub < width - half)
            for (int v = 0; v < nv; ++v) {
                V* p = sh + v * BLOCK;
                p[sub * k + c] = add2(p[sub * k + c], p[(sub + half) * k + c]);
            }
        width = half;
    }
    __syncthreads();
}
