// dev_common.h -- device helpers: complex FP64 arithmetic on double2.
#pragma once

#include <hip/hip_runtime.h>

namespace dsm {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline double2 cz() { return make_double2(0.0, 0.0); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by i*s
__host__ __device__ inline double2 cmul_i(double2 a, double s) { return make_double2(-a.y * s, a.x * s); }
// Complex quotient by Smith's method -- the algorithm of the C/Fortran
// runtime complex division (libgcc __divdc3, compiler-rt) that the reference
// and the oracle use for every complex quotient, including the reciprocal
// pivots of the band factorization.  Keeping it makes this code follow the
// reference's rounding where the problem is ill-conditioned.
__host__ __device__ inline double2 cdiv(double2 p, double2 q)
{
    const double a = p.x, b = p.y, c = q.x, d = q.y;
    if (fabs(c) < fabs(d)) {
        const double ratio = c / d;
        const double denom = (c * ratio) + d;
        return make_double2(((a * ratio) + b) / denom, ((b * ratio) - a) / denom);
    } else {
        const double ratio = d / c;
        const double denom = (d * ratio) + c;
        return make_double2(((b * ratio) + a) / denom, (b - (a * ratio)) / denom);
    }
}
__host__ __device__ inline double2 crecip(double2 q) { return cdiv(make_double2(1.0, 0.0), q); }
__host__ __device__ inline double cabs2(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }
// a + s*b with real s
__host__ __device__ inline double2 caxpy(double2 a, double s, double2 b)
{
    return make_double2(a.x + s * b.x, a.y + s * b.y);
}

// ---- bfloat16 operand terms of the mixed forms: x = t0 + t1 (+ t2) +
// residual, each term the round-to-nearest-even bf16 of what the previous
// ones leave (the remainder is exact in FP64).
__host__ __device__ inline unsigned short bf16_rn(float f)
{
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline double bf16_value(unsigned short h)
{
    const unsigned u = (unsigned)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return (double)f;
}
__host__ __device__ inline void bf16_terms(double x, int ns, unsigned short *t)
{
    double r = x;
    for (int s = 0; s < ns; s++) {
        t[s] = bf16_rn((float)r);
        r -= bf16_value(t[s]);
    }
}
#ifdef __HIPCC__  // device code only (the host set-up includes this header too)
// catalog stores of the synthesis: a non-finite value counts 1
__device__ __forceinline__ int nonfinite2(double2 v) { return !(__builtin_isfinite(v.x) && __builtin_isfinite(v.y)); }
// Catalog element idx (complex): FP64 (c64 = 0) or complex64 (c64 = 1, the
// files' precision: the mixed forms keep their catalog so).  accum adds v
// into the stored value (the sum in FP64, then rounded).  Returns what was
// stored.
__device__ __forceinline__ double2 cat_store(void *out, size_t idx, double2 v, int accum, int c64)
{
    if (c64) {
        float2 *o = (float2 *)out + idx;
        if (accum) {
            const float2 w = *o;
            v = make_double2((double)w.x + v.x, (double)w.y + v.y);
        }
        const float2 f = make_float2((float)v.x, (float)v.y);
        *o = f;
        return make_double2((double)f.x, (double)f.y);
    }
    double2 *o = (double2 *)out + idx;
    if (accum) {
        const double2 w = *o;
        v = make_double2(w.x + v.x, w.y + v.y);
    }
    *o = v;
    return v;
}
// The real (reim 0) or imaginary part of catalog element idx, as cat_store.
__device__ __forceinline__ double cat_store_part(void *out, size_t idx, int reim, double v, int accum, int c64)
{
    if (c64) {
        float *o = (float *)out + 2 * idx + reim;
        const double s = accum ? (double)*o + v : v;
        const float f = (float)s;
        *o = f;
        return (double)f;
    }
    double *o = (double *)out + 2 * idx + reim;
    const double s = accum ? *o + v : v;
    *o = s;
    return s;
}
__device__ __forceinline__ void add_nonfinite(unsigned long long *count, int n)
{
    if (n && count) atomicAdd(count, (unsigned long long)n);
}
#endif
// degree order of the bf16 operands: even l first, then odd l
__host__ __device__ inline int lperm(int l, int lcap) { return (l & 1) ? (lcap >> 1) + (l >> 1) : (l >> 1); }
// Layout of the bf16 operands of the synthesis: planes of M rows of lcap
// degree positions (lperm), stored chunk-major -- the 32 positions of one
// chunk of every row of the plane together, [plane][lcap / 32][M][32] -- so
// that one K step of a synthesis tile (16 rows x 32 degrees per wave load)
// is 1 KiB of contiguous memory.  lcap is a multiple of 64 (the even and
// the odd halves start on a chunk); a buffer of one-degree rows (lcap 1)
// is a plain [plane][M] array.
__host__ __device__ inline size_t b16_at(size_t plane, size_t M, size_t m, int p, int lcap)
{
    if (lcap < 32) return (plane * M + m) * lcap + p;
    return ((plane * (size_t)(lcap >> 5) + (p >> 5)) * M + m) * 32 + (p & 31);
}
// Destination of the bf16 coefficient rows: Eb, planes q ns + term of
// receiver-major rows m = (r nslot + slot) 2 + re|im.
struct B16Rows {
    unsigned short *Eb;
    int ns, nslot, r_n, lcap;
    __host__ __device__ void put(int q, int r, int slot, int l, double2 v) const
    {
        const size_t M = (size_t)nslot * r_n * 2;
        const size_t m = ((size_t)r * nslot + slot) * 2;
        const int lp = lperm(l, lcap);
        unsigned short tr[3], ti[3];
        bf16_terms(v.x, ns, tr);
        bf16_terms(v.y, ns, ti);
        for (int s = 0; s < ns; s++) {
            const size_t pl = (size_t)q * ns + s;
            Eb[b16_at(pl, M, m, lp, lcap)] = tr[s];
            Eb[b16_at(pl, M, m + 1, lp, lcap)] = ti[s];
        }
    }
};

// ---- operation families: F = false keeps the reference's rounding
// (separate products and sums, Smith quotients); F = true is the fast FP64
// family of the mixed forms (fused multiply-adds, reciprocal by v_rcp_f64
// and two Newton steps, ~1 ulp, no lane divergence).
__host__ __device__ inline double rcp_newton(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
#else
    double r = 1.0 / x;
#endif
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
template <bool F>
__host__ __device__ inline double2 cmul_t(double2 a, double2 b)
{
    if (!F) return cmul(a, b);
    return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
// c + a b
template <bool F>
__host__ __device__ inline double2 cmadd_t(double2 c, double2 a, double2 b)
{
    if (!F) return cadd(c, cmul(a, b));
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
template <bool F>
__host__ __device__ inline double2 crecip_t(double2 q)
{
    if (!F) return crecip(q);
    const double r = rcp_newton(fma(q.x, q.x, q.y * q.y));
    return make_double2(q.x * r, -q.y * r);
}
// a + s b with real s
template <bool F>
__host__ __device__ inline double2 caxpy_t(double2 a, double s, double2 b)
{
    if (!F) return caxpy(a, s, b);
    return make_double2(fma(s, b.x, a.x), fma(s, b.y, a.y));
}

}  // namespace dsm
