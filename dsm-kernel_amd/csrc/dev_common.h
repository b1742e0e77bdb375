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

}  // namespace dsm
