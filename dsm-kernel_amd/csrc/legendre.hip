// legendre.hip -- FP64 associated-Legendre recurrence and the real
// theta-basis functions of the synthesis contraction.
//
// Computed once per run (the reference recomputes the recurrence for every
// frequency, SGTsh-MPI.f90:406,425-427): the functions depend only on
// (l, |m|, theta).  One lane per distance runs the upward recurrence in l
// (calplm, src/SGTsh/trial.f90:85-117, state P_l, P_{l-1}, P_{l-2}), so the
// recurrence is bit-for-bit the reference's sequence of operations.
#include "dev_common.h"
#include "kernels.h"

#include "legendre_lane.h"

namespace dsm {

__global__ void legendre_sh_kernel(const double *theta_rad, int theta_n, int theta_pad, int lmax, int lcap,
                                   double *tab)
{
    legendre_sh_column(theta_rad, theta_n, theta_pad, lmax, lcap, tab, blockIdx.x * blockDim.x + threadIdx.x);
}

void launch_legendre_sh(const double *theta_rad, int theta_n, int theta_pad, int lmax, int lcap, double *tab,
                        hipStream_t s)
{
    hipLaunchKernelGGL(legendre_sh_kernel, dim3((theta_pad + 63) / 64), dim3(64), 0, s, theta_rad, theta_n,
                       theta_pad, lmax, lcap, tab);
}

// P-SV table in two passes: (1) per distance and m, the recurrence of
// P_l^m (m = 0..3) over l, the only serial part, into P[m][l][theta];
// (2) one thread per (l, distance): F0..F9 from those values
// (legendre_psv_values, the per-degree work of legendre_psv_column).
__global__ __launch_bounds__(64) void legendre_plm_kernel(const double *theta_rad, int theta_n, int lmax, int lcap,
                                                          int theta_pad, double *P)
{
    // one lane per (distance, m): the recurrences of different m are independent
    const int it = blockIdx.x * 64 + threadIdx.x, m = blockIdx.y;
    if (it >= theta_n) return;
    const double th = theta_rad[it];
    if (th == 0.0 || th == DSM_PI) return;  // poles: no recurrence
    const double x = cos(th);
    double plm[3] = {0, 0, 0};
    double *Pm = P + (size_t)m * lcap * theta_pad + it;
    for (int l = 0; l <= lmax; l++) {
        if (l >= m) calplm(l, m, x, plm);
        Pm[(size_t)l * theta_pad] = plm[0];
    }
}

__global__ __launch_bounds__(256) void legendre_psv_fill_kernel(const double *theta_rad, int theta_n, int theta_pad,
                                                                int lmax, int lcap, const double *P, double *tab)
{
    const int it = blockIdx.x * 64 + (threadIdx.x & 63);
    const int l = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (it >= theta_pad || l >= lcap) return;
    const size_t plane = (size_t)lcap * theta_pad;
    double f[3][10];
    if (it >= theta_n || l > lmax) {
        for (int mu = 0; mu < 3; mu++)
            for (int k = 0; k < 10; k++) f[mu][k] = 0.0;
    } else {
        const double th = theta_rad[it];
        const bool pole = (th == 0.0 || th == DSM_PI);
        const double x = cos(th), s = sin(th);
        const double sf = (double)(float)s, cf = (double)(float)(x / s);
        double pl[4] = {0.0, 0.0, 0.0, 0.0};
        if (!pole)
            for (int m = 0; m < 4; m++) pl[m] = P[m * plane + (size_t)l * theta_pad + it];
        legendre_psv_values(l, th, pole, x, s, sf, cf, pl, f);
    }
    legendre_psv_store(f, tab, plane, l, theta_pad, it);
}

// P: scratch of 4 lcap theta_pad doubles
void launch_legendre_psv(const double *theta_rad, int theta_n, int theta_pad, int lmax, int lcap, double *P,
                         double *tab, hipStream_t s)
{
    hipLaunchKernelGGL(legendre_plm_kernel, dim3((theta_n + 63) / 64, 4), dim3(64), 0, s, theta_rad, theta_n, lmax,
                       lcap, theta_pad, P);
    hipLaunchKernelGGL(legendre_psv_fill_kernel, dim3((theta_pad + 63) / 64, (lcap + 3) / 4), dim3(256), 0, s,
                       theta_rad, theta_n, theta_pad, lmax, lcap, P, tab);
}

}  // namespace dsm
