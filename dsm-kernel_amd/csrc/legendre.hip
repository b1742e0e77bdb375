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

__global__ void legendre_psv_kernel(const double *theta_rad, int theta_n, int theta_pad, int lmax, int lcap,
                                    double *tab)
{
    legendre_psv_column(theta_rad, theta_n, theta_pad, lmax, lcap, tab, blockIdx.x * blockDim.x + threadIdx.x);
}

void launch_legendre_psv(const double *theta_rad, int theta_n, int theta_pad, int lmax, int lcap, double *tab,
                         hipStream_t s)
{
    hipLaunchKernelGGL(legendre_psv_kernel, dim3((theta_pad + 63) / 64), dim3(64), 0, s, theta_rad, theta_n,
                       theta_pad, lmax, lcap, tab);
}

}  // namespace dsm
