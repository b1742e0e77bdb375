// sh_kernels.hip -- SH (toroidal) per-(omega,l) kernels: band solve and the
// coefficient rows of the synthesis contraction.
#include "dev_common.h"
#include "kernels.h"

namespace dsm {

__global__ __launch_bounds__(64) void sh_solve_kernel(ShSolveArgs a) { sh_solve_lane(a, blockIdx.x, threadIdx.x); }

void launch_sh_solve(const ShSolveArgs &a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(sh_solve_kernel, dim3(nblocks), dim3(64), 0, s, a);
}

__global__ __launch_bounds__(64) void sh_coef_kernel(ShCoefArgs a) { sh_coef_lane(a, blockIdx.x, threadIdx.x); }

void launch_sh_coef(const ShCoefArgs &a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(sh_coef_kernel, dim3(nblocks), dim3(64), 0, s, a);
}

// Zero E[slot][row][l] for l >= nl[slot] (the solve chunks run past the cut-off).
__global__ void zero_tail_kernel(double2 *E, int rows_per_slot, const int *slots, const int *nl, int lcap)
{
    const int row = blockIdx.y;
    const int sl = slots[blockIdx.z];
    const int n = nl[sl];
    double2 *p = E + ((size_t)sl * rows_per_slot + row) * lcap;
    for (int l = n + blockIdx.x * blockDim.x + threadIdx.x; l < lcap; l += gridDim.x * blockDim.x) p[l] = cz();
}

void launch_zero_tail(double2 *E, int nslots, int rows_per_slot, const int *slots, const int *nl, int lcap,
                      hipStream_t s)
{
    hipLaunchKernelGGL(zero_tail_kernel, dim3(4, rows_per_slot, nslots), dim3(256), 0, s, E, rows_per_slot, slots,
                       nl, lcap);
}

}  // namespace dsm
