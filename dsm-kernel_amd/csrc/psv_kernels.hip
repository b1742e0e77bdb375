// psv_kernels.hip -- P-SV (spheroidal) per-(omega,l) kernels: band solve,
// coefficient rows of the synthesis contraction and the l = 0 system.
// Built without FMA contraction (see psv_lane.h).
#include "dev_common.h"
#include "kernels.h"

namespace dsm {

__global__ __launch_bounds__(64) void psv_solve_kernel(PsvSolveArgs a) { psv_solve_lane(a, blockIdx.x, threadIdx.x); }

void launch_psv_solve(const PsvSolveArgs &a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(psv_solve_kernel, dim3(nblocks), dim3(64), 0, s, a);
}

__global__ __launch_bounds__(64) void psv_coef_kernel(PsvCoefArgs a) { psv_coef_lane(a, blockIdx.x, threadIdx.x); }

void launch_psv_coef(const PsvCoefArgs &a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(psv_coef_kernel, dim3(nblocks), dim3(64), 0, s, a);
}

// l = 0: one 64-lane workgroup per frequency, the phases of psv_lane.h
// separated by barriers
__global__ __launch_bounds__(64) void psv_l0_kernel(PsvL0Args a)
{
    const int t = blockIdx.x, tx = threadIdx.x;
    for (int i = tx; i < a.nn0; i += 64) psv_l0_assemble(a, t, i);
    __syncthreads();
    if (tx == 0) psv_l0_factor(a, t);
    __syncthreads();
    if (tx < 3) psv_l0_subst(a, t, tx);
    __syncthreads();
    for (int r = tx; r <= a.r_n; r += 64) psv_l0_rows(a, t, r);
}

void launch_psv_l0(const PsvL0Args &a, hipStream_t s)
{
    hipLaunchKernelGGL(psv_l0_kernel, dim3(a.nslots), dim3(64), 0, s, a);
}

}  // namespace dsm
