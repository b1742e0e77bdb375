// sh_kernels.hip -- SH (toroidal) per-(omega,l) kernels: band solve and the
// coefficient rows of the synthesis contraction.
#include "dev_common.h"
#include "kernels.h"

namespace dsm {

__global__ __launch_bounds__(64) void sh_solve_kernel(ShSolveArgs a)
{
    const int small = sh_solve_lane(a, blockIdx.x, threadIdx.x);
    if (small && a.small_pivots) atomicAdd(a.small_pivots, (unsigned long long)small);
}

void launch_sh_solve(const ShSolveArgs &a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(sh_solve_kernel, dim3(nblocks), dim3(64), 0, s, a);
}

// Depth cut (mixed forms), the P-SV rule on SH's node rows: for the
// smallest degree of a block of 64 (the least decay), D = sum over node
// intervals below the source of min(dcap, sqrt((l+1/2)^2/r^2 - omega^2/v^2)
// dr) where positive; rows below the first interval that takes D past dcut
// are dropped, and at least three rows below the source are kept.  The
// lanes take 64 intervals at a time going down from the source, a wave
// prefix sum finds the crossing.
__global__ __launch_bounds__(64) void sh_depth_kernel(ShDepthArgs a)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= a.nslots * a.nlb) return;
    const int ks = t / a.nlb, blk = t - ks * a.nlb;
    const int sl = a.slot_list[ks];
    const double om = a.omega[sl];
    const double al = (double)max(1, a.l0 + 64 * blk) + 0.5;
    int jlo = 1;
    double carry = 0.0;
    for (int j0 = a.k; j0 >= 2; j0 -= 64) {
        const int j = j0 - lane;  // interval between rows j-1 and j (1-based)
        double c = 0.0;
        if (j >= 2) {
            const double rm = 0.5 * (a.rrow[j - 2] + a.rrow[j - 1]);
            const double vm = 0.5 * (a.vrow[j - 2] + a.vrow[j - 1]);
            const double dr = a.rrow[j - 1] - a.rrow[j - 2];
            const double q = al / rm, pp = om / vm;
            const double k2 = q * q - pp * pp;
            if (k2 > 0.0 && dr > 0.0) c = fmin(sqrt(k2) * dr, a.dcap);
        }
        double ps = c;  // inclusive prefix sum, lane 0 the interval nearest the source
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(ps, d, 64);
            if (lane >= d) ps += o;
        }
        const unsigned long long over = __ballot(j >= 2 && carry + ps > a.dcut);
        if (over) {
            jlo = j0 - (__ffsll((long long)over) - 1);
            break;
        }
        carry += __shfl(ps, 63, 64);
    }
    jlo = max(1, min(jlo, a.k - 3));
    if (lane == 0) a.jstart[t] = jlo;
}

void launch_sh_depth(const ShDepthArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(sh_depth_kernel, dim3(a.nslots * a.nlb), dim3(64), 0, s, a);
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
