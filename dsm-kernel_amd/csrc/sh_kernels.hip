// sh_kernels.hip -- SH (toroidal) per-(omega,l) kernels: band solve and the
// coefficient rows of the synthesis contraction.
#include <climits>

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
    if (a.lstop && a.l0 + 64 * blk >= a.lstop[sl]) return;  // past the slot's cut-off: not solved
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

// The l cut-off of SH (calamp, src/SGTsh/others.f90:805-824, in the order
// of SGTsh-MPI.f90:423-478; the host form is calamp in engine.cpp) over the
// degrees [l0, lend] of one chunk, one 64-lane workgroup per slot, 64
// degrees at a time.  Per degree the entries are m = -2, -1, 1, 2 with
// |m| <= l, of amplitude a = amp[l] at |m| = 2 and 0 at |m| = 1.  The
// running maximum starts at -1 and the first entry (l = 1, amplitude 0)
// lifts it to 0, so at an entry of degree l it is the maximum of 0 and the
// amplitudes through l (a wave prefix maximum, carried across chunks); an
// entry is small when its ratio (a / maximum; 0 for a = 0) is below ratl and
// l > lsuf, and a small entry adds one to the count, any other resets it.
// So a degree maps the count x before its entries to
//   l = 0:  x;
//   l = 1:  x + 2 if 0 < ratl and l > lsuf, else 0;
//   l >= 2: x + 4 if ratio < ratl and l > lsuf, else 0 (the m = 2 entry,
//           not small, resets the count last),
// maps of the form "set c" or "add k", composed by a wave scan; the cut-off
// is the first degree that starts with a count above 20.  A slot whose
// cut-off an earlier chunk found is left as it is.
__global__ __launch_bounds__(64) void sh_cut_kernel(ShCutArgs a)
{
    const int k = blockIdx.x, tx = threadIdx.x;
    if (a.lstop[k] != INT_MAX) return;
    const int lsuf = a.lsuf[k];
    double carry = a.carry[k];
    int xin = a.count[k];
    for (int b0 = a.l0; b0 <= a.lend; b0 += 64) {
        const int l = b0 + tx;
        const bool in = l <= a.lend;
        const double a0 = (in && l >= 2) ? a.amp[(size_t)k * a.lcap + l] : 0.0;
        double pm = a0;  // inclusive prefix maximum over the 64 degrees
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(pm, d, 64);
            if (tx >= d) pm = fmax(pm, o);
        }
        const double mx = fmax(carry, pm);
        const double ratio = (a0 != 0.0 && mx != 0.0) ? a0 / mx : 0.0;
        const bool past = l > lsuf;
        int set = 0, val = 0;  // lanes past lend and l = 0: add 0
        if (in && l == 1) {
            set = !(0.0 < a.ratl && past);
            val = set ? 0 : 2;
        } else if (in && l >= 2) {
            const bool f = ratio < a.ratl && past;
            set = !f;
            val = f ? 4 : 0;
        }
        for (int d = 1; d < 64; d <<= 1) {  // inclusive scan of "earlier, then this"
            const int os = __shfl_up(set, d, 64);
            const int ov = __shfl_up(val, d, 64);
            if (tx >= d) {
                val = set ? val : ov + val;
                set = set | os;
            }
        }
        const int after = set ? val : xin + val;  // count after this degree's entries
        const int before0 = __shfl_up(after, 1, 64);
        const int before = (tx == 0) ? xin : before0;
        const unsigned long long stop = __ballot(in && before > 20);
        if (stop) {
            if (tx == 0) a.lstop[k] = b0 + (__ffsll((long long)stop) - 1);
            return;
        }
        carry = fmax(carry, __shfl(pm, 63, 64));
        xin = __shfl(after, 63, 64);
    }
    if (tx == 0) {
        a.carry[k] = carry;
        a.count[k] = xin;
        if (a.lend >= a.maxlmax) a.lstop[k] = a.maxlmax + 1;  // no cut-off: the series runs to lmax
    }
}

void launch_sh_cut(const ShCutArgs &a, int nslots, hipStream_t s)
{
    hipLaunchKernelGGL(sh_cut_kernel, dim3(nslots), dim3(64), 0, s, a);
}

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
