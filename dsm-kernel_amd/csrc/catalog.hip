// catalog.hip -- the catalog records of a run in the reference's file layout.
//
// The synthesis writes component-major rows, [pos][comp][rec][theta], so
// that a tile stores whole rows of distances: complex FP64 in the plain
// form, complex64 in the mixed forms.  The files hold one complex(4) array
// a(num, theta_n) per (frequency, receiver), component fastest
// (SGTpsv-MPI.f90:1113-1149, SGTsh-MPI.f90:577-611), cast from the FP64
// accumulators.  pack_records_kernel does that cast (plain form) and the
// transpose on the device, so the host receives each frequency's records as
// one block:
//   [TSGT: rec][theta][num_t] [RSGT: rec][theta][num_r] [SYNN: theta][num_s]
// HBM-bound: 16 (8) B read and 8 B written per value, staged through LDS so
// that both sides are contiguous.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dsm {

namespace {
constexpr int PK_T = 64;        // distances per workgroup
constexpr int PK_MAXC = 20;     // components per record (TSGT_PSV)
constexpr int PK_NT = 256;

__global__ __launch_bounds__(PK_NT) void pack_records_kernel(PackArgs a)
{
    __shared__ float2 tile[PK_MAXC][PK_T + 1];
    const int t0 = blockIdx.x * PK_T, rec = blockIdx.y, k = blockIdx.z;
    const int pos = a.pos_list[k];
    const int nt = min(PK_T, a.theta_n - t0), nc = a.ncomp;
    const size_t s0 = ((size_t)pos * nc * a.nrec + rec) * a.theta_n + t0;
    const size_t cstride = (size_t)a.nrec * a.theta_n;
    for (int idx = threadIdx.x; idx < nc * PK_T; idx += PK_NT) {
        const int c = idx / PK_T, t = idx - c * PK_T;
        if (t < nt) {
            if (a.c64) {
                tile[c][t] = ((const float2 *)a.src)[s0 + c * cstride + t];
            } else {
                const double2 v = ((const double2 *)a.src)[s0 + c * cstride + t];
                tile[c][t] = make_float2((float)v.x, (float)v.y);
            }
        }
    }
    __syncthreads();
    float2 *dst = a.dst + k * a.dst_stride + a.dst_off + ((size_t)rec * a.theta_n + t0) * nc;
    for (int o = threadIdx.x; o < nt * nc; o += PK_NT) {
        const int t = o / nc, c = o - t * nc;
        dst[o] = tile[c][t];
    }
}
}  // namespace

void launch_pack_records(const PackArgs &a, int npos, hipStream_t s)
{
    if (npos <= 0 || a.theta_n <= 0 || a.nrec <= 0) return;
    const dim3 g((a.theta_n + PK_T - 1) / PK_T, a.nrec, npos);
    hipLaunchKernelGGL(pack_records_kernel, g, dim3(PK_NT), 0, s, a);
}

}  // namespace dsm
