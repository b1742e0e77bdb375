// synth.hip -- the strain synthesis as a real GEMM on FP64 MFMA.
//
// The reference evaluates, for every (omega,l,m,source,receiver,theta), the
// displacement gradient from the vector harmonics and adds the strain maps
// into the catalog (SGTsh-MPI.f90:491-549; SGTpsv-MPI.f90:917-1084).  Here the
// l-dependent complex coefficients of every catalog component are formed
// once per (omega,l,receiver) (sh_coef) and the sum over l becomes
//     out[(omega,r,re|im)][theta] = sum_{j,l} A[(omega,r,re|im)][(j,l)] * T[(j,l)][theta]
// with A the re/im parts of the coefficient rows and T the real theta
// functions of the component's |m| (legendre.hip).
//
// Tile: 64 rows x 64 distances per 256-thread workgroup, 4 waves in a 2x2
// arrangement, each wave 2x2 MFMA 16x16x4 FP64 blocks; K staged through LDS
// in chunks of 16.
#include "dev_common.h"
#include "kernels.h"

namespace dsm {

constexpr int BM = 64, BN = 64, BK = 16;

__global__ __launch_bounds__(256) void synth_f64_kernel(SynthArgs a)
{
    __shared__ double As[BK][BM];
    __shared__ double Bs[BK][BN];
    __shared__ int kmax_s;
    const SynthJob job = a.jobs[blockIdx.z];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int R = a.r_eff;
    const int M = a.nslot * R * 2;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    if (tid == 0) kmax_s = 0;
    __syncthreads();
    if (tid < BM && m0 + tid < M) atomicMax(&kmax_s, a.nl[(m0 + tid) / (2 * R)]);
    __syncthreads();
    const int kmax = kmax_s;
    const double *tab_mu = a.tab + (size_t)(job.mu - 1) * 6 * a.lcap * a.theta_pad;
    const size_t plane = (size_t)a.lcap * a.theta_pad;

    v4d acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    const int wr = wave >> 1, wc = wave & 1;

    for (int jb = 0; jb < job.nb; jb++) {
        const double2 *Eq = a.E + (size_t)(job.q0 + jb) * R * a.lcap;
        const double *T = tab_mu + (size_t)job.b[jb] * plane;
        for (int l0 = 0; l0 < kmax; l0 += BK) {
            // A tile: 32 complex rows x 16 l
            for (int u = 0; u < 2; u++) {
                const int idx = tid + 256 * u;
                const int pr = idx >> 4, kk = idx & 15;
                const int prow = (m0 >> 1) + pr;  // (slot, r) pair index
                double2 v = make_double2(0.0, 0.0);
                if (2 * prow < M) {
                    const int sl = prow / R, r = prow - sl * R;
                    v = Eq[((size_t)sl * a.nq * R + r) * a.lcap + l0 + kk];
                }
                As[kk][2 * pr] = v.x;
                As[kk][2 * pr + 1] = v.y;
            }
            // B tile: 16 l x 64 theta
            for (int u = 0; u < 2; u++) {
                const int idx = (tid + 256 * u) * 2;
                const int kk = idx >> 6, c = idx & 63;
                const double2 v = *(const double2 *)(T + (size_t)(l0 + kk) * a.theta_pad + n0 + c);
                Bs[kk][c] = v.x;
                Bs[kk][c + 1] = v.y;
            }
            __syncthreads();
#pragma unroll
            for (int k4 = 0; k4 < BK; k4 += 4) {
                const int kk = k4 + (lane >> 4);
                double av[2], bv[2];
                for (int i = 0; i < 2; i++) av[i] = As[kk][wr * 32 + i * 16 + (lane & 15)];
                for (int j = 0; j < 2; j++) bv[j] = Bs[kk][wc * 32 + j * 16 + (lane & 15)];
                for (int i = 0; i < 2; i++)
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
            }
            __syncthreads();
        }
    }
    // epilogue: row = (slot, r, re|im) -> out[pos][r][theta][comp].
    // v_mfma_f64_16x16x4f64 leaves D[i][j] in lane j + 16*(i%4), register i/4.
    double *out = (double *)job.out;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < 4; q++) {
                const int row = m0 + wr * 32 + i * 16 + (lane >> 4) + 4 * q;
                const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                if (row < M && col < a.theta_n) {
                    const int pr = row >> 1, reim = row & 1;
                    const int sl = pr / R, r = pr - sl * R;
                    const size_t o = (((size_t)a.slot_out[sl] * R + r) * a.theta_n + col) * job.ncomp + job.comp;
                    out[2 * o + reim] = acc[i][j][q];
                }
            }
}

void launch_synth_f64(const SynthArgs &a, hipStream_t s)
{
    const int M = a.nslot * a.r_eff * 2;
    dim3 grid(a.theta_pad / BN, (M + BM - 1) / BM, a.njobs);
    hipLaunchKernelGGL(synth_f64_kernel, grid, dim3(256), 0, s, a);
}

}  // namespace dsm
