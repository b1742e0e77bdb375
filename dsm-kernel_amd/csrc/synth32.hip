// synth32.hip -- the synthesis contraction in mixed precision on MFMA.
//
// Same contraction and job table as synth.hip (FP64).  The operands are
// rounded to FP32 (the coefficient rows while they are staged into LDS, the
// theta functions once per run) and multiplied on v_mfma_f32_16x16x4f32;
// every 32-wide K chunk of FP32 partial sums is added into FP64 accumulators,
// so the rounding error does not grow with the length of the l series.
// Tile: 128 x 128 per 512-thread workgroup, 8 waves in 2x4, each 4x2 blocks
// of 16x16; register-staged double buffer as in synth.hip.
#include "dev_common.h"
#include "kernels.h"

namespace dsm {

namespace {

constexpr int BM = 128, BN = 128, BK = 32, NT = 512;

struct TileLoad32 {
    double2 a[4];  // 64 complex rows x 32 l / 512 threads
    float4 b[2];   // 32 l x 128 theta / 512 threads (quads)
};

__device__ __forceinline__ void load_tile32(const SynthArgs &a, const SynthJob &job, const float *tab32,
                                            size_t plane, int m0, int n0, int M, int R, int it, int nk, int tid,
                                            TileLoad32 &t)
{
    const int jb = it / nk, l0 = (it - jb * nk) * BK;
    const double2 *Eq = a.E + (size_t)(job.q0 + jb) * R * a.lcap;
    const float *T = tab32 + (size_t)job.b[jb] * plane;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int idx = tid + NT * u;
        const int pr = idx >> 5, kk = idx & 31;
        const int prow = (m0 >> 1) + pr;
        double2 v = make_double2(0.0, 0.0);
        if (2 * prow < M) {
            const int sl = prow / R, r = prow - sl * R;
            v = Eq[((size_t)sl * a.nq * R + r) * a.lcap + l0 + kk];
        }
        t.a[u] = v;
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int idx = (tid + NT * u) * 4;
        const int kk = idx >> 7, c = idx & 127;
        t.b[u] = *(const float4 *)(T + (size_t)(l0 + kk) * a.theta_pad + n0 + c);
    }
}

__device__ __forceinline__ void store_tile32(float (*As)[BM], float (*Bs)[BN], int tid, const TileLoad32 &t)
{
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int idx = tid + NT * u;
        const int pr = idx >> 5, kk = idx & 31;
        As[kk][2 * pr] = (float)t.a[u].x;
        As[kk][2 * pr + 1] = (float)t.a[u].y;
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int idx = (tid + NT * u) * 4;
        const int kk = idx >> 7, c = idx & 127;
        *(float4 *)&Bs[kk][c] = t.b[u];
    }
}

// XCD-sliced tile order, as in synth.hip
__device__ __forceinline__ bool xcd_tile32(int nx, int ny, int &tx, int &ty, int &tz)
{
    constexpr int NXCD = 8;
    const int T = nx * ny, cpx = (T + NXCD - 1) / NXCD;
    const int pid = blockIdx.x, xcd = pid % NXCD, local = pid / NXCD;
    tz = local / cpx;
    const int t = xcd * cpx + (local - tz * cpx);
    if (t >= T) return false;
    ty = t / nx;
    tx = t - ty * nx;
    return true;
}

}  // namespace

template <int KIND>
__global__ __launch_bounds__(512) void synth_f32_kernel(SynthArgs a, const float *tab32, int nx, int ny)
{
    __shared__ float As[2][BK][BM];
    __shared__ float Bs[2][BK][BN];
    int tx, ty, tz;
    if (!xcd_tile32(nx, ny, tx, ty, tz)) return;
    const SynthJob job = a.jobs[tz];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int R = a.r_eff;
    const int M = a.nslot * R * 2;
    const int m0 = ty * BM, n0 = tx * BN;
    int kmax = 0;
    {
        const int s0 = m0 / (2 * R), s1 = min(M - 1, m0 + BM - 1) / (2 * R);
        for (int sl = s0; sl <= s1; sl++) kmax = max(kmax, a.nl[sl]);
    }
    const int nk = (kmax + BK - 1) / BK;
    const int niter = job.nb * nk;
    const size_t plane = (size_t)a.lcap * a.theta_pad;
    const int wr = wave >> 2, wc = wave & 3;

    v4f acc[4][2];
    v4d acc64[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
            acc64[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
        }

    TileLoad32 tl;
    if (niter > 0) {
        load_tile32(a, job, tab32, plane, m0, n0, M, R, 0, nk, tid, tl);
        store_tile32(As[0], Bs[0], tid, tl);
    }
    __syncthreads();
    for (int it = 0; it < niter; it++) {
        const int st = it & 1;
        if (it + 1 < niter) load_tile32(a, job, tab32, plane, m0, n0, M, R, it + 1, nk, tid, tl);
#pragma unroll
        for (int k4 = 0; k4 < BK; k4 += 4) {
            const int kk = k4 + (lane >> 4);
            float av[4], bv[2];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[st][kk][wr * 64 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; j++) bv[j] = Bs[st][kk][wc * 32 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        // fold the chunk's FP32 sums into FP64
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    acc64[i][j][q] += (double)acc[i][j][q];
                    acc[i][j][q] = 0.f;
                }
        if (it + 1 < niter) store_tile32(As[st ^ 1], Bs[st ^ 1], tid, tl);
        __syncthreads();
    }
    // epilogue: v_mfma_f32_16x16x4f32 leaves D[i][j] in lane j + 16*(i/4), register i%4
    // out[pos][comp][r][theta] (complex): registers 2h, 2h+1 are re, im of
    // one (slot, r) row, 16 lanes write 16 consecutive distances
    void *out = job.out;
    const int ntgt = job.comp[1] >= 0 ? 2 : 1;
    int bad = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++) {
            const int row = m0 + wr * 64 + i * 16 + 4 * (lane >> 4) + 2 * h2;
            if (row >= M) continue;
            const int pr = row >> 1;
            const int sl = pr / R, r = pr - sl * R;
            const size_t pos = a.slot_out[sl];
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                if (col >= a.theta_n) continue;
                for (int t = 0; t < ntgt; t++) {
                    const size_t o = ((pos * job.ncomp + job.comp[t]) * R + r) * a.theta_n + col;
                    const double2 v = make_double2(job.sign[t] * acc64[i][j][2 * h2], job.sign[t] * acc64[i][j][2 * h2 + 1]);
                    bad += nonfinite2(cat_store(out, o, v, job.accumulate, job.c64));
                }
            }
        }
    add_nonfinite(a.nonfinite, bad);
}

__global__ void tab_to_f32_kernel(const double *src, float *dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

void launch_tab_to_f32(const double *src, float *dst, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(tab_to_f32_kernel, dim3(4096), dim3(256), 0, s, src, dst, n);
}

void launch_synth_f32(const SynthArgs &a, const float *tab32, hipStream_t s)
{
    const int M = a.nslot * a.r_eff * 2;
    const int nx = a.theta_pad / BN, ny = (M + BM - 1) / BM;
    const dim3 grid(8 * ((nx * ny + 7) / 8) * a.njobs);
    if (a.r_eff > 1) hipLaunchKernelGGL(synth_f32_kernel<0>, grid, dim3(NT), 0, s, a, tab32, nx, ny);
    else hipLaunchKernelGGL(synth_f32_kernel<1>, grid, dim3(NT), 0, s, a, tab32, nx, ny);
}

}  // namespace dsm
