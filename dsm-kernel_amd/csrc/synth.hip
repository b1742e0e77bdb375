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
// Tile: 128 rows x 128 distances per 512-thread workgroup; 8 waves in a 2x4
// arrangement, each wave 4x2 blocks of v_mfma_f64_16x16x4f64 (64x32).  K
// (= (theta function, l)) is staged through LDS in chunks of 16 with a
// register-staged double buffer: the global loads of chunk t+1 are in
// flight while the MFMAs of chunk t run.  128x128 tiles keep the operand
// stream at ~5 TB/s chip-wide at the FP64 MFMA rate (64x64 tiles needed ~10).
//
// Distance symmetry (MIR = 1, a grid symmetric about 90 degrees): every
// job's theta functions satisfy F(l, pi - theta) = msign (-1)^l F(l, theta),
// so the tiles cover the distances up to 90 degrees only, each K chunk is
// staged with its 8 even degrees first and its 8 odd ones last, the waves
// keep the even and the odd sums apart, and the epilogue writes distance c
// as even + odd and distance theta_n - 1 - c as msign (even - odd).
#include "dev_common.h"
#include "kernels.h"

namespace dsm {

constexpr int BM = 128, BN = 128, BK = 16;

constexpr int NT = 512;
struct TileLoad {
    double2 a[2];  // 64 complex rows x 16 l / 512 threads
    double2 b[2];  // 16 l x 128 theta / 512 threads (pairs)
};

__device__ __forceinline__ void load_tile(const SynthArgs &a, const SynthJob &job, size_t plane, int m0, int n0,
                                          int M, int R, int it, int nk, int tid, TileLoad &t)
{
    const int jb = it / nk, l0 = (it - jb * nk) * BK;
    const double2 *Eq = a.E + (size_t)(job.q0 + jb) * R * a.lcap;
    const double *T = a.tab + (size_t)job.b[jb] * plane;
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int idx = tid + NT * u;
        const int pr = idx >> 4, kk = idx & 15;
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
        const int idx = (tid + NT * u) * 2;
        const int kk = idx >> 7, c = idx & 127;
        t.b[u] = *(const double2 *)(T + (size_t)(l0 + kk) * a.theta_pad + n0 + c);
    }
}

// K slot of degree l0 + kk in LDS: in order, or (MIR) the even degrees of the
// chunk first (l0 is a multiple of BK, so kk has the parity of l)
template <int MIR> __device__ __forceinline__ int kslot(int kk) { return MIR ? ((kk & 1) ? 8 + (kk >> 1) : (kk >> 1)) : kk; }

template <int MIR>
__device__ __forceinline__ void store_tile(double (*As)[BM], double (*Bs)[BN], int tid, const TileLoad &t)
{
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int idx = tid + NT * u;
        const int pr = idx >> 4, kk = kslot<MIR>(idx & 15);
        As[kk][2 * pr] = t.a[u].x;
        As[kk][2 * pr + 1] = t.a[u].y;
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int idx = (tid + NT * u) * 2;
        const int kk = kslot<MIR>(idx >> 7), c = idx & 127;
        Bs[kk][c] = t.b[u].x;
        Bs[kk][c + 1] = t.b[u].y;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Every job's tiles (row panel major, distance tile minor) are cut into
// 8 equal contiguous slices, one per XCD, and each XCD walks its slices job
// by job: the tiles resident on one XCD share A panels (same row panel) and
// B panels (neighbouring row panels), and all XCDs get the same work.
// Returns false for the padding slots of the last slice.
__device__ __forceinline__ bool xcd_tile(int nx, int ny, int &tx, int &ty, int &tz)
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

// accumulator registers per lane: [1 + MIR][4][2] v4d
template <int MIR> constexpr int kpart_regs() { return (1 + MIR) * 4 * 2 * 4; }

// Split K (a.kphase 1): workgroup (tile, part) sums the chunks
// [pp kchunk, (pp + 1) kchunk) of theta function jb = part / ksplit (pp =
// part % ksplit) and stores its accumulators in kpart; a.kphase 2: one
// workgroup per tile adds the parts in order and runs the epilogue.  The
// parts are fixed degree ranges and their sums are added in a fixed order,
// so the result depends neither on the scheduling nor on the tile's other
// rows (chunks past a row's series add exact zeros).
template <int KIND, int MIR>  // KIND 0: receiver rows, 1: surface rows (separate symbol for profiling)
__global__ __launch_bounds__(512) void synth_f64_kernel(SynthArgs a, int nx, int ny)
{
    __shared__ double As[2][BK][BM];
    __shared__ double Bs[2][BK][BN];
    int tx, ty, tz;
    if (!xcd_tile(nx, ny, tx, ty, tz)) return;
    const SynthJob job = a.jobs[tz];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int R = a.r_eff;
    const int M = a.nslot * R * 2;
    const int m0 = ty * BM, n0 = tx * BN;
    // K extent: the longest l series among the tile's frequencies
    int kmax = 0;
    {
        const int s0 = m0 / (2 * R), s1 = min(M - 1, m0 + BM - 1) / (2 * R);
        for (int sl = s0; sl <= s1; sl++) kmax = max(kmax, a.nl[sl]);
    }
    const int nk = (kmax + BK - 1) / BK;
    const int part = (a.kphase == 1) ? (int)blockIdx.y : 0;
    int it0 = 0, niter = (a.kphase == 2) ? 0 : job.nb * nk;
    if (a.kphase == 1) {
        const int jbp = part / a.ksplit, c0 = (part - jbp * a.ksplit) * a.kchunk;
        const int c1 = min(nk, c0 + a.kchunk);
        it0 = jbp * nk + c0;
        niter = (jbp < job.nb && c1 > c0) ? jbp * nk + c1 : it0;
    }
    const size_t tile_id = ((size_t)tz * ny + ty) * nx + tx;
    const int nparts = a.knb * a.ksplit;
    const size_t plane = (size_t)a.lcap * a.theta_pad;
    const int wr = wave >> 2, wc = wave & 3;

    // acc[0]: all degrees (MIR = 0) or the even ones; acc[1]: the odd ones
    v4d acc[1 + MIR][4][2];
#pragma unroll
    for (int p = 0; p <= MIR; p++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[p][i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

    TileLoad tl;
    if (niter > it0) {
        load_tile(a, job, plane, m0, n0, M, R, it0, nk, tid, tl);
        store_tile<MIR>(As[0], Bs[0], tid, tl);
    }
    __syncthreads();
    for (int it = it0; it < niter; it++) {
        const int st = (it - it0) & 1;
        if (it + 1 < niter) load_tile(a, job, plane, m0, n0, M, R, it + 1, nk, tid, tl);
#pragma unroll
        for (int k4 = 0; k4 < BK; k4 += 4) {
            const int p = (MIR && k4 >= 8) ? 1 : 0;
            const int kk = k4 + (lane >> 4);
            double av[4], bv[2];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[st][kk][wr * 64 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; j++) bv[j] = Bs[st][kk][wc * 32 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[p][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[p][i][j], 0, 0, 0);
        }
        if (it + 1 < niter) store_tile<MIR>(As[st ^ 1], Bs[st ^ 1], tid, tl);
        __syncthreads();
    }
    constexpr int NREG = kpart_regs<MIR>();
    if (a.kphase == 1) {  // partial sums of this part, [tile][part][reg][thread]
        double *kp = a.kpart + (tile_id * nparts + part) * NREG * NT + tid;
#pragma unroll
        for (int p = 0; p <= MIR; p++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) kp[(size_t)(((p * 4 + i) * 2 + j) * 4 + q) * NT] = acc[p][i][j][q];
        return;
    }
    if (a.kphase == 2) {  // the parts' sums, in part order
        for (int pt = 0; pt < nparts; pt++) {
            const double *kp = a.kpart + (tile_id * nparts + pt) * NREG * NT + tid;
#pragma unroll
            for (int p = 0; p <= MIR; p++)
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
#pragma unroll
                        for (int q = 0; q < 4; q++) acc[p][i][j][q] += kp[(size_t)(((p * 4 + i) * 2 + j) * 4 + q) * NT];
        }
    }
    // epilogue: row = (slot, r, re|im) -> out[pos][comp][r][theta].
    // v_mfma_f64_16x16x4f64 leaves D[i][j] in lane j + 16*(i%4), register i/4.
    void *out = job.out;
    const int ntgt = job.comp[1] >= 0 ? 2 : 1;
    int bad = 0;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = m0 + wr * 64 + i * 16 + (lane >> 4) + 4 * q;
            if (row >= M) continue;
            const int pr = row >> 1, reim = row & 1;
            const int sl = pr / R, r = pr - sl * R;
            const size_t pos = a.slot_out[sl];
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                if (col >= (MIR ? a.nh : a.theta_n)) continue;
                const int mcol = a.theta_n - 1 - col;  // (MIR) the mirrored distance
                for (int t = 0; t < ntgt; t++) {
                    const size_t base = ((pos * job.ncomp + job.comp[t]) * R + r) * a.theta_n;
                    if (!MIR) {
                        const double v = job.sign[t] * acc[0][i][j][q];
                        bad += !__builtin_isfinite(cat_store_part(out, base + col, reim, v, job.accumulate, job.c64));
                    } else {
                        const double e = acc[0][i][j][q], od = acc[MIR][i][j][q];
                        const double v = job.sign[t] * (e + od);
                        bad += !__builtin_isfinite(cat_store_part(out, base + col, reim, v, job.accumulate, job.c64));
                        if (mcol != col) {
                            const double w = job.sign[t] * (job.msign * (e - od));
                            bad += !__builtin_isfinite(
                                cat_store_part(out, base + mcol, reim, w, job.accumulate, job.c64));
                        }
                    }
                }
            }
        }
    add_nonfinite(a.nonfinite, bad);
}

static void synth_f64_geometry(const SynthArgs &a, int &nx, int &ny)
{
    const int M = a.nslot * a.r_eff * 2;
    nx = a.mirror ? (a.nh + BN - 1) / BN : a.theta_pad / BN;
    ny = (M + BM - 1) / BM;
}

size_t synth_f64_tiles(const SynthArgs &a)
{
    int nx, ny;
    synth_f64_geometry(a, nx, ny);
    return (size_t)nx * ny * a.njobs;
}

size_t synth_f64_kpart_size(const SynthArgs &a)
{
    return synth_f64_tiles(a) * (size_t)(a.ksplit > 1 ? a.ksplit * a.knb : 0) * kpart_regs<1>() * NT;
}

static void synth_f64_launch(const SynthArgs &a, dim3 grid, int nx, int ny, hipStream_t s)
{
    if (a.mirror) {
        if (a.r_eff > 1) hipLaunchKernelGGL((synth_f64_kernel<0, 1>), grid, dim3(NT), 0, s, a, nx, ny);
        else hipLaunchKernelGGL((synth_f64_kernel<1, 1>), grid, dim3(NT), 0, s, a, nx, ny);
    } else {
        if (a.r_eff > 1) hipLaunchKernelGGL((synth_f64_kernel<0, 0>), grid, dim3(NT), 0, s, a, nx, ny);
        else hipLaunchKernelGGL((synth_f64_kernel<1, 0>), grid, dim3(NT), 0, s, a, nx, ny);
    }
}

void launch_synth_f64(const SynthArgs &a, hipStream_t s)
{
    int nx, ny;
    synth_f64_geometry(a, nx, ny);
    const unsigned g = 8 * ((nx * ny + 7) / 8) * a.njobs;
    SynthArgs b = a;
    if (a.ksplit > 1 && a.kpart && a.kchunk > 0 && a.knb > 0) {
        b.kphase = 1;
        synth_f64_launch(b, dim3(g, a.ksplit * a.knb), nx, ny, s);
        b.kphase = 2;
        synth_f64_launch(b, dim3(g), nx, ny, s);
    } else {
        b.kphase = 0;
        synth_f64_launch(b, dim3(g), nx, ny, s);
    }
}

}  // namespace dsm
