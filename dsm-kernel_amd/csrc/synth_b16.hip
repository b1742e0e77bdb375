// synth_b16.hip -- the synthesis contraction of the mixed forms on
// v_mfma_f32_16x16x32_bf16.
//
// Operands.  Every FP64 operand x (coefficient rows E, theta functions T)
// is split into NS bfloat16 terms x = x0 + x1 (+ x2) + residual, each term
// the bf16 rounding of what the previous ones leave (residual ~2^-17 |x|
// for NS = 2, ~2^-25 |x| for NS = 3).  The product is the sum of the
// cross products of order < NS:
//   NS = 2 (BF16X3): x0 y0 + x0 y1 + x1 y0
//   NS = 3 (BF16X6): the above + x0 y2 + x1 y1 + x2 y0
// issued small terms first into one FP32 MFMA accumulator per 32-wide K
// chunk, which is then added into FP64; so the error does not grow with
// the length of the l series.  (Two chunks per FP32 accumulator measured
// no faster and ~3x the error on the 5 s P-SV catalog.)
//
// Degree order.  Both operands store the degrees permuted, even l first
// (l -> l/2) and odd l after (l -> lcap/2 + l/2), so every K chunk holds
// degrees of one parity.  On a distance grid symmetric about 90 degrees
// the functions obey F(l, pi - theta) = sigma (-1)^l F(l, theta) with
// sigma = +-1 fixed per catalog component (the parity of the function
// times (-1)^|m|), so with e and o the even-l and odd-l partial sums at
// theta,
//   out(theta) = e + o,   out(pi - theta) = sigma (e - o),
// and only the columns theta <= 90 degrees are computed.  The even half is
// accumulated first and parked in a per-workgroup stash (coalesced, one
// FP64 per lane and register) while the odd half accumulates in the same
// registers.
//
// Rows are receiver-major, m = (r nslot + slot) 2 + re|im, and every
// (receiver, slot) row carries its own l extent nlr (the l cut-off, and in
// the batched forms the depth cut: a receiver below the solved rows of a
// block of l has no coefficients there), so a tile's K extent is the
// longest series among its rows.
//
// Tile: 128 x 128 per 512-thread workgroup, 8 waves in 2 x 4, each wave
// 4 x 2 blocks of 16 x 16.  Operands go from global memory straight into
// LDS (global_load_lds_dwordx4: a wave writes 1 KiB, 16 rows of 32
// degrees) in a ring of three stages, two K steps ahead of the MFMAs; a
// wave waits for its own loads of the current stage (s_waitcnt vmcnt) and
// the barrier makes every wave's loads visible.  Rows are 64 bytes, their
// four 16-byte degree segments permuted per row (sw_seg), so the 16-byte
// fragment reads of 8 consecutive rows hit distinct banks (a permutation
// that also spread 16 rows measured 4% slower in this kernel).
#include <climits>

#include "dev_common.h"
#include "kernels.h"

namespace dsm {

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short u16;

constexpr int BM = 128, BN = 128, BK = 32, NT = 512;
constexpr int NSTAGE = 3;

// degree segment (8 bf16) seg of tile row `row` sits at LDS segment
// seg ^ sw_seg(row): rows 2k, 2k+1 share a permutation, so any 8
// consecutive rows take all four, each on two rows 4 segments apart
__device__ __forceinline__ int sw_seg(int row) { return (row >> 1) & 3; }

#define LDS_PTR(x) ((__attribute__((address_space(3))) void *)(x))

// s_waitcnt: vmcnt <= N (the loads of later stages stay in flight), all LDS
// reads returned (lgkmcnt 0), expcnt not waited (gfx9 encoding)
template <int N>
__device__ __forceinline__ void wait_vm_lgkm0()
{
    static_assert(N >= 0 && N < 16, "vmcnt field");
    __builtin_amdgcn_s_waitcnt(N | (0x7 << 4));
}

// XCD-sliced tile order, as in synth.hip
__device__ __forceinline__ bool xcd_tile_b16(int nx, int ny, int &tx, int &ty, int &tz)
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

// E [slot][q][r][lcap] (FP64 complex) -> Eb, planes q ns + s (bf16 terms,
// degrees permuted, b16_at), receiver-major rows m = (r nslot + slot) 2 + re|im,
// for the degrees a synthesis tile reads: l < nl(slot) split, l in
// [nl(slot), lz) zero, lz = min(lcap, longest nl + 128) (a tile reads its
// rows' degrees in chunks of 32 even and 32 odd ones up to its longest
// series; degrees from lz up keep what earlier batches left).  Rows in
// turn per workgroup, lanes along l.
__global__ __launch_bounds__(256) void split_rows_kernel(const double2 *E, int nslot, int nq, int R, int lcap, int ns,
                                                         u16 *Eb, const int *nl)
{
    __shared__ int smax;
    if (threadIdx.x == 0) smax = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < nslot; i += blockDim.x) atomicMax(&smax, nl[i]);
    __syncthreads();
    const int lz = min(lcap, smax + 128);
    const size_t rows = (size_t)nslot * nq * R;
    const B16Rows dst{Eb, ns, nslot, R, lcap};
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const int r = (int)(row % R);
        const size_t rest = row / R;
        const int q = (int)(rest % nq);
        const int sl = (int)(rest / nq);
        const int n = nl[sl];
        const double2 *src = E + row * lcap;
        for (int l = threadIdx.x; l < lz; l += blockDim.x)
            dst.put(q, r, sl, l, l < n ? src[l] : make_double2(0.0, 0.0));
    }
}

// zero the degrees l in [nl(slot), min(lend, longest nl + 128)] of every
// row (the l cut-off).  A synthesis tile reads its rows' degrees in chunks
// of 32 even and 32 odd ones up to its longest series, so never past the
// batch's longest cut-off + 128; rows from earlier batches may hold other
// values there.  Rows in turn per workgroup, lanes along l (the even and
// the odd degrees of a row are two contiguous runs).
__global__ __launch_bounds__(256) void zero_tail_b16_kernel(u16 *Eb, int nq, int ns, int nslot, int R, int lcap,
                                                            const int *nl, int lend)
{
    __shared__ int smax, smin;
    if (threadIdx.x == 0) {
        smax = 0;
        smin = INT_MAX;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nslot; i += blockDim.x) {
        atomicMax(&smax, nl[i]);
        atomicMin(&smin, nl[i]);
    }
    __syncthreads();
    const int lz = min(lend, smax + 128);
    if (smin > lz) return;  // every slot runs past what the tiles read
    const size_t rows = (size_t)nq * ns * nslot * R * 2, M = (size_t)nslot * R * 2;
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const size_t m = row % M, pl = row / M;
        const int sl = (int)((m >> 1) % nslot);
        for (int l = nl[sl] + (int)threadIdx.x; l <= lz; l += blockDim.x) Eb[b16_at(pl, M, m, lperm(l, lcap), lcap)] = 0;
    }
}

// degree 0 of nrows bf16 rows (planes of M rows) from one-degree rows
// (lperm(0, lcap) = 0)
__global__ void b16_put_l0_kernel(const u16 *one, u16 *Eb, size_t nrows, size_t M, int lcap)
{
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (size_t)gridDim.x * blockDim.x)
        Eb[b16_at(r / M, M, r % M, 0, lcap)] = one[r];
}

// zero the padding degrees l in [lfrom, lcap) of nrows rows (planes of M
// rows; never written)
__global__ void zero_pad_b16_kernel(u16 *Eb, size_t nrows, size_t M, int lcap, int lfrom)
{
    const int np = lcap - lfrom;
    const size_t total = nrows * (size_t)np;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t row = idx / np;
        const int l = lfrom + (int)(idx - row * np);
        Eb[b16_at(row / M, M, row % M, lperm(l, lcap), lcap)] = 0;
    }
}

// tab [plane][lcap][theta_pad] (FP64) -> Tb, planes s nplanes + plane of
// nhp rows (b16_at), row c from distance c < nh, the others zero.  A transpose through LDS:
// a workgroup reads 64 degrees x 64 distances of one plane along the
// distances and writes, per split term and distance, the tile's 32 even
// and 32 odd degrees (two contiguous runs of the permuted row) along l.
// lcap is a multiple of 64.
constexpr int ST_L = 64, ST_C = 64, ST_NT = 256;
__global__ __launch_bounds__(ST_NT) void split_table_kernel(const double *tab, int lcap, int theta_pad, int nh, int nhp,
                                                            int ns, u16 *Tb, int nplanes)
{
    __shared__ u16 t[3][ST_C][ST_L + 4];
    const int l0 = blockIdx.x * ST_L, c0 = blockIdx.y * ST_C, p = blockIdx.z;
    const double *src = tab + ((size_t)p * lcap + l0) * theta_pad + c0;
    for (int idx = threadIdx.x; idx < ST_L * ST_C; idx += ST_NT) {
        const int ll = idx / ST_C, cc = idx - ll * ST_C;
        const double v = (c0 + cc < nh) ? src[(size_t)ll * theta_pad + cc] : 0.0;
        u16 tr[3];
        bf16_terms(v, ns, tr);
        for (int sp = 0; sp < ns; sp++) t[sp][cc][ll] = tr[sp];
    }
    __syncthreads();
    // per (term, distance, parity) 32 degrees = 8 lanes x 4 values
    const int LH = lcap >> 1;
    for (int idx = threadIdx.x; idx < ns * ST_C * 2 * 8; idx += ST_NT) {
        const int q = idx & 7, par = (idx >> 3) & 1, rest = idx >> 4;
        const int cc = rest % ST_C, sp = rest / ST_C;
        u16 *dst = Tb + b16_at((size_t)sp * nplanes + p, nhp, c0 + cc, (par ? LH : 0) + (l0 >> 1) + 4 * q, lcap);
        ushort4 w;
        w.x = t[sp][cc][2 * (4 * q + 0) + par];
        w.y = t[sp][cc][2 * (4 * q + 1) + par];
        w.z = t[sp][cc][2 * (4 * q + 2) + par];
        w.w = t[sp][cc][2 * (4 * q + 3) + par];
        *(ushort4 *)dst = w;
    }
}

// One K step of the tile into LDS stage `dst` (u16 offset base): 2 NS 8
// groups of 16 rows (A: splits of the coefficient rows, then B: splits of
// the table columns), group g by wave g % 8, 1 KiB per group.  loff[u]:
// the lane's row offset (degree segment included) in group wave + 8 u;
// ea, tb: the job's coefficient-row and table-plane bases at degree lp0.
// with_b false: the wave loads its A groups only (NS loads instead of 2 NS).
// Nothing here may wait on a vector load: the loads of the later stages
// are in flight.
template <int NS>
__device__ __forceinline__ void issue_stage(const u16 *ea, const u16 *tb, size_t a_split, size_t t_split,
                                            const unsigned *loff, int wave, bool with_b, u16 *dst)
{
#pragma unroll
    for (int u = 0; u < NS * 2; u++) {
        const int g = wave + 8 * u;
        if (g >= NS * 8 && !with_b) continue;
        const int s = (g >> 3) % NS;
        const u16 *src = (g < NS * 8) ? ea + s * a_split : tb + s * t_split;
        __builtin_amdgcn_global_load_lds((const void *)(src + loff[u]), LDS_PTR(dst + g * 512), 16, 0, 0);
    }
}

// One 128 x 128 tile (rows m0.., distance columns n0..) of job `job`.
// Wave tiles: 64 x 32 (4 x 2 blocks of 16 x 16; waves in 2 x 4), or, for
// the last distance tile when it holds at most 32 columns (NARROW), 16 x 32
// (1 x 2 blocks; waves stacked along the rows): the tile's MFMAs cover the
// 32 columns only, and only waves 0 and 1 load table rows (the 32 columns).
template <int NS, bool NARROW>
__device__ __forceinline__ void synth_b16_tile(const SynthB16Args &a, const SynthJob *job, int tile, int m0, int n0,
                                               u16 *lds)
{
    constexpr int STAGE = NS * (BM + BN) * BK;  // u16
    constexpr int WI = NARROW ? 1 : 4;          // 16-row blocks per wave
    const int nb = job->nb;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int R = a.r_eff;
    const int M = a.nslot * R * 2;
    int kmax = 0;  // longest l series among the tile's (receiver, slot) rows
    for (int pr = m0 >> 1; pr <= (min(M - 1, m0 + BM - 1) >> 1); pr++) kmax = max(kmax, a.nlr[pr]);
    const int ce = ((kmax + 1) / 2 + BK - 1) / BK;  // chunks of even degrees l < kmax
    const int co = (kmax / 2 + BK - 1) / BK;        // odd
    const int LH = a.lcap >> 1;
    const int nit_e = nb * ce, nit = nit_e + nb * co;
    const int rbase = NARROW ? wave * 16 : (wave >> 2) * 64;  // the wave's first row, column in the tile
    const int cbase = NARROW ? 0 : (wave & 3) * 32;
    const bool with_b = !NARROW || wave < 2;
    const int fr = lane & 15;

    v4d acc64[WI][2];
#pragma unroll
    for (int i = 0; i < WI; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc64[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    // stash of tile t (< nx ny njobs, within synth_b16_grid)
    double *stash = a.stash + (((size_t)tile * 8 + wave) * 32) * 64 + lane;
    bool stashed = false;

    auto itmap = [&](int it, int &jb, int &lp0) {
        if (it < nit_e) {
            jb = it / ce;
            lp0 = (it - jb * ce) * BK;
        } else {
            const int u = it - nit_e;
            jb = u / co;
            lp0 = LH + (u - jb * co) * BK;
        }
    };

    // loads: the job's fields in scalar registers, the lanes' row offsets
    // (rows past M read row M - 1: their products land in rows not stored)
    const int q0 = __builtin_amdgcn_readfirstlane(job->q0);
    const int pl0 = __builtin_amdgcn_readfirstlane(job->b[0]), pl1 = __builtin_amdgcn_readfirstlane(job->b[1]);
    const int pl2 = __builtin_amdgcn_readfirstlane(job->b[2]), pl3 = __builtin_amdgcn_readfirstlane(job->b[3]);
    unsigned loff[2 * NS];
#pragma unroll
    for (int u = 0; u < 2 * NS; u++) {
        const int g = wave + 8 * u;
        const int row = (g & 7) * 16 + (lane >> 2);
        const int seg = (lane & 3) ^ sw_seg(row);
        const int src_row = (g < NS * 8) ? min(m0 + row, M - 1) : n0 + row;
        loff[u] = (unsigned)src_row * 32u + 8 * seg;  // within the chunk of the stage's 32 degrees
    }
    const size_t a_split = (size_t)M * a.lcap, t_split = (size_t)a.nplanes * a.nhp * a.lcap;
    auto issue = [&](int it2) {
        int jb, lp0;
        itmap(it2, jb, lp0);
        const int plane = jb == 0 ? pl0 : jb == 1 ? pl1 : jb == 2 ? pl2 : pl3;
        // the stage's chunk of degrees lp0 .. lp0 + 31 (b16_at)
        const u16 *ea = a.Eb + (size_t)(q0 + jb) * NS * a_split + (size_t)lp0 * M;
        const u16 *tb = a.Tb + (size_t)plane * a.nhp * a.lcap + (size_t)lp0 * a.nhp;
        issue_stage<NS>(ea, tb, a_split, t_split, loff, wave, with_b, lds + (it2 % NSTAGE) * STAGE);
    };
    if (nit > 0) issue(0);
    if (nit > 1) issue(1);
    for (int it = 0; it < nit; it++) {
        // this stage's loads landed (each wave its own 2 NS, or NS without
        // table rows), and every wave's fragment reads of the stage refilled
        // next are done
        if (it + 1 >= nit) wait_vm_lgkm0<0>();
        else if (with_b) wait_vm_lgkm0<2 * NS>();
        else wait_vm_lgkm0<NS>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (it + 2 < nit) issue(it + 2);
        const u16 *As = lds + (it % NSTAGE) * STAGE, *Bs = As + NS * BM * BK;
        v8bf af[NS][WI], bfg[NS][2];
#pragma unroll
        for (int s = 0; s < NS; s++) {
#pragma unroll
            for (int i = 0; i < WI; i++) {
                const int row = rbase + i * 16 + fr;
                af[s][i] = *(const v8bf *)&As[(s * BM + row) * BK + 8 * ((lane >> 4) ^ sw_seg(row))];
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int row = cbase + j * 16 + fr;
                bfg[s][j] = *(const v8bf *)&Bs[(s * BN + row) * BK + 8 * ((lane >> 4) ^ sw_seg(row))];
            }
        }
        v4f acc[WI][2];
#pragma unroll
        for (int i = 0; i < WI; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        // cross products of order < NS, smallest first
#pragma unroll
        for (int ord = NS - 1; ord >= 0; ord--)
#pragma unroll
            for (int sa = ord; sa >= 0; sa--) {
                const int sb = ord - sa;
#pragma unroll
                for (int i = 0; i < WI; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[sa][i], bfg[sb][j], acc[i][j], 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < WI; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc64[i][j][q] += (double)acc[i][j][q];
        if (a.mirror && it == nit_e - 1) {
#pragma unroll
            for (int i = 0; i < WI; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        stash[(size_t)((i * 2 + j) * 4 + q) * 64] = acc64[i][j][q];
                        acc64[i][j][q] = 0.0;
                    }
            stashed = true;
        }
    }
    // epilogue: v_mfma_f32_16x16x32_bf16 leaves D[i][j] in lane j + 16*(i/4),
    // register i%4: registers 2h, 2h+1 are re, im of one (receiver, slot)
    // row, written as one complex to out[pos][comp][r][theta], 16 lanes on
    // 16 consecutive distances (the mirror columns likewise, descending)
    void *out = job->out;
    const int ncomp = job->ncomp, accum = job->accumulate, c64 = job->c64;
    const int comp0 = job->comp[0], comp1 = job->comp[1];
    const double sign0 = job->sign[0], sign1 = job->sign[1], msign = job->msign;
    int bad = 0;
#pragma unroll
    for (int i = 0; i < WI; i++)
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++) {
            const int row = m0 + rbase + i * 16 + 4 * (lane >> 4) + 2 * h2;  // even; M is even
            if (row >= M) continue;
            const int pr = row >> 1;
            const int r = pr / a.nslot, sl = pr - r * a.nslot;
            const size_t pos = a.slot_out[sl];
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int col = n0 + cbase + j * 16 + fr;
                if (col >= a.nh) continue;
                const int q0 = 2 * h2, q1 = q0 + 1;
                const double o0 = acc64[i][j][q0], o1 = acc64[i][j][q1];
                const double e0 = stashed ? stash[(size_t)((i * 2 + j) * 4 + q0) * 64] : 0.0;
                const double e1 = stashed ? stash[(size_t)((i * 2 + j) * 4 + q1) * 64] : 0.0;
                const int mc = a.theta_n - 1 - col;
                const bool mir = a.mirror && mc != col;
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int comp = t ? comp1 : comp0;
                    if (comp < 0) continue;
                    const double sg = t ? sign1 : sign0;
                    const size_t od = ((pos * ncomp + comp) * R + r) * a.theta_n;
                    const double2 v = make_double2(sg * (e0 + o0), sg * (e1 + o1));
                    bad += nonfinite2(cat_store(out, od + col, v, accum, c64));
                    if (mir) {
                        const double2 w = make_double2(sg * msign * (e0 - o0), sg * msign * (e1 - o1));
                        bad += nonfinite2(cat_store(out, od + mc, w, accum, c64));
                    }
                }
            }
        }
    add_nonfinite(a.nonfinite, bad);
}

template <int NS, int KIND>  // KIND 0: receiver rows, 1: surface rows (separate symbol for profiling)
__global__ __launch_bounds__(512) void synth_b16_kernel(SynthB16Args a, int nx, int ny)
{
    __shared__ __attribute__((aligned(1024))) u16 lds[NSTAGE * NS * (BM + BN) * BK];
    int tx, ty, tz, tile = blockIdx.x;
    if (a.tiles) {
        // host order: groups of tiles that share operands on one XCD,
        // longest l series first; -1 marks an empty slot
        if ((int)blockIdx.x >= a.ntiles) return;
        const int t = a.tiles[blockIdx.x];
        if (t < 0) return;
        tile = t;
        tx = t % nx;
        ty = (t / nx) % ny;
        tz = t / (nx * ny);
    } else if (!xcd_tile_b16(nx, ny, tx, ty, tz)) {
        return;
    }
    const int m0 = ty * BM, n0 = tx * BN;
    if (a.nh - n0 <= SYNTH_B16_NARROW) synth_b16_tile<NS, true>(a, a.jobs + tz, tile, m0, n0, lds);
    else synth_b16_tile<NS, false>(a, a.jobs + tz, tile, m0, n0, lds);
}

void launch_split_rows(const double2 *E, int nslot, int nq, int r_eff, int lcap, int ns, unsigned short *Eb,
                       const int *nl, hipStream_t s)
{
    hipLaunchKernelGGL(split_rows_kernel, dim3(8192), dim3(256), 0, s, E, nslot, nq, r_eff, lcap, ns, Eb, nl);
}

void launch_zero_tail_b16(unsigned short *Eb, int nq, int ns, int nslot, int r_eff, int lcap, const int *nl, int lend,
                          hipStream_t s)
{
    hipLaunchKernelGGL(zero_tail_b16_kernel, dim3(8192), dim3(256), 0, s, Eb, nq, ns, nslot, r_eff, lcap, nl, lend);
}

void launch_b16_put_l0(const unsigned short *one, unsigned short *Eb, size_t nrows, size_t M, int lcap,
                       hipStream_t s)
{
    hipLaunchKernelGGL(b16_put_l0_kernel, dim3(1024), dim3(256), 0, s, one, Eb, nrows, M, lcap);
}

void launch_zero_pad_b16(unsigned short *Eb, size_t nrows, size_t M, int lcap, int lfrom, hipStream_t s)
{
    if (lfrom >= lcap || nrows == 0) return;
    hipLaunchKernelGGL(zero_pad_b16_kernel, dim3(2048), dim3(256), 0, s, Eb, nrows, M, lcap, lfrom);
}

void launch_split_table(const double *tab, int nplanes, int lcap, int theta_pad, int nh, int nhp, int ns,
                        unsigned short *Tb, hipStream_t s)
{
    // the tiles read distances c0 .. c0 + 63 < nhp of rows of theta_pad >= nh
    // values: columns past nh are not read
    const dim3 g(lcap / ST_L, nhp / ST_C, nplanes);
    hipLaunchKernelGGL(split_table_kernel, g, dim3(ST_NT), 0, s, tab, lcap, theta_pad, nh, nhp, ns, Tb, nplanes);
}

size_t synth_b16_grid(int nslot, int r_eff, int nhp, int njobs)
{
    const int M = nslot * r_eff * 2;
    const int nx = nhp / BN, ny = (M + BM - 1) / BM;
    return (size_t)8 * ((nx * ny + 7) / 8) * njobs;
}

void synth_b16_tiles(int nslot, int r_eff, int nhp, int &nx, int &ny)
{
    nx = nhp / BN;
    ny = (nslot * r_eff * 2 + BM - 1) / BM;
}

void launch_synth_b16(const SynthB16Args &a, int ns, hipStream_t s)
{
    const int M = a.nslot * a.r_eff * 2;
    const int nx = a.nhp / BN, ny = (M + BM - 1) / BM;
    const dim3 grid(a.tiles ? (unsigned)a.ntiles : (unsigned)synth_b16_grid(a.nslot, a.r_eff, a.nhp, a.njobs));
    if (ns == 3) {
        if (a.r_eff > 1) hipLaunchKernelGGL((synth_b16_kernel<3, 0>), grid, dim3(NT), 0, s, a, nx, ny);
        else hipLaunchKernelGGL((synth_b16_kernel<3, 1>), grid, dim3(NT), 0, s, a, nx, ny);
    } else {
        if (a.r_eff > 1) hipLaunchKernelGGL((synth_b16_kernel<2, 0>), grid, dim3(NT), 0, s, a, nx, ny);
        else hipLaunchKernelGGL((synth_b16_kernel<2, 1>), grid, dim3(NT), 0, s, a, nx, ny);
    }
}

}  // namespace dsm
