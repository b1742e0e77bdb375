// psv_kernels.hip -- P-SV (spheroidal) per-(omega,l) kernels: source
// excitation, band solve, coefficient rows of the synthesis contraction and
// the l = 0 system.  Built without FMA contraction (see psv_lane.h).
#include <type_traits>

#include "dev_common.h"
#include "kernels.h"

namespace dsm {

constexpr int SV_NT = 256, SV_CH = 32;  // lanes per solve workgroup, band columns per LDS chunk

// ---------------------------------------------------------------------------
// Source excitation g[s][i] (calya + calg, psv_excitation) of every
// (frequency, l): grid (l blocks of 64, active frequencies).  Two waves per
// SIMD (46 VGPRs spilled): 0.59 ms per 128-frequency batch against 0.74 ms
// at the one wave its full register set allows.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void psv_excite_kernel(PsvSolveArgs a)
{
    const int sl = a.slot_list[blockIdx.y];
    const int l = a.l0 + blockIdx.x * 64 + threadIdx.x;
    if (l > a.lend || l == 0) return;
    double2 g[4][4];
    psv_excitation(l, a.src[sl], a.sc, g);
    double2 *o = a.gexc + (size_t)sl * 16 * a.lcap + l;
    for (int s = 0; s < 4; s++)
        for (int i = 0; i < 4; i++) o[(size_t)(4 * s + i) * a.lcap] = g[s][i];
}

// One wave per (slot, block of 256 l): the lanes take 64 rows at a time
// going down from the source, form their rows' increments of D, and a
// wave prefix sum finds the first row where D passes dcut.
__global__ __launch_bounds__(64) void psv_depth_kernel(PsvDepthArgs a)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= a.nslots * a.nblk) return;
    const int k = t / a.nblk, b = t - k * a.nblk;
    const int sl = a.slot_list[k];
    const double om = a.omega[sl];
    const double al = (double)max(1, a.l0 + 256 * b) + 0.5;
    // walk down from the source row; keep at least three rows below it
    int kst = a.row0;
    double carry = 0.0;
    for (int j0 = a.jtmp - 1; j0 >= a.row0; j0 -= 64) {
        const int j = j0 - lane;  // rows j (1-based), j+1
        double c = 0.0;
        if (j >= a.row0) {
            const double rm = 0.5 * (a.rrow[j - 1] + a.rrow[j]);
            const double vm = 0.5 * (a.vrow[j - 1] + a.vrow[j]);
            const double dr = a.rrow[j] - a.rrow[j - 1];
            const double q = al / rm, p = om / vm;
            const double k2 = q * q - p * p;
            if (k2 > 0.0 && dr > 0.0) c = fmin(sqrt(k2) * dr, a.dcap);
        }
        double ps = c;  // inclusive prefix sum, lane 0 the row nearest the source
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(ps, d, 64);
            if (lane >= d) ps += o;
        }
        const unsigned long long over = __ballot(j >= a.row0 && carry + ps > a.dcut);
        if (over) {
            kst = j0 - (__ffsll((long long)over) - 1);
            break;
        }
        carry += __shfl(ps, 63, 64);
    }
    kst = min(kst, a.jtmp - 3);
    if (lane == 0) a.kstart[t] = max(1, kst - a.row0 + 1);
}

__global__ __launch_bounds__(64) void psv_need_kernel(PsvNeedArgs a)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= a.nblk * a.r_n) return;
    const int b = t / a.r_n, r = t - b * a.r_n;
    const int S = a.nslots, npair = a.r_n * S;
    const int lo = ((r * S) / 64) * 64, hi = min(npair, ((r * S + S - 1) / 64 + 1) * 64) - 1;
    unsigned char f = 0;
    for (int p = lo; p <= hi && !f; p++) {
        const int rr = p / S, sl = p - rr * S;
        if (a.kstart[sl * a.nblk + b] <= a.rlow[rr]) f = 1;
    }
    a.rneed[t] = f;
}

void launch_psv_need(const PsvNeedArgs &a, hipStream_t s)
{
    const int n = a.nblk * a.r_n;
    hipLaunchKernelGGL(psv_need_kernel, dim3((n + 63) / 64), dim3(64), 0, s, a);
}

// Offsets of the per-block factor storage: block b stores
// psv_block_store(n, fs, max(klo, kstart[b]), ckpt) complex per lane of its
// 256; out[0] = total in double2.  One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void psv_loff_kernel(const int *kstart, int nblocks, int n, int klo, int fs,
                                                        int ckpt, long long *loff, long long *total)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x, c = (nblocks + 1023) / 1024;
    const int b0 = min(nblocks, t * c), b1 = min(nblocks, b0 + c);
    long long sum = 0;
    for (int b = b0; b < b1; b++) sum += SV_NT * psv_block_store(n, fs, max(klo, kstart[b]), ckpt != 0);
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan
        const long long v = (t >= d) ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long off = part[t] - sum;
    for (int b = b0; b < b1; b++) {
        loff[b] = off;
        off += SV_NT * psv_block_store(n, fs, max(klo, kstart[b]), ckpt != 0);
    }
    if (t == 1023) total[0] = part[1023];
}

void launch_psv_loff(const int *kstart, int nblocks, int n, int klo, int fs, int ckpt, long long *loff,
                     long long *total, hipStream_t s)
{
    hipLaunchKernelGGL(psv_loff_kernel, dim3(1), dim3(1024), 0, s, kstart, nblocks, n, klo, fs, ckpt, loff, total);
}

// One 64-lane workgroup per slot, 64 degrees at a time.  In the count
// (calamp, src/SGTsh/others.f90:805-824, over m = -2..2, |m| <= l; the
// host form is calamp in engine.cpp) the entries m != 0 have amplitude 0:
// they never raise the running maximum (which starts at -1 and so is
// max(0, amplitudes so far)), their ratio is 0, so they count as small
// exactly when 0 < ratl and l >= lsuf.  The lanes form their degree's
// amplitude, the running maximum through it (wave prefix maximum) and the
// m = 0 test f0.  A degree then maps the count x before its entries to
//   x + 2 side + 1   if f0 (which implies l >= lsuf),
//   side             if l >= lsuf (and 0 < ratl) but not f0,
//   0                otherwise          (side: entries m < 0, 2 or 1 at l = 1),
// maps of the form "set c" or "add k", which compose associatively: a wave
// scan gives the count before every degree, and the cut-off is the first
// degree that starts with a count above 20.
__global__ __launch_bounds__(64) void psv_cut_kernel(PsvCutArgs a)
{
    const int k = blockIdx.x, tx = threadIdx.x;
    const int lsuf = a.lsuf[k];
    const bool zero_small = 0.0 < a.ratl;  // a ratio of 0 counts as small
    int nl = a.maxlmax + 1, llog = a.maxlmax;
    double carry = 0.0;  // running maximum over the degrees of earlier chunks (and the m = -1 entry of l = 1)
    int xin = 0;         // count before the chunk
    for (int l0 = 1; l0 <= a.maxlmax; l0 += 64) {
        const int l = l0 + tx;
        const bool in = l <= a.maxlmax;
        double a0 = 0.0;
        if (in) {
            const double2 u = a.amp[((size_t)k * a.lcap + l) * 2], v = a.amp[((size_t)k * a.lcap + l) * 2 + 1];
            const double a1 = hypot(u.x, u.y), a2 = hypot(v.x, v.y);
            a0 = sqrt(a1 * a1 + a2 * a2);
        }
        double pm = a0;  // inclusive prefix maximum over the chunk
        for (int d = 1; d < 64; d <<= 1) {
            const double o = __shfl_up(pm, d, 64);
            if (tx >= d) pm = fmax(pm, o);
        }
        const double mx = fmax(carry, pm);
        const double ratio = (a0 != 0.0 && mx != 0.0) ? a0 / mx : 0.0;
        const bool f0 = in && ratio < a.ratl && l >= lsuf;
        const bool sm = in && zero_small && l >= lsuf;
        const int side = (l >= 2) ? 2 : 1;
        // this degree's map; lanes past maxlmax: identity (add 0)
        bool set = in && !f0;
        int val = f0 ? 2 * side + 1 : (sm ? side : 0);
        for (int d = 1; d < 64; d <<= 1) {  // inclusive scan of "earlier, then this"
            const bool os = __shfl_up(set, d, 64);
            const int ov = __shfl_up(val, d, 64);
            if (tx >= d) {
                val = set ? val : ov + val;
                set = set || os;
            }
        }
        const int after = set ? val : xin + val;  // count after this degree's entries
        const int before0 = __shfl_up(after, 1, 64);
        const int before = (tx == 0) ? xin : before0;
        const unsigned long long stop = __ballot(in && before > 20);
        if (stop) {
            const int ll = l0 + (__ffsll((long long)stop) - 1);
            llog = min(llog, ll);
            nl = ll;
            break;
        }
        carry = fmax(carry, __shfl(pm, 63, 64));
        xin = __shfl(after, 63, 64);
    }
    if (tx == 0) {
        a.nl[k] = nl;
        a.llog[k] = llog;
    }
}

void launch_psv_cut(const PsvCutArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(psv_cut_kernel, dim3(a.nslots), dim3(64), 0, s, a);
}

// l extent of every (receiver, slot) row of the bf16 synthesis: the l
// cut-off and the leading blocks of 256 degrees whose solved rows reach
// the receiver's stencil
__global__ __launch_bounds__(256) void psv_nlr_kernel(const int *kstart, const int *rlow, const int *nl, int nslots,
                                                      int nblk, int r_n, int *nlr)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nslots * r_n) return;
    const int r = t / nslots, k = t - r * nslots;
    int nb = 0;
    while (nb < nblk && kstart[k * nblk + nb] <= rlow[r]) nb++;
    nlr[t] = min(nl[k], 256 * nb);
}

void launch_psv_nlr(const int *kstart, const int *rlow, const int *nl, int nslots, int nblk, int r_n, int *nlr,
                    hipStream_t s)
{
    hipLaunchKernelGGL(psv_nlr_kernel, dim3((nslots * r_n + 255) / 256), dim3(256), 0, s, kstart, rlow, nl, nslots,
                       nblk, r_n, nlr);
}

void launch_psv_depth(const PsvDepthArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(psv_depth_kernel, dim3(a.nslots * a.nblk), dim3(64), 0, s, a);
}

// ---------------------------------------------------------------------------
// Band solve for l >= 1 (the formulation of psv_solve_lane, same
// operations in the same order), in two kernels: the upward sweep
// (factorization fused with the forward substitution of the four source
// solutions) and the backward substitution of the six solutions.  One lane
// per (frequency, l); lane gid keeps its factors in scratch column gid
// between the two.  F selects the operation family (dev_common.h): false,
// the reference's rounding (the FP64 parity form); true, fused FP64
// arithmetic for the mixed forms.
//
// Upward sweep: a workgroup holds 256 consecutive l of one frequency, so
// every lane walks the same band columns at the same time: the columns are
// staged through LDS in chunks of SV_CH (two buffers, the next chunk's
// global loads in flight during the current one) and read as broadcasts.
// Lanes past the l range shadow a valid l and store nothing, so every lane
// reaches every barrier.
// ---------------------------------------------------------------------------

// Compact band of the fused family: entry ii of column j is A + s B with
// A = e0 (+ e2), B = e1 (+ e3) folded from the full column (two zones'
// terms summed before the scale); one block per (slot, 64 columns).
__global__ __launch_bounds__(256) void psv_compact_kernel(const double2 *band, const int *kind, int nn,
                                                          const int *slot_list, double2 *cband)
{
    const int sl = slot_list[blockIdx.y];
    const int t = blockIdx.x * 256 + threadIdx.x;  // (column, ii)
    if (t >= 4 * nn) return;
    const int j = t >> 2, ii = t & 3;
    const int kb = kind[j];
    const int kk = (kb >> (4 * ii)) & 3, two = (kb >> (4 * ii + 2)) & 1;
    const double2 *e = band + ((size_t)sl * nn + j) * 16 + 4 * ii;
    double2 A = cz(), B = cz();
    if (kk == 3) {
        A = e[0];
    } else if (kk != 0) {
        A = two ? cadd(e[0], e[2]) : e[0];
        B = two ? cadd(e[1], e[3]) : e[1];
    }
    double2 *o = cband + (((size_t)sl * nn + j) * 4 + ii) * 2;
    o[0] = A;
    o[1] = B;
}

void launch_psv_compact(const double2 *band, const int *kind, int nn, const int *slot_list, int nslots,
                        double2 *cband, hipStream_t s)
{
    hipLaunchKernelGGL(psv_compact_kernel, dim3((4 * nn + 255) / 256, nslots), dim3(256), 0, s, band, kind, nn,
                       slot_list, cband);
}

template <bool F, bool CK>  // CK: checkpointed storage (batched mixed forms, F only)
__global__ __launch_bounds__(SV_NT) void psv_factor_kernel(PsvSolveArgs a)
{
    constexpr int CW = F ? 8 : 16;  // double2 per staged column
    __shared__ double2 sband[2][SV_CH * CW];
    __shared__ int skind[2][SV_CH];
    const int tid = threadIdx.x;
    const int sl = a.slot_list[blockIdx.x];
    const int lraw = a.l0 + blockIdx.y * SV_NT + tid;
    const bool act = lraw <= a.lend && lraw != 0;
    const int l = act ? lraw : max(1, a.lend);
    const long gid = ((long)blockIdx.x * gridDim.y + blockIdx.y) * SV_NT + tid;
    const long ss = a.sstride;
    const double l2 = (double)l * (double)(l + 1);
    const double lsq = sqrt(l2);
    const int n = a.nn - a.row0 + 1;
    const int fs = a.jtmp - a.row0 + 1;    // local row of the first source row (>= 3)
    const int klo = a.rowlo - a.row0 + 1;  // lowest local row needed
    const int nsrc = (l >= 2) ? 4 : 3;
    // local column j (1..n) is band row row0+j-1
    const double2 *band0 = (F ? a.cband : a.band) + ((size_t)sl * a.nn + (a.row0 - 1)) * CW;
    const int *kind0 = (F ? a.csel : a.kind) + (a.row0 - 1);
    const double2 *gx = a.gexc + (size_t)sl * 16 * a.lcap + l;
    // first solved local row of this block (depth cut of the batched forms)
    const int blk = blockIdx.x * gridDim.y + blockIdx.y;
    const int kst = a.kstart ? a.kstart[blk] : 1;
    // factor rows k >= kL at Lb[(3 (k - kL) + j) Lst + Lln], forward values
    // at Zs[(4 (k - fs) + s) ss + gid]
    double2 *Lb = a.loff ? a.scratch + a.loff[blk] : a.scratch;
    const int kL = a.loff ? max(klo, kst) : klo;
    const long Lst = a.loff ? SV_NT : ss, Lln = a.loff ? tid : gid;
    double2 *Zs = a.zs;

    const int nch = (n + SV_CH - 1) / SV_CH;
    constexpr int PER = SV_CH * CW / SV_NT;  // staged values per thread
    double2 pf[PER];
    int pk = 0;
    auto fetch = [&](int c) {
        const int base = c * SV_CH;
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + u * SV_NT;
            pf[u] = (base + e / CW < n) ? band0[(size_t)base * CW + e] : cz();
        }
        if (tid < SV_CH) pk = (base + tid < n) ? kind0[base + tid] : 0;
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int u = 0; u < PER; u++) sband[buf][tid + u * SV_NT] = pf[u];
        if (tid < SV_CH) skind[buf][tid] = pk;
    };
    // entries of column j (1..n, visited in increasing order): advances the
    // staging by one chunk on the first column of the next chunk
    int cur = (kst - 1) / SV_CH;
    auto column = [&](int j, double2 e[4]) {
        const int c = (j - 1) / SV_CH;
        if (c != cur) {
            if (cur + 1 < nch) {
                put((cur + 1) & 1);
                if (cur + 2 < nch) fetch(cur + 2);
            }
            __syncthreads();
            cur = c;
        }
        const int o = (j - 1) - c * SV_CH;
        if (F) psv_col_fast(&sband[c & 1][o * CW], skind[c & 1][o], l2, lsq, e);
        else psv_col_exact(&sband[c & 1][o * CW], skind[c & 1][o], l2, lsq, e);
    };

    fetch(cur);
    put(cur & 1);
    if (cur + 1 < nch) fetch(cur + 1);
    __syncthreads();

    PsvWin W;
    {
        double2 e[4];
        column(kst, e);
        W.d0 = e[3];
        column(kst + 1, e);
        W.e01 = e[2];
        W.d1 = e[3];
        column(kst + 2, e);
        W.e02 = e[1];
        W.e12 = e[2];
        W.d2 = e[3];
    }
    W.w4m3 = W.w4m2 = W.w4m1 = W.w3m2 = W.w3m1 = W.w2m1 = cz();
    // checkpointed storage (batched mixed forms): windows of the segments
    // below the source, then the factors of rows xb = max(fs, kL) .. n
    const int nseg = (CK && fs > kL) ? (fs - kL + PSV_CKPT - 1) / PSV_CKPT : 0;
    const int xb = CK ? max(fs, kL) : kL;
    double2 *Lx = CK ? Lb + (size_t)6 * nseg * SV_NT : Lb;
    auto storeL = [&](int k, double2 w2, double2 w3, double2 w4) {
        if (act && k >= klo && k >= xb) {
            Lx[(3 * (k - xb) + 0) * Lst + Lln] = w2;
            Lx[(3 * (k - xb) + 1) * Lst + Lln] = w3;
            Lx[(3 * (k - xb) + 2) * Lst + Lln] = w4;
        }
    };
    auto storeW = [&](int k) {  // window before the step of row k, at the start of a segment
        if (act && ((k - kL) % PSV_CKPT) == 0) {
            double2 *c = Lb + (size_t)6 * ((k - kL) / PSV_CKPT) * SV_NT + Lln;
            c[0 * SV_NT] = W.d0;
            c[1 * SV_NT] = W.e01;
            c[2 * SV_NT] = W.e02;
            c[3 * SV_NT] = W.d1;
            c[4 * SV_NT] = W.e12;
            c[5 * SV_NT] = W.d2;
        }
    };
    // phase 1: rows below the source, factorization only (fs <= n - 3);
    // factors stored from klo up
    for (int k = kst; k < min(fs, klo); k++) {
        double2 e[4], w2, w3, w4;
        column(k + 3, e);
        psv_factor_step<F>(W, k, n, e, crecip_t<F>(W.d0), w2, w3, w4);
    }
    for (int k = max(kst, klo); k < fs; k++) {
        double2 e[4], w2, w3, w4;
        column(k + 3, e);
        if (CK) storeW(k);
        psv_factor_step<F>(W, k, n, e, crecip_t<F>(W.d0), w2, w3, w4);
        if (!CK) storeL(k, w2, w3, w4);
    }
    // the window after the last row below the source, for psv_factor_top_kernel
    // (the factors of the last three rows, which forward() of the first
    // source rows would read, multiply forward values that are zero there)
    {
        const double2 wv[6] = {W.d0, W.e01, W.e02, W.d1, W.e12, W.d2};
#pragma unroll
        for (int c = 0; c < 6; c++) a.wsave[(size_t)c * ss + gid] = wv[c];
    }
}

// Rows fs .. n of the factorization (the source rows and the end of the
// band) with the forward values of the source solutions, from the window
// psv_factor_kernel leaves below the source: a separate kernel so the long
// sweep below runs with the few registers it needs.
template <bool F, bool CK>
__global__ __launch_bounds__(SV_NT) void psv_factor_top_kernel(PsvSolveArgs a)
{
    constexpr int CW = F ? 8 : 16;  // double2 per column
    const int tid = threadIdx.x;
    const int sl = a.slot_list[blockIdx.x];
    const int lraw = a.l0 + blockIdx.y * SV_NT + tid;
    const bool act = lraw <= a.lend && lraw != 0;
    const int l = act ? lraw : max(1, a.lend);
    const long gid = ((long)blockIdx.x * gridDim.y + blockIdx.y) * SV_NT + tid;
    const long ss = a.sstride;
    const double l2 = (double)l * (double)(l + 1);
    const double lsq = sqrt(l2);
    const int n = a.nn - a.row0 + 1;
    const int fs = a.jtmp - a.row0 + 1;    // local row of the first source row (>= 3)
    const int klo = a.rowlo - a.row0 + 1;  // lowest local row needed
    const int nsrc = (l >= 2) ? 4 : 3;
    // local column j (1..n) is band row row0+j-1
    const double2 *band0 = (F ? a.cband : a.band) + ((size_t)sl * a.nn + (a.row0 - 1)) * CW;
    const int *kind0 = (F ? a.csel : a.kind) + (a.row0 - 1);
    const double2 *gx = a.gexc + (size_t)sl * 16 * a.lcap + l;
    // first solved local row of this block (depth cut of the batched forms)
    const int blk = blockIdx.x * gridDim.y + blockIdx.y;
    const int kst = a.kstart ? a.kstart[blk] : 1;
    // factor rows k >= kL at Lb[(3 (k - kL) + j) Lst + Lln], forward values
    // at Zs[(4 (k - fs) + s) ss + gid]
    double2 *Lb = a.loff ? a.scratch + a.loff[blk] : a.scratch;
    const int kL = a.loff ? max(klo, kst) : klo;
    const long Lst = a.loff ? SV_NT : ss, Lln = a.loff ? tid : gid;
    double2 *Zs = a.zs;

    // checkpointed storage (batched mixed forms): windows of the segments
    // below the source, then the factors of rows xb = max(fs, kL) .. n
    const int nseg = (CK && fs > kL) ? (fs - kL + PSV_CKPT - 1) / PSV_CKPT : 0;
    const int xb = CK ? max(fs, kL) : kL;
    double2 *Lx = CK ? Lb + (size_t)6 * nseg * SV_NT : Lb;
    auto storeL = [&](int k, double2 w2, double2 w3, double2 w4) {
        if (act && k >= klo && k >= xb) {
            Lx[(3 * (k - xb) + 0) * Lst + Lln] = w2;
            Lx[(3 * (k - xb) + 1) * Lst + Lln] = w3;
            Lx[(3 * (k - xb) + 2) * Lst + Lln] = w4;
        }
    };
    PsvWin W;
    W.d0 = a.wsave[0 * ss + gid];
    W.e01 = a.wsave[1 * ss + gid];
    W.e02 = a.wsave[2 * ss + gid];
    W.d1 = a.wsave[3 * ss + gid];
    W.e12 = a.wsave[4 * ss + gid];
    W.d2 = a.wsave[5 * ss + gid];
    // factors of the rows below the source: forward() multiplies them by the
    // zero forward values of rows below fs only
    W.w4m3 = W.w4m2 = W.w4m1 = W.w3m2 = W.w3m1 = W.w2m1 = cz();
    auto column = [&](int j, double2 e[4]) {
        if (F) psv_col_fast(band0 + (size_t)(j - 1) * CW, kind0[j - 1], l2, lsq, e);
        else psv_col_exact(band0 + (size_t)(j - 1) * CW, kind0[j - 1], l2, lsq, e);
    };
    // phase 2: forward values of the source solutions at row k (their
    // right-hand sides sit at rows fs..fs+3), then the factorization step
    double2 z1[4], z2[4], z3[4];
    for (int s = 0; s < 4; s++) z1[s] = z2[s] = z3[s] = cz();
    auto forward = [&](int k, double2 piv) {
        const double2 a1 = W.w4m3, a2 = W.w3m2, a3 = W.w2m1;
        const bool src_row = (k - fs) < 4;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            if (s >= nsrc) continue;
            const double2 bk = src_row ? gx[(size_t)(4 * s + (k - fs)) * a.lcap] : cz();
            double2 sum;
            if (k == 3) sum = cmadd_t<F>(cmul_t<F>(a2, z2[s]), a3, z3[s]);
            else sum = cmadd_t<F>(cmadd_t<F>(cmul_t<F>(a1, z1[s]), a2, z2[s]), a3, z3[s]);
            const double2 zk = csub(bk, sum);
            z1[s] = z2[s];
            z2[s] = z3[s];
            z3[s] = zk;
            if (act) Zs[(size_t)(4 * (k - fs) + s) * ss + gid] = F ? cmul_t<F>(zk, piv) : cdiv(zk, W.d0);
        }
    };
    for (int k = fs; k <= n - 3; k++) {
        double2 e[4], w2, w3, w4;
        column(k + 3, e);
        const double2 piv = crecip_t<F>(W.d0);
        forward(k, piv);
        psv_factor_step<F>(W, k, n, e, piv, w2, w3, w4);
        storeL(k, w2, w3, w4);
    }
    // rows n-2, n-1 (end-of-band updates) and n (forward value only)
    double2 Dn1, Dn, w2n1;
    {
        double2 w2, w3, w4;
        double2 piv = crecip_t<F>(W.d0);
        forward(n - 2, piv);
        psv_factor_step<F>(W, n - 2, n, nullptr, piv, w2, w3, w4);
        storeL(n - 2, w2, w3, w4);
        piv = crecip_t<F>(W.d0);
        forward(n - 1, piv);
        Dn1 = W.d0;
        psv_factor_step<F>(W, n - 1, n, nullptr, piv, w2, w3, w4);
        storeL(n - 1, w2, w3, w4);
        w2n1 = w2;
        piv = crecip_t<F>(W.d0);
        forward(n, piv);
        Dn = W.d0;
    }
    if (act) {
        a.ends[gid] = Dn1;
        a.ends[ss + gid] = Dn;
        a.ends[2 * ss + gid] = w2n1;
    }
}

// Backward substitution from row n down to klo for the six solutions,
// recording them at the rows in need[]: the four source solutions from
// their forward values, the back-propagated forces YB (row n-1, value
// -float(c0)) and S (row n) from their forward values at rows n-1, n
// only.  One lane per (frequency, l), 64-lane workgroups.  The rows from
// the source up (forward values present) come first; below the source the
// factors of four rows ahead are in flight in a ring of four registers
// sets that is never copied (a copy would wait for its load).
template <bool F>
__global__ __launch_bounds__(64) void psv_back_kernel(PsvSolveArgs a)
{
    const int sub = blockIdx.y & 3, xb = blockIdx.y >> 2;
    const int sl = a.slot_list[blockIdx.x];
    const int lraw = a.l0 + xb * SV_NT + sub * 64 + threadIdx.x;
    if (lraw > a.lend || lraw == 0) return;
    const int l = lraw;
    const long gid = ((long)blockIdx.x * (gridDim.y >> 2) + xb) * SV_NT + sub * 64 + threadIdx.x;
    const long ss = a.sstride;
    const double l2 = (double)l * (double)(l + 1);
    const double lsq = sqrt(l2);
    const int n = a.nn - a.row0 + 1;
    const int fs = a.jtmp - a.row0 + 1;
    const int klo = a.rowlo - a.row0 + 1;
    const int nsrc = (l >= 2) ? 4 : 3;
    const double2 *Zs = a.zs;
    const double2 Dn1 = a.ends[gid], Dn = a.ends[ss + gid], w2n1 = a.ends[2 * ss + gid];
    const int blk = blockIdx.x * (gridDim.y >> 2) + xb;
    const int kst = a.kstart ? a.kstart[blk] : 1;
    const int kbot = max(klo, kst);  // rows below kst are outside the solved system
    const double2 *Lb = a.loff ? a.scratch + a.loff[blk] : a.scratch;
    const int kL = a.loff ? kbot : klo;
    const long Lst = a.loff ? SV_NT : ss, Lln = a.loff ? sub * 64 + threadIdx.x : gid;

    // back-propagated forces: forward values at rows n-1, n only
    const double c0f = (double)(float)sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / 1.0);
    const double cm1 = sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / (1.0 * (double)l * (double)(l + 1)));
    const double hs = l2 * cm1 / 2.0;
    const double2 yb_rhs = make_double2(-c0f, 0.0);
    const double2 s_rhs = cdiv(make_double2(hs, 0.0), make_double2(lsq, 0.0));
    const double2 zb_n1 = cdiv(yb_rhs, Dn1);
    // z(n) = 0 - (a(1,n) 0 + a(2,n) 0 + a(3,n) z(n-1)); the zero products are exact
    const double2 zb_n = cdiv(csub(cz(), cadd(cadd(cz(), cz()), cmul(w2n1, yb_rhs))), Dn);
    const double2 zs_n = cdiv(s_rhs, Dn);

    double2 x1[PSV_NSOL], x2[PSV_NSOL], x3[PSV_NSOL];  // b(k+1), b(k+2), b(k+3)
    for (int s = 0; s < PSV_NSOL; s++) x1[s] = x2[s] = x3[s] = cz();
    int p = 0;
    double2 *xo = a.xn + (size_t)sl * a.nneed * PSV_NSOL * a.lcap + l;
    struct Row {
        double2 L2, L3, L4;
    };
    auto ld = [&](int k) {
        Row r{cz(), cz(), cz()};
        if (k >= kbot) {
            r.L2 = (k < n) ? Lb[(3 * (k - kL) + 0) * Lst + Lln] : cz();
            r.L3 = (k < n - 1) ? Lb[(3 * (k - kL) + 1) * Lst + Lln] : cz();
            r.L4 = (k < n - 2) ? Lb[(3 * (k - kL) + 2) * Lst + Lln] : cz();
        }
        return r;
    };
    auto record = [&](int k, const double2 *bk) {
        const int row = a.row0 + k - 1;
        if (p < a.nneed && a.need[p] == row) {
            for (int s = 0; s < PSV_NSOL; s++) xo[(size_t)(PSV_NSOL * p + s) * a.lcap] = bk[s];
            p++;
        }
    };
    auto shift = [&](const double2 *bk) {
#pragma unroll
        for (int s = 0; s < PSV_NSOL; s++) {
            x3[s] = x2[s];
            x2[s] = x1[s];
            x1[s] = bk[s];
        }
    };
    // rows n .. fs: forward values present
    for (int k = n; k >= max(fs, kbot); k--) {
        const Row q = ld(k);
        double2 zd[PSV_NSOL];
#pragma unroll
        for (int s = 0; s < 4; s++) zd[s] = (s < nsrc) ? Zs[(size_t)(4 * (k - fs) + s) * ss + gid] : cz();
        zd[4] = (k == n) ? zb_n : (k == n - 1) ? zb_n1 : cz();
        zd[5] = (k == n) ? zs_n : cz();
        double2 bk[PSV_NSOL];
        if (k == n) {
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++) bk[s] = zd[s];
        } else if (F) {
            const double2 m2 = cneg(q.L2), m3 = cneg(q.L3), m4 = cneg(q.L4);
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++)
                bk[s] = cmadd_t<F>(cmadd_t<F>(cmadd_t<F>(zd[s], m4, x3[s]), m3, x2[s]), m2, x1[s]);
        } else if (k == n - 1) {
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++) bk[s] = csub(zd[s], cmul(q.L2, x1[s]));
        } else if (k == n - 2) {
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++) bk[s] = csub(zd[s], cadd(cmul(q.L3, x2[s]), cmul(q.L2, x1[s])));
        } else {
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++)
                bk[s] = csub(zd[s], cadd(cadd(cmul(q.L4, x3[s]), cmul(q.L3, x2[s])), cmul(q.L2, x1[s])));
        }
        shift(bk);
        record(k, bk);
    }
    // rows fs-1 .. kbot: no right-hand side (fs - 1 <= n - 4)
    auto step = [&](int k, const Row &q) {
        double2 bk[PSV_NSOL];
        if (F) {
            const double2 m2 = cneg(q.L2), m3 = cneg(q.L3), m4 = cneg(q.L4);
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++)
                bk[s] = cmadd_t<F>(cmadd_t<F>(cmul_t<F>(m4, x3[s]), m3, x2[s]), m2, x1[s]);
        } else {
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++)
                bk[s] = csub(cz(), cadd(cadd(cmul(q.L4, x3[s]), cmul(q.L3, x2[s])), cmul(q.L2, x1[s])));
        }
        shift(bk);
        record(k, bk);
    };
    int k = fs - 1;
    Row r0 = ld(k), r1 = ld(k - 1), r2 = ld(k - 2), r3 = ld(k - 3);
    for (; k - 3 >= kbot; k -= 4) {
        step(k, r0);
        r0 = ld(k - 4);
        step(k - 1, r1);
        r1 = ld(k - 5);
        step(k - 2, r2);
        r2 = ld(k - 6);
        step(k - 3, r3);
        r3 = ld(k - 7);
    }
    if (k >= kbot) step(k, r0);
    if (k - 1 >= kbot) step(k - 1, r1);
    if (k - 2 >= kbot) step(k - 2, r2);
    // recorded rows below the cut are not written: the coefficient kernel
    // reads them as zero (PsvCoefArgs::kstart)
}

// Backward substitution of the checkpointed storage (batched mixed forms):
// rows n .. fs from the stored factors as in psv_back_kernel; below the
// source, segment by segment from the top, the factors of a segment's
// PSV_CKPT rows recomputed from its window (psv_factor_step on the compact
// band, the factor kernel's operations on the same values, so the same
// bits) into registers, then the substitution down through them.
__global__ __launch_bounds__(64) void psv_back_ckpt_kernel(PsvSolveArgs a)
{
    constexpr bool F = true;
    const int sub = blockIdx.y & 3, xb_ = blockIdx.y >> 2;
    const int sl = a.slot_list[blockIdx.x];
    const int lraw = a.l0 + xb_ * SV_NT + sub * 64 + threadIdx.x;
    if (lraw > a.lend || lraw == 0) return;
    const int l = lraw;
    const long gid = ((long)blockIdx.x * (gridDim.y >> 2) + xb_) * SV_NT + sub * 64 + threadIdx.x;
    const long ss = a.sstride;
    const double l2 = (double)l * (double)(l + 1);
    const double lsq = sqrt(l2);
    const int n = a.nn - a.row0 + 1;
    const int fs = a.jtmp - a.row0 + 1;
    const int klo = a.rowlo - a.row0 + 1;
    const int nsrc = (l >= 2) ? 4 : 3;
    const double2 *Zs = a.zs;
    const double2 Dn1 = a.ends[gid], Dn = a.ends[ss + gid], w2n1 = a.ends[2 * ss + gid];
    const int blk = blockIdx.x * (gridDim.y >> 2) + xb_;
    const int kL = max(klo, a.kstart[blk]);  // rows below kL are outside the solved system or not needed
    const int nseg = fs > kL ? (fs - kL + PSV_CKPT - 1) / PSV_CKPT : 0;
    const int xb = max(fs, kL);
    const int lane = sub * 64 + threadIdx.x;
    const double2 *Ck = a.scratch + a.loff[blk];
    const double2 *Lx = Ck + (size_t)6 * nseg * SV_NT;
    const double2 *cb0 = a.cband + ((size_t)sl * a.nn + (a.row0 - 1)) * 8;
    const int *cs0 = a.csel + (a.row0 - 1);

    const double c0f = (double)(float)sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / 1.0);
    const double cm1 = sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / (1.0 * (double)l * (double)(l + 1)));
    const double hs = l2 * cm1 / 2.0;
    const double2 yb_rhs = make_double2(-c0f, 0.0);
    const double2 s_rhs = cdiv(make_double2(hs, 0.0), make_double2(lsq, 0.0));
    const double2 zb_n1 = cdiv(yb_rhs, Dn1);
    const double2 zb_n = cdiv(csub(cz(), cadd(cadd(cz(), cz()), cmul(w2n1, yb_rhs))), Dn);
    const double2 zs_n = cdiv(s_rhs, Dn);

    double2 x1[PSV_NSOL], x2[PSV_NSOL], x3[PSV_NSOL];  // b(k+1), b(k+2), b(k+3)
    for (int s = 0; s < PSV_NSOL; s++) x1[s] = x2[s] = x3[s] = cz();
    int p = 0;
    double2 *xo = a.xn + (size_t)sl * a.nneed * PSV_NSOL * a.lcap + l;
    auto record = [&](int k, const double2 *bk) {
        const int row = a.row0 + k - 1;
        if (p < a.nneed && a.need[p] == row) {
            for (int s = 0; s < PSV_NSOL; s++) xo[(size_t)(PSV_NSOL * p + s) * a.lcap] = bk[s];
            p++;
        }
    };
    auto shift = [&](const double2 *bk) {
#pragma unroll
        for (int s = 0; s < PSV_NSOL; s++) {
            x3[s] = x2[s];
            x2[s] = x1[s];
            x1[s] = bk[s];
        }
    };
    // rows n .. xb: forward values present, factors stored
    for (int k = n; k >= xb; k--) {
        double2 L2 = cz(), L3 = cz(), L4 = cz();
        if (k < n) L2 = Lx[(3 * (k - xb) + 0) * SV_NT + lane];
        if (k < n - 1) L3 = Lx[(3 * (k - xb) + 1) * SV_NT + lane];
        if (k < n - 2) L4 = Lx[(3 * (k - xb) + 2) * SV_NT + lane];
        double2 zd[PSV_NSOL];
#pragma unroll
        for (int s = 0; s < 4; s++) zd[s] = (s < nsrc && k >= fs) ? Zs[(size_t)(4 * (k - fs) + s) * ss + gid] : cz();
        zd[4] = (k == n) ? zb_n : (k == n - 1) ? zb_n1 : cz();
        zd[5] = (k == n) ? zs_n : cz();
        double2 bk[PSV_NSOL];
        if (k == n) {
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++) bk[s] = zd[s];
        } else {
            const double2 m2 = cneg(L2), m3 = cneg(L3), m4 = cneg(L4);
#pragma unroll
            for (int s = 0; s < PSV_NSOL; s++)
                bk[s] = cmadd_t<F>(cmadd_t<F>(cmadd_t<F>(zd[s], m4, x3[s]), m3, x2[s]), m2, x1[s]);
        }
        shift(bk);
        record(k, bk);
    }
    // below the source: segments from the top
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int k0 = kL + sg * PSV_CKPT;
        const double2 *c = Ck + (size_t)6 * sg * SV_NT + lane;
        PsvWin W;
        W.d0 = c[0 * SV_NT];
        W.e01 = c[1 * SV_NT];
        W.e02 = c[2 * SV_NT];
        W.d1 = c[3 * SV_NT];
        W.e12 = c[4 * SV_NT];
        W.d2 = c[5 * SV_NT];
        double2 w2[PSV_CKPT], w3[PSV_CKPT], w4[PSV_CKPT];
#pragma unroll
        for (int u = 0; u < PSV_CKPT; u++) {
            const int k = k0 + u;
            if (k < fs) {
                double2 e[4];
                psv_col_fast(cb0 + (size_t)(k + 2) * 8, cs0[k + 2], l2, lsq, e);  // column k + 3
                psv_factor_step<F>(W, k, n, e, crecip_t<F>(W.d0), w2[u], w3[u], w4[u]);
            }
        }
#pragma unroll
        for (int u = PSV_CKPT - 1; u >= 0; u--) {
            const int k = k0 + u;
            if (k < fs) {
                const double2 m2 = cneg(w2[u]), m3 = cneg(w3[u]), m4 = cneg(w4[u]);
                double2 bk[PSV_NSOL];
#pragma unroll
                for (int s = 0; s < PSV_NSOL; s++)
                    bk[s] = cmadd_t<F>(cmadd_t<F>(cmul_t<F>(m4, x3[s]), m3, x2[s]), m2, x1[s]);
                shift(bk);
                record(k, bk);
            }
        }
    }
    // recorded rows below the cut are not written: the coefficient kernel
    // reads them as zero (PsvCoefArgs::kstart)
}

void launch_psv_excite(const PsvSolveArgs &a, int nslots, hipStream_t s)
{
    const int nb = (a.lend - a.l0 + 64) / 64;
    hipLaunchKernelGGL(psv_excite_kernel, dim3(nb, nslots), dim3(64), 0, s, a);
}

int psv_solve_blocks(int l0, int lend) { return (lend - l0 + SV_NT) / SV_NT; }

void launch_psv_solve(const PsvSolveArgs &a, int nslots, bool fast, hipStream_t s)
{
    const int nb = psv_solve_blocks(a.l0, a.lend);
    // slots fastest: every slot's lowest (longest, full-depth) block of l is
    // dispatched first
    const dim3 g1(nslots, nb), g2(nslots, 4 * nb);
    if (fast) {
        if (a.ckpt && a.loff && a.kstart) {
            hipLaunchKernelGGL((psv_factor_kernel<true, true>), g1, dim3(SV_NT), 0, s, a);
            hipLaunchKernelGGL((psv_factor_top_kernel<true, true>), g1, dim3(SV_NT), 0, s, a);
            hipLaunchKernelGGL(psv_back_ckpt_kernel, g2, dim3(64), 0, s, a);
        } else {
            hipLaunchKernelGGL((psv_factor_kernel<true, false>), g1, dim3(SV_NT), 0, s, a);
            hipLaunchKernelGGL((psv_factor_top_kernel<true, false>), g1, dim3(SV_NT), 0, s, a);
            hipLaunchKernelGGL(psv_back_kernel<true>, g2, dim3(64), 0, s, a);
        }
    } else {
        hipLaunchKernelGGL((psv_factor_kernel<false, false>), g1, dim3(SV_NT), 0, s, a);
        hipLaunchKernelGGL((psv_factor_top_kernel<false, false>), g1, dim3(SV_NT), 0, s, a);
        hipLaunchKernelGGL(psv_back_kernel<false>, g2, dim3(64), 0, s, a);
    }
}

// Coefficient rows of the bf16 forms, one 64-lane wave per block of 64
// degrees (l0 a multiple of 64).  A row's NS terms x (re, im) are staged in
// LDS, the wave's 32 even and 32 odd degrees being contiguous in the two
// halves of the permuted row, and written by 16 NS lanes as 16-byte
// chunks: one store instruction per row where each lane would issue 2 NS
// 2-byte stores.  Lanes past lend or at l = 0 stage zeros (l = 0 is
// written afterwards by psv_l0_kernel; degrees above lmax are padding).
namespace {
// lane part of a staged row's store address: term s, re|im, half (even or
// odd degrees) and 16-byte chunk of lane t (t < 16 NS), for receiver r
__device__ __forceinline__ unsigned short *wave_row_base(const B16Rows &d, int r, int slot, int lw0, int t)
{
    const int sg = (t >> 2) % (4 * d.ns), c = t & 3;
    const int pl = sg >> 1, hh = sg & 1, s = pl >> 1, reim = pl & 1;
    const size_t M = (size_t)d.nslot * d.r_n * 2, m = ((size_t)r * d.nslot + slot) * 2 + reim;
    return d.Eb + b16_at(s, M, m, (hh ? (d.lcap >> 1) : 0) + (lw0 >> 1) + 8 * c, d.lcap);
}
// stage row q's terms of lane t's value, then the wave writes the row
// (lanes t < 16 NS, 16 bytes each at base + q NS M lcap)
__device__ __forceinline__ void wave_put(const B16Rows &d, unsigned short (*st)[32], unsigned short *base, size_t qstride,
                                         int q, double2 v, int t)
{
    unsigned short tr[3], ti[3];
    bf16_terms(v.x, d.ns, tr);
    bf16_terms(v.y, d.ns, ti);
    const int h = t & 1, pos = t >> 1;
    for (int s = 0; s < d.ns; s++) {
        st[(s * 2 + 0) * 2 + h][pos] = tr[s];
        st[(s * 2 + 1) * 2 + h][pos] = ti[s];
    }
    if (t < 16 * d.ns) {
        const uint4 val = *(const uint4 *)&st[t >> 2][8 * (t & 3)];
        *(uint4 *)(base + (size_t)q * qstride) = val;
    }
}
}  // namespace

__global__ __launch_bounds__(64) void psv_coef_b16_kernel(PsvCoefArgs a)
{
    __shared__ __attribute__((aligned(16))) unsigned short st[12][32];
    const int nsl = gridDim.x / a.nlb;
    const int b = (blockIdx.x % nsl) * a.nlb + blockIdx.x / nsl;  // l-major, as psv_coef_kernel
    const int t = threadIdx.x;
    const int sl = a.slot_list[b / a.nlb];
    const int lw0 = a.l0 + (b % a.nlb) * 64;
    const int lraw = lw0 + t;
    const bool live = lraw <= a.lend && lraw != 0;
    const int l = live ? lraw : max(1, min(lraw, a.lend));  // dead lanes read a valid degree, stage zeros
    const int R = a.r_n, LC = a.lcap;
    const double il = 1.0 / sqrt((double)l * (double)(l + 1));
    const bool mu1 = true, mu2 = (l >= 2);
    const unsigned char *rneed = a.rneed ? a.rneed + (size_t)((lw0 - a.l0) >> 8) * R : nullptr;
    const double2 *X = a.xn + (size_t)sl * a.nneed * PSV_NSOL * LC + l;
    auto xv = [&](int p, int s) { return X[(size_t)(PSV_NSOL * p + s) * LC]; };
    if (live) {  // amplitude rows of the top receiver (calamp on imt 1, m = 0)
        a.ampuv[((size_t)sl * LC + l) * 2 + 0] = xv(a.amp_i0, 0);
        a.ampuv[((size_t)sl * LC + l) * 2 + 1] = xv(a.amp_i1, 0);
    }
    {
        double2 us[PSV_NSOL], vs[PSV_NSOL], es[PSV_NQS];
        for (int s = 0; s < PSV_NSOL; s++) {
            us[s] = xv(a.syn_i0, s);
            vs[s] = xv(a.syn_i1, s);
        }
        psv_synn_rows(us, vs, il, mu1, mu2, es);
        if (a.Es) {  // SYNN rows in FP64 (degree 0 is psv_l0_kernel's)
            double2 *Es = a.Es + (size_t)sl * PSV_NQS * LC + lraw;
            if (lraw != 0)
                for (int q = 0; q < PSV_NQS; q++) Es[(size_t)q * LC] = live ? es[q] : cz();
        } else {
            unsigned short *base = wave_row_base(a.esb, 0, sl, lw0, t);
            const size_t qs = (size_t)a.esb.ns * a.esb.nslot * 2 * a.esb.lcap;
            for (int q = 0; q < PSV_NQS; q++) wave_put(a.esb, st, base, qs, q, live ? es[q] : cz(), t);
        }
    }
    const size_t qstride = (size_t)a.eb.ns * a.eb.nslot * a.eb.r_n * 2 * a.eb.lcap;
    const int kst = a.kstart ? a.kstart[b / a.nlb * a.nblk + ((lw0 - a.l0) >> 8)] : 1;
    const int nq = a.nq ? a.nq : PSV_NQ;  // (rows past nq are read by no job)
    for (int r = 0; r < R; r++) {
        if (rneed && !rneed[r]) continue;  // below the solved rows of every slot the tile holds
        unsigned short *base = wave_row_base(a.eb, r, sl, lw0, t);
        if (a.kstart && kst > a.rlow[r]) {  // below this block's solved rows: zero rows
            for (int q = 0; q < nq; q++) wave_put(a.eb, st, base, qstride, q, cz(), t);
            continue;
        }
        const int *rr = a.rec_row + 6 * r;
        const double *w0 = a.w0 + 3 * r, *w1 = a.w1 + 3 * r;
        const bool fl = a.fluid[r] != 0;
        const double2 K = a.fluidK[(size_t)sl * R + r];
        const double irf = a.irf[r];
        auto put = [&](int q, double2 v) {
            if (q < nq) wave_put(a.eb, st, base, qstride, q, live ? v : cz(), t);  // (q is uniform)
        };
        // one solution at a time (psv_rows grouped by solution): its
        // interpolated values, then its rows; the stencil values of the next
        // solution are loaded before the rows of this one
        auto ld = [&](int S, double2 *g) {
            for (int k = 0; k < 6; k++) g[k] = xv(rr[k], S);
        };
        auto one = [&](auto sc, const double2 *g) {
            constexpr int S = decltype(sc)::value;
            PsvVals z;
            z.U = cadd(cadd(cscale(g[0], w0[0]), cscale(g[2], w0[1])), cscale(g[4], w0[2]));
            z.V = cadd(cadd(cscale(g[1], w0[0]), cscale(g[3], w0[1])), cscale(g[5], w0[2]));
            z.Ud = cadd(cadd(cscale(g[0], w1[0]), cscale(g[2], w1[1])), cscale(g[4], w1[2]));
            z.Vd = cadd(cadd(cscale(g[1], w1[0]), cscale(g[3], w1[1])), cscale(g[5], w1[2]));
            psv_rows_one<S>(z, fl, K, il, irf, mu1, mu2, put);
        };
        double2 ga[6], gb[6];
        ld(0, ga);
        ld(1, gb);
        one(std::integral_constant<int, 0>{}, ga);
        ld(2, ga);
        one(std::integral_constant<int, 1>{}, gb);
        ld(3, gb);
        one(std::integral_constant<int, 2>{}, ga);
        ld(4, ga);
        one(std::integral_constant<int, 3>{}, gb);
        ld(5, gb);
        one(std::integral_constant<int, 4>{}, ga);
        one(std::integral_constant<int, 5>{}, gb);
    }
}

// blocks of 64 l in l-major order (the low l, with every receiver's rows, first)
__global__ __launch_bounds__(64) void psv_coef_kernel(PsvCoefArgs a)
{
    const int ns = gridDim.x / a.nlb;
    const int b = (blockIdx.x % ns) * a.nlb + blockIdx.x / ns;
    psv_coef_lane(a, b, threadIdx.x);
}

void launch_psv_coef(const PsvCoefArgs &a, int nblocks, hipStream_t s)
{
    if (a.eb.Eb && (a.esb.Eb || a.Es) && (a.l0 % 64) == 0 && a.eb.ns <= 3)
        hipLaunchKernelGGL(psv_coef_b16_kernel, dim3(nblocks), dim3(64), 0, s, a);
    else
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
