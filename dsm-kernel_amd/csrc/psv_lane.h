// psv_lane.h -- per-lane bodies of the P-SV (omega,l) kernels, shared by the
// HIP kernels (psv_kernels.hip) and the host build used to test the
// formulation on the CPU (hostcheck.cpp).  Compiled without FMA
// contraction: the band factorization, the source problem and the
// substitutions keep the reference's rounding (see sh_lane.h).
#pragma once

#include "dev_common.h"

namespace dsm {

#ifndef DSM_PI_DEFINED
#define DSM_PI_DEFINED
static constexpr double DSM_PI = 3.1415926535897932;
#endif

// The six distinct right-hand sides of one (omega,l) system.  The
// reference solves 15 (SGTpsv-MPI.f90:867-1080); the others are exact
// multiples of these (the source vector of calg is linear in the moment
// tensor, and the band solve commutes exactly with a factor -1, +-i or 2):
//   0 Y1: m=0, imt=1            1 Y2: m=0, imt=2 (= imt=3)
//   2 X1: m=-1, imt=4  (m=+1 imt4: -X1; imt5 both m: i X1)
//   3 X2: m=+2, imt=2  (m=-2 imt2: X2; imt3: -X2; imt6: -2i sgn(m) X2)
//   4 YB: back-propagated surface force icomp=1 (row nn-1)
//   5 S : back-propagated surface force at row nn with value
//         l(l+1)c/2 / sqrt(l(l+1)) (icomp=2: +-S, icomp=3: -iS)
constexpr int PSV_NSOL = 6;
constexpr int PSV_NQ = 63;   // coefficient rows per receiver (TSGT 42 + RSGT 21)
constexpr int PSV_NQT = 42;  // the TSGT rows, first
constexpr int PSV_NQS = 10;  // SYNN rows

// per-slot excitation data: calabnum tables and source-zone moduli
struct PsvSrc {
    double2 anum[160], bnum[160];
    double2 coef1, coef2;  // coef1(spn), coef2(spn)
};
struct PsvSrcConst {
    double ecC0, ecF0, ecL0, r0, ra1, ra2;
};

struct PsvSolveArgs {
    const double2 *band;   // [slot][nn][4 band rows][4]: P0, P1, Q0, Q1
    const int *kind;       // [nn]: 4 bits per band row: kind (2 bits) | second term (bit 2)
    const int *slot_list;
    int nlb, l0, lend;
    int nn, row0, jtmp;    // band rows, first solved row, first source row (1-based)
    int rowlo;             // lowest row the backward pass must produce
    int nneed;
    const int *need;       // rows, descending, 1-based
    double2 *scratch;      // L: [n_l rows][3][nthreads], then zd: [n_z rows][4][nthreads]
    long sstride;
    double2 *xn;           // [slot][nneed][PSV_NSOL][lcap]
    int lcap;
    const PsvSrc *src;     // [slot]
    PsvSrcConst sc;
    double2 *gexc;         // [slot][4 sources][4 rows][lcap]: excitation (psv_excite_kernel)
    double2 *ends;         // [3][sstride]: d(n-1), d(n), w2(n-1) of the factorization
    const int *kstart;     // [active slot][l block of 256]: first solved local row (null: 1)
    const double2 *cband;  // fused family: [slot][nn][4][A, B] (psv_compact_kernel)
    const int *csel;       // [nn]: bit ii set where entry ii scales with sqrt(l(l+1))
    double2 *zs;           // forward values: [k - fs][4][sstride]
    double2 *wsave;        // [6][sstride]: the factorization window below the source (d0 .. d2 of PsvWin)
    // batched forms: factors of (active slot, l block) blk from scratch + loff[blk],
    // [k - max(klo, kstart)][3][256]; null: [k - klo][3][sstride]
    const long long *loff;
    // batched mixed forms: below the source row the factorization's window
    // every PSV_CKPT rows instead of the factors (psv_block_store); the
    // backward pass recomputes a segment's factors from its checkpoint
    int ckpt;
};

constexpr int PSV_CKPT = 8;

// Factor storage of one block of the batched forms, double2 per lane, the
// lowest stored row kL = max(klo, kstart): explicit, the factors (3 per
// row) of rows kL .. n; checkpointed, the windows (6) of the segments of
// PSV_CKPT rows from kL below the source row fs, then the factors of rows
// max(fs, kL) .. n.
__host__ __device__ inline long long psv_block_store(int n, int fs, int kL, bool ckpt)
{
    if (!ckpt) return 3LL * (n - kL + 1);
    const int nseg = fs > kL ? (fs - kL + PSV_CKPT - 1) / PSV_CKPT : 0;
    return 6LL * nseg + 3LL * (n - (fs > kL ? fs : kL) + 1);
}

// Depth cut of the batched forms: for the smallest l of every block of 256
// degrees, the deepest band row whose evanescent decay below the source,
//   D = sum over rows of min(dcap, sqrt(max(0, ((l+1/2)/r)^2 - (omega/v)^2)) dr),
// stays within dcut (calmdr's local quantity, others.f90:904-927, summed
// row by row; v the slowest wave speed of the row).  The cap per row
// follows the discrete system rather than the continuum: where the grid
// does not resolve the decay length (kappa h >> 1) a 3-point element
// operator decays by at most ~e^-1.3 per node whatever kappa is.  Rows
// below are dropped from the system: their share of the solution is below
// e^-dcut.
// Receivers whose coefficient rows some synthesis tile reads in each
// block of 256 degrees: a tile holds the rows (r', slot) of 64 consecutive
// receiver-major pairs, and reads a block if for one of them the block's
// solved rows reach the receiver's stencil (kstart <= rlow(r')).
struct PsvNeedArgs {
    const int *kstart;     // [slot][nblk]
    const int *rlow;       // [r]: local row of ksta - 1
    int nslots, nblk, r_n;
    unsigned char *rneed;  // [nblk][r_n]
};

// l cut-off of every slot (calamp replay, SGTpsv-MPI.f90:616-634,
// 884-888): from the amplitudes of the top receiver's rows, stop after
// more than 20 consecutive degrees whose amplitude ratio is below ratl
// past lsuf; m = +-1, +-2 contribute zero amplitudes.
struct PsvCutArgs {
    const double2 *amp;  // [slot][lcap][2]
    const int *lsuf;     // [slot]
    int nslots, lcap, maxlmax;
    double ratl;
    int *nl, *llog;      // [slot]: number of degrees, the reference's llog
};

struct PsvDepthArgs {
    const double *rrow, *vrow;  // [nn] (psv_row_profile)
    const double *omega;        // [slot] real angular frequency
    const int *slot_list;
    int nslots, nblk, l0, row0, jtmp;
    double dcut, dcap;
    int *kstart;                // [active slot][nblk] local rows
};

struct PsvCoefArgs {
    const double2 *xn;
    const int *slot_list;
    int nlb, l0, lend, lcap, nneed, r_n;
    const int *rec_row;    // [r][6] -> index into need of rows ksta-1 .. ksta+4
    const int *fluid;      // [r]
    const double *w0, *w1; // [r][3]
    const double *irf;     // [r]: 1 / float(r)
    const double2 *fluidK; // [slot][r]
    int amp_i0, amp_i1;    // need indices of rows ksta(r_n)-1, ksta(r_n)
    int syn_i0, syn_i1;    // need indices of rows nn-1, nn
    double2 *E;            // [slot][PSV_NQ][r_n][lcap]
    double2 *Es;           // [slot][PSV_NQS][lcap]
    double2 *ampuv;        // [slot][lcap][2]
    B16Rows eb, esb;       // mixed forms (eb.Eb non-null): bf16 rows instead of E, Es
    const unsigned char *rneed;  // batched forms: [l block of 256][r] rows some synthesis tile reads (null: all)
    // batched forms: first solved local row per (slot, l block) and the
    // receivers' lowest stencil row (local); a receiver below its block's
    // first row has zero solutions there (not stored by the back substitution)
    const int *kstart, *rlow;
    int nblk;
    int nq;                // rows written: PSV_NQ, or PSV_NQT without RSGT (0: PSV_NQ)
};

struct PsvL0Args {
    const double2 *band;
    const int *kind;
    const int *slot_list;
    int nslots;
    int nn, nn0, row0_0, jtmp;
    int js;                // node index of the source row jtmp (js, js+1: the source's U rows)
    const int *node_row;   // [nn0]
    const int *node_c1;    // [nn0]
    double2 *work;         // [slot][2*nn0 + 3*nn0]
    int r_n;
    const int *jsta0;      // [r]
    const int *fluid;
    const double *w0, *w1, *irf;
    const double2 *fluidK;
    const PsvSrc *src;
    PsvSrcConst sc;
    double2 *E, *Es;
    int lcap;
    B16Rows eb, esb;       // mixed forms: bf16 rows (see PsvCoefArgs)
};

// ---------------------------------------------------------------------------
// band entry a(ii,j) = A(j-4+ii, j) of A(omega,l) (cala + calbc,
// calmat.f90:747-809): contributions added in zone order, each formed as
// the reference forms it.
// ---------------------------------------------------------------------------
template <bool F = false>
__host__ __device__ inline double2 psv_entry(const double2 *col, int kbits, int ii, double l2, double lsq)
{
    const int k = (kbits >> (4 * ii)) & 3;
    const int two = (kbits >> (4 * ii + 2)) & 1;
    const double2 *e = col + 4 * ii;
    if (k == 0) return cz();
    if (k == 3) return e[0];
    const double s = (k == 1) ? l2 : lsq;
    double2 v = caxpy_t<F>(e[0], s, e[1]);
    if (two) v = cadd(v, caxpy_t<F>(e[2], s, e[3]));
    return v;
}

// ---------------------------------------------------------------------------
// Source excitation (calya/calg, calmat.f90:908-1243 with rk3.f90): the
// four fundamental solutions by one RK4 step per half interval, then the
// 4x4 pivoted elimination (glu2.f90) of the boundary conditions for the
// four distinct source vectors.  g[s][0..3] are the values at the rows
// jtmp..jtmp+3.
// ---------------------------------------------------------------------------
__host__ __device__ inline double psv_cabs(double2 a) { return hypot(a.x, a.y); }

// eqmotion1 (calmat.f90:1176-1208): f = (a + l2 b) y with the 4x4 matrix
// starting at entry `base` of the calabnum tables
__host__ __device__ inline void psv_eqmotion(const PsvSrc &S, int base, double l2, const double2 y[4], double2 f[4])
{
    for (int i = 0; i < 4; i++) {
        double2 acc = cz();
        for (int j = 0; j < 4; j++) {
            const double2 a = S.anum[base + i + 4 * j], b = S.bnum[base + i + 4 * j];
            const double2 c = make_double2(a.x + l2 * b.x, a.y + l2 * b.y);
            acc = cadd(acc, cmul(c, y[j]));
        }
        f[i] = acc;
    }
}

// one rkstep (rk3.f90:46-86) through the four matrices at bases b[0..3]
__host__ __device__ inline void psv_rk(const PsvSrc &S, const int b[4], double l2, double h, double2 y[4])
{
    const double c1 = 1.0 / 6.0, c2 = 1.0 / 3.0, c3 = c2, c4 = c1, b2 = 0.5, b3 = b2;
    double2 ak[4], w[4], yn[4];
    psv_eqmotion(S, b[0], l2, y, ak);
    for (int i = 0; i < 4; i++) yn[i] = cadd(y[i], cscale(ak[i], h * c1));
    for (int i = 0; i < 4; i++) w[i] = cadd(y[i], cscale(ak[i], h * b2));
    psv_eqmotion(S, b[1], l2, w, ak);
    for (int i = 0; i < 4; i++) yn[i] = cadd(yn[i], cscale(ak[i], h * c2));
    for (int i = 0; i < 4; i++) w[i] = cadd(y[i], cscale(ak[i], h * b3));
    psv_eqmotion(S, b[2], l2, w, ak);
    for (int i = 0; i < 4; i++) yn[i] = cadd(yn[i], cscale(ak[i], h * c3));
    for (int i = 0; i < 4; i++) w[i] = cadd(y[i], cscale(ak[i], h));
    psv_eqmotion(S, b[3], l2, w, ak);
    for (int i = 0; i < 4; i++) y[i] = cadd(yn[i], cscale(ak[i], h * c4));
}

// calya (calmat.f90:908-1010): each RK step uses a window of four of the
// ten tabulated matrices, selected by the reference's index arithmetic
// (window entry i of step k is table entry i + ktmp(segment of i))
__host__ __device__ inline void psv_calya(const PsvSrc &S, double l2, double ra1, double ra2, double r0,
                                          double2 ya[4], double2 yb[4], double2 yc[4], double2 yd[4])
{
    for (int i = 0; i < 4; i++) ya[i] = yb[i] = yc[i] = yd[i] = cz();
    ya[1] = make_double2(1.0, 0.0);
    yb[3] = make_double2(1.0, 0.0);
    yc[1] = make_double2(1.0, 0.0);
    yd[3] = make_double2(1.0, 0.0);
    double dr = (r0 - ra1) / 2.0;
    if (dr > 0.0) {
        double xs = ra1;
        for (int k = 1; k <= 2; k++) {
            const int kt1 = 32 * (k - 1), kt2 = kt1 - 16;
            const int b[4] = {0 + kt1, 16 + kt1, 32 + kt2, 48 + kt2};
            const double xe = xs + dr;
            psv_rk(S, b, l2, (xe - xs) / 1.0, ya);
            xs = xe - dr;
            psv_rk(S, b, l2, (xe - xs) / 1.0, yb);
            xs = xe;
        }
    }
    dr = (ra2 - r0) / 2.0;
    if (dr > 0.0) {
        double xs = ra2;
        for (int k = 1; k <= 2; k++) {
            const int kt1 = 144 - 32 * (k - 1), kt2 = kt1 - 32, kt3 = kt1 - 48, kt4 = kt1 - 80;
            const int b[4] = {0 + kt1, 16 + kt2, 32 + kt3, 48 + kt4};
            const double xe = xs - dr;
            psv_rk(S, b, l2, (xe - xs) / 1.0, yc);
            xs = xe + dr;
            psv_rk(S, b, l2, (xe - xs) / 1.0, yd);
            xs = xe;
        }
    }
}

// glu (glu2.f90:2-137) on an n x n column-major complex matrix with nrhs
// right-hand sides: the elimination steps on the matrix do not depend on
// the right-hand side, so one elimination serves all of them.
template <int N>
__host__ __device__ inline void psv_glu(double2 a[N * N], double2 (*b)[4], int nrhs)
{
#define AA(i, j) a[(i) + N * (j)]
    const double eps = 3.52e-15;
    int ip[N];
    for (int k = 0; k < N; k++) {
        double amax = psv_cabs(AA(k, k));
        int ipk = k;
        for (int i = k + 1; i < N; i++) {
            const double aik = psv_cabs(AA(i, k));
            if (aik > amax) {
                ipk = i;
                amax = aik;
            }
        }
        ip[k] = ipk;
        if (!(amax > eps)) {  // singular: the reference stops here
            for (int i = k + 1; i < N; i++) AA(i, k) = cz();
            ip[k] = k;
            break;
        }
        if (ipk != k) {
            const double2 t = AA(ipk, k);
            AA(ipk, k) = AA(k, k);
            AA(k, k) = t;
        }
        double2 wk[N];
        for (int i = k + 1; i < N; i++) {
            AA(i, k) = cdiv(cneg(AA(i, k)), AA(k, k));
            wk[i] = AA(i, k);
        }
        for (int j = k + 1; j < N; j++) {
            if (ipk != k) {
                const double2 t = AA(ipk, j);
                AA(ipk, j) = AA(k, j);
                AA(k, j) = t;
            }
            const double2 t = AA(k, j);
            for (int i = k + 1; i < N; i++) AA(i, j) = cadd(AA(i, j), cmul(wk[i], t));
        }
    }
    for (int r = 0; r < nrhs; r++) {
        double2 *bb = b[r];
        for (int k = 0; k < N; k++) {
            if (ip[k] != k) {
                const double2 t = bb[ip[k]];
                bb[ip[k]] = bb[k];
                bb[k] = t;
            }
            const double2 t = bb[k];
            for (int i = k + 1; i < N; i++) bb[i] = cadd(bb[i], cmul(AA(i, k), t));
        }
        bb[N - 1] = cdiv(bb[N - 1], AA(N - 1, N - 1));
        for (int k = N - 2; k >= 0; k--) {
            const double2 t = bb[k + 1];
            for (int i = 0; i <= k; i++) bb[i] = csub(bb[i], cmul(AA(i, k + 1), t));
            bb[k] = cdiv(bb[k], AA(k, k));
        }
    }
#undef AA
}

// calg (calmat.f90:1075-1171) for the four distinct source vectors of
// l >= 1: g[0] Y1 (m=0 imt1), g[1] Y2 (m=0 imt2), g[2] X1 (m=-1 imt4),
// g[3] X2 (m=+2 imt2, l >= 2).
__host__ __device__ inline void psv_excitation(int l, const PsvSrc &S, const PsvSrcConst &C, double2 g[4][4])
{
    const double l2 = (double)l * (double)(l + 1);
    const double lsq = sqrt(l2);
    double2 ya[4], yb[4], yc[4], yd[4];
    psv_calya(S, l2, C.ra1, C.ra2, C.r0, ya, yb, yc, yd);
    const double r0 = C.r0;
    const double b1 = sqrt((double)(2 * l + 1) / (16.0 * DSM_PI));
    const double b2 = sqrt((double)(2 * l + 1) * (double)(l - 1) * (double)(l + 2) / (64.0 * DSM_PI));
    const double2 c1 = S.coef1, c2 = S.coef2;
    double2 s[4][4];
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < 4; i++) s[r][i] = cz();
    {
        // m = 0: dd and xtmp
        const double2 den = cadd(cscale(c2, C.ecC0 - 4.0 / 3.0 * C.ecL0), cscale(c1, 4.0 / 3.0 * C.ecL0));
        const double2 num = csub(cscale(c2, C.ecF0 + 2.0 / 3.0 * C.ecL0), cscale(c1, 2.0 / 3.0 * C.ecL0));
        const double2 xtmp = cdiv(num, den);
        const double r03 = r0 * r0 * r0;
        // imt 1 (mt11 = 1)
        s[0][0] = cdiv(make_double2(2.0 * b1 * 1.0 / (r0 * r0), 0.0), den);
        const double2 s1a = cadd(make_double2(-2.0 * b1 * (0.0 + 0.0) / r03, 0.0), cscale(xtmp, 4.0 * b1 * 1.0 / r03));
        const double2 s2a = csub(make_double2(b1 * lsq * (0.0 + 0.0) / r03, 0.0), cscale(xtmp, 2.0 * b1 * lsq * 1.0 / r03));
        s[0][1] = s1a;
        s[0][3] = make_double2(s2a.x / lsq, s2a.y / lsq);
        // imt 2 (mt22 = 1)
        s[1][0] = cdiv(make_double2(2.0 * b1 * 0.0 / (r0 * r0), 0.0), den);
        const double2 s1b = cadd(make_double2(-2.0 * b1 * (1.0 + 0.0) / r03, 0.0), cscale(xtmp, 4.0 * b1 * 0.0 / r03));
        const double2 s2b = csub(make_double2(b1 * lsq * (1.0 + 0.0) / r03, 0.0), cscale(xtmp, 2.0 * b1 * lsq * 0.0 / r03));
        s[1][1] = s1b;
        s[1][3] = make_double2(s2b.x / lsq, s2b.y / lsq);
    }
    {
        // m = -1, imt 4 (mt12 = 1, sgn = -1)
        const double2 den = cscale(c1, r0 * r0 * C.ecL0);
        const double2 ee = cdiv(make_double2(-b1 * -1.0 * 1.0, b1 * 0.0), den);
        s[2][2] = make_double2(ee.x / lsq, ee.y / lsq);
    }
    if (l >= 2) {
        // m = +2, imt 2 (mt22 = 1)
        const double r03 = r0 * r0 * r0;
        const double2 s2 = make_double2(-b2 * (1.0 - 0.0) / r03, 1.0 * 2.0 * b2 * 0.0 / r03);
        s[3][3] = make_double2(s2.x / lsq, s2.y / lsq);
    }
    double2 a[16];
    for (int i = 0; i < 4; i++) {
        a[i + 0] = cneg(ya[i]);
        a[i + 4] = cneg(yb[i]);
        a[i + 8] = yc[i];
        a[i + 12] = yd[i];
    }
    psv_glu<4>(a, s, (l >= 2) ? 4 : 3);
    double dt[4];
    dt[0] = -C.ra1 * C.ra1;
    dt[1] = dt[0] * lsq;
    dt[2] = C.ra2 * C.ra2;
    dt[3] = dt[2] * lsq;
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < 4; i++) g[r][i] = (r == 3 && l < 2) ? cz() : make_double2(s[r][i].x * dt[i], s[r][i].y * dt[i]);
}

// l = 0 source vectors (sab2 branch of calg): Y1 (imt 1) and Y2 (imt 2) at
// the U rows jtmp and jtmp+2.
__host__ __device__ inline void psv_excitation_l0(const PsvSrc &S, const PsvSrcConst &C, double2 g[2][2])
{
    double2 ya[4], yb[4], yc[4], yd[4];
    psv_calya(S, 0.0, C.ra1, C.ra2, C.r0, ya, yb, yc, yd);
    const double r0 = C.r0;
    const double b1 = sqrt(1.0 / (16.0 * DSM_PI));
    const double2 c1 = S.coef1, c2 = S.coef2;
    const double2 den = cadd(cscale(c2, C.ecC0 - 4.0 / 3.0 * C.ecL0), cscale(c1, 4.0 / 3.0 * C.ecL0));
    const double2 num = csub(cscale(c2, C.ecF0 + 2.0 / 3.0 * C.ecL0), cscale(c1, 2.0 / 3.0 * C.ecL0));
    const double2 xtmp = cdiv(num, den);
    const double r03 = r0 * r0 * r0;
    double2 s[2][4];
    s[0][0] = cdiv(make_double2(2.0 * b1 * 1.0 / (r0 * r0), 0.0), den);
    s[0][1] = cadd(make_double2(-2.0 * b1 * (0.0 + 0.0) / r03, 0.0), cscale(xtmp, 4.0 * b1 * 1.0 / r03));
    s[1][0] = cdiv(make_double2(2.0 * b1 * 0.0 / (r0 * r0), 0.0), den);
    s[1][1] = cadd(make_double2(-2.0 * b1 * (1.0 + 0.0) / r03, 0.0), cscale(xtmp, 4.0 * b1 * 0.0 / r03));
    double2 a[4];
    a[0] = cneg(ya[0]);
    a[1] = cneg(ya[1]);
    a[2] = yc[0];
    a[3] = yc[1];
    psv_glu<2>(a, s, 2);
    const double dt0 = -C.ra1 * C.ra1, dt2 = C.ra2 * C.ra2;
    for (int r = 0; r < 2; r++) {
        g[r][0] = make_double2(s[r][0].x * dt0, s[r][0].y * dt0);
        g[r][1] = make_double2(s[r][1].x * dt2, s[r][1].y * dt2);  // b(3) = b(2)
    }
}

// ---------------------------------------------------------------------------
// Band solve for l >= 1.  One lane per (omega,l); 64 consecutive l of one
// frequency per wavefront, so the band pieces of the frequency are read at
// one address by all lanes.  Factorization dcsymbdl0 (dcsymbdl.f90:2-88,
// m=3) in the reference's order, the forward substitution of dcsbdlv0
// (dcsymbdl.f90:93-148) fused into the same upward sweep, then the
// backward substitution from the surface down to the deepest row a
// receiver reads, recording the six solutions at the rows in need[].
// Local row k (1..n) is band row row0+k-1.
// ---------------------------------------------------------------------------
// Factorization window of dcsymbdl0 at step k (values before the step):
//   d0 = a(4,k)  e01 = a(3,k+1)  e02 = a(2,k+2)
//   d1 = a(4,k+1)  e12 = a(3,k+2)  d2 = a(4,k+2)
// and the L entries of the last three steps (a(1,k) = w4 of step k-3,
// a(2,k) = w3 of step k-2, a(3,k) = w2 of step k-1).
struct PsvWin {
    double2 d0, e01, e02, d1, e12, d2;
    double2 w4m3, w4m2, w4m1, w3m2, w3m1, w2m1;
};

// The four band entries a(1..4, j) of column j for this lane's l: from the
// full column (16 values, entry ii = e0 + s e1 (+ e2 + s e3), psv_entry;
// all values loaded first so the reads issue together), or, for the fused
// family, from the compact column of psv_compact_kernel (entry ii =
// A + s B with s = lsq where bit ii of sel is set, else l(l+1)).
__host__ __device__ inline void psv_col_exact(const double2 *col, int kbits, double l2, double lsq, double2 a[4])
{
    double2 e[16];
    for (int t = 0; t < 16; t++) e[t] = col[t];
    for (int ii = 0; ii < 4; ii++) a[ii] = psv_entry<false>(e, kbits, ii, l2, lsq);
}
__host__ __device__ inline void psv_col_fast(const double2 *col, int sel, double l2, double lsq, double2 a[4])
{
    double2 e[8];
    for (int t = 0; t < 8; t++) e[t] = col[t];
    for (int ii = 0; ii < 4; ii++) a[ii] = caxpy_t<true>(e[2 * ii], ((sel >> ii) & 1) ? lsq : l2, e[2 * ii + 1]);
}

// One step k < n of dcsymbdl0 (m = 3): the main update for k <= n-3 with
// the entries a3[ii] = a(ii+1, k+3), the end-of-band updates of
// dcsymbdl.f90:62-84 for k = n-2, n-1 (whose product order differs).
// piv = 1 / W.d0 (crecip_t<F>).  Returns w2, w3, w4 of the step.
template <bool F = false>
__host__ __device__ inline void psv_factor_step(PsvWin &W, int k, int n, const double2 *a3, double2 piv, double2 &w2,
                                                double2 &w3, double2 &w4)
{
    w3 = cz();
    w4 = cz();
    if (k <= n - 3) {
        const double2 a13 = a3[0], a23 = a3[1], a33 = a3[2], a43 = a3[3];
        const double2 zz2 = cneg(W.e01), zz3 = cneg(W.e02), zz4 = cneg(a13);
        w2 = cmul_t<F>(W.e01, piv);
        w3 = cmul_t<F>(W.e02, piv);
        w4 = cmul_t<F>(a13, piv);
        const double2 nd1 = cmadd_t<F>(W.d1, w2, zz2);
        const double2 nd2 = cmadd_t<F>(W.d2, w3, zz3);
        const double2 nd3 = cmadd_t<F>(a43, w4, zz4);
        const double2 ne12 = cmadd_t<F>(W.e12, w2, zz3);
        const double2 ne23 = cmadd_t<F>(a33, w3, zz4);
        const double2 ne13 = cmadd_t<F>(a23, w2, zz4);
        W.d0 = nd1;
        W.e01 = ne12;
        W.e02 = ne13;
        W.d1 = nd2;
        W.e12 = ne23;
        W.d2 = nd3;
    } else if (k == n - 2) {
        const double2 zz2 = cneg(W.e01), zz3 = cneg(W.e02);
        w2 = cmul_t<F>(W.e01, piv);
        const double2 nd1 = cmadd_t<F>(W.d1, w2, zz2);
        w3 = cmul_t<F>(W.e02, piv);
        const double2 ne12 = cmadd_t<F>(W.e12, w3, zz2);
        const double2 nd2 = cmadd_t<F>(W.d2, w3, zz3);
        W.d0 = nd1;
        W.e01 = ne12;
        W.d1 = nd2;
    } else {  // k == n - 1
        const double2 zz2 = cneg(W.e01);
        w2 = cmul_t<F>(W.e01, piv);
        W.d0 = cmadd_t<F>(W.d1, w2, zz2);
    }
    W.w4m3 = W.w4m2;
    W.w4m2 = W.w4m1;
    W.w4m1 = w4;
    W.w3m2 = W.w3m1;
    W.w3m1 = w3;
    W.w2m1 = w2;
}

// ---------------------------------------------------------------------------
// Band solve for l >= 1.  One lane per (omega,l); 64 consecutive l of one
// frequency per wavefront, so the band pieces of the frequency are read at
// one address by all lanes.  Factorization dcsymbdl0 (dcsymbdl.f90:2-88,
// m=3) in the reference's order, the forward substitution of dcsbdlv0
// (dcsymbdl.f90:93-148) fused into the same upward sweep from the source
// rows on, then the backward substitution from the surface down to the
// deepest row a receiver reads, recording the six solutions at the rows
// in need[].  Local row k (1..n) is band row row0+k-1.
// ---------------------------------------------------------------------------
__host__ __device__ inline void psv_solve_lane(const PsvSolveArgs &a, int b, int tx)
{
    const int sl = a.slot_list[b / a.nlb];
    const int l = a.l0 + (b % a.nlb) * 64 + tx;
    if (l > a.lend || l == 0) return;
    const long gid = (long)b * 64 + tx;
    const long ss = a.sstride;
    const double l2 = (double)l * (double)(l + 1);
    const double lsq = sqrt(l2);
    const int n = a.nn - a.row0 + 1;
    const int fs = a.jtmp - a.row0 + 1;   // local row of the first source row (>= 3)
    const int klo = a.rowlo - a.row0 + 1; // lowest local row needed
    const double2 *band = a.band + (size_t)sl * a.nn * 16;
    double2 *Ls = a.scratch;                                   // [k - klo][3]
    double2 *Zs = a.scratch + (size_t)3 * (n - klo + 1) * ss;  // [k - fs][4]
    auto col = [&](int k) { return band + (size_t)(a.row0 + k - 2) * 16; };
    auto kb = [&](int k) { return a.kind[a.row0 + k - 2]; };
    auto storeL = [&](int k, double2 w2, double2 w3, double2 w4) {
        if (k >= klo) {
            Ls[(size_t)(3 * (k - klo) + 0) * ss + gid] = w2;
            Ls[(size_t)(3 * (k - klo) + 1) * ss + gid] = w3;
            Ls[(size_t)(3 * (k - klo) + 2) * ss + gid] = w4;
        }
    };
    PsvWin W;
    W.d0 = psv_entry(col(1), kb(1), 3, l2, lsq);
    W.e01 = psv_entry(col(2), kb(2), 2, l2, lsq);
    W.d1 = psv_entry(col(2), kb(2), 3, l2, lsq);
    W.e02 = psv_entry(col(3), kb(3), 1, l2, lsq);
    W.e12 = psv_entry(col(3), kb(3), 2, l2, lsq);
    W.d2 = psv_entry(col(3), kb(3), 3, l2, lsq);
    W.w4m3 = W.w4m2 = W.w4m1 = W.w3m2 = W.w3m1 = W.w2m1 = cz();
    // phase 1: rows below the source, factorization only
    for (int k = 1; k < fs; k++) {
        double2 w2, w3, w4;
        double2 a3[4];
        psv_col_exact(col(k + 3), kb(k + 3), l2, lsq, a3);
        psv_factor_step(W, k, n, a3, crecip(W.d0), w2, w3, w4);
        storeL(k, w2, w3, w4);
    }
    // phase 2: from the source rows up, factorization + forward values of
    // the four source solutions (their right-hand sides sit at rows
    // fs..fs+3; below, their forward values are zero)
    const int nsrc = (l >= 2) ? 4 : 3;
    double2 Dn1 = cz(), Dn = cz(), w2n1 = cz();
    {
        double2 g[4][4];
        psv_excitation(l, a.src[sl], a.sc, g);
        double2 z1[4], z2[4], z3[4];
        for (int s = 0; s < 4; s++) z1[s] = z2[s] = z3[s] = cz();
        for (int k = fs; k <= n; k++) {
            const double2 a1 = W.w4m3, a2 = W.w3m2, a3 = W.w2m1;
            for (int s = 0; s < nsrc; s++) {
                const double2 bk = (k - fs < 4) ? g[s][k - fs] : cz();
                double2 sum;
                if (k == 3) sum = cadd(cmul(a2, z2[s]), cmul(a3, z3[s]));
                else sum = cadd(cadd(cmul(a1, z1[s]), cmul(a2, z2[s])), cmul(a3, z3[s]));
                const double2 zk = csub(bk, sum);
                z1[s] = z2[s];
                z2[s] = z3[s];
                z3[s] = zk;
                Zs[(size_t)(4 * (k - fs) + s) * ss + gid] = cdiv(zk, W.d0);
            }
            if (k == n - 1) Dn1 = W.d0;
            if (k == n) {
                Dn = W.d0;
                break;
            }
            double2 w2, w3, w4;
            double2 c3v[4];
            psv_col_exact(col(k + 3 <= n ? k + 3 : n), kb(k + 3 <= n ? k + 3 : n), l2, lsq, c3v);
            psv_factor_step(W, k, n, c3v, crecip(W.d0), w2, w3, w4);
            storeL(k, w2, w3, w4);
            if (k == n - 1) w2n1 = w2;
        }
    }
    // back-propagated forces: forward values at rows n-1, n only
    const double c0f = (double)(float)sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / 1.0);
    const double cm1 = sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / (1.0 * (double)l * (double)(l + 1)));
    const double hs = l2 * cm1 / 2.0;
    const double2 yb_rhs = make_double2(-c0f, 0.0);
    const double2 s_rhs = cdiv(make_double2(hs, 0.0), make_double2(lsq, 0.0));
    const double2 zb_n1 = cdiv(yb_rhs, Dn1);
    // z(n) = 0 - (a(1,n) 0 + a(2,n) 0 + a(3,n) z(n-1))
    const double2 zb_n = cdiv(csub(cz(), cadd(cadd(cmul(W.w4m3, cz()), cmul(W.w3m2, cz())), cmul(w2n1, yb_rhs))), Dn);
    const double2 zs_n = cdiv(s_rhs, Dn);

    // backward substitution: b(n), b(n-1), then rows n-2 .. klo
    double2 x1[PSV_NSOL], x2[PSV_NSOL], x3[PSV_NSOL];  // b(k+1), b(k+2), b(k+3)
    for (int s = 0; s < PSV_NSOL; s++) x1[s] = x2[s] = x3[s] = cz();
    int p = 0;
    double2 *xo = a.xn + (size_t)sl * a.nneed * PSV_NSOL * a.lcap + l;
    for (int k = n; k >= klo; k--) {
        double2 zd[PSV_NSOL];
        for (int s = 0; s < 4; s++)
            zd[s] = (k >= fs && s < nsrc) ? Zs[(size_t)(4 * (k - fs) + s) * ss + gid] : cz();
        zd[4] = (k == n) ? zb_n : (k == n - 1) ? zb_n1 : cz();
        zd[5] = (k == n) ? zs_n : cz();
        double2 bk[PSV_NSOL];
        if (k == n) {
            for (int s = 0; s < PSV_NSOL; s++) bk[s] = zd[s];
        } else {
            const double2 L2 = Ls[(size_t)(3 * (k - klo) + 0) * ss + gid];
            if (k == n - 1) {
                for (int s = 0; s < PSV_NSOL; s++) bk[s] = csub(zd[s], cmul(L2, x1[s]));
            } else {
                const double2 L3 = Ls[(size_t)(3 * (k - klo) + 1) * ss + gid];
                if (k == n - 2) {
                    for (int s = 0; s < PSV_NSOL; s++)
                        bk[s] = csub(zd[s], cadd(cmul(L3, x2[s]), cmul(L2, x1[s])));
                } else {
                    const double2 L4 = Ls[(size_t)(3 * (k - klo) + 2) * ss + gid];
                    for (int s = 0; s < PSV_NSOL; s++)
                        bk[s] = csub(zd[s], cadd(cadd(cmul(L4, x3[s]), cmul(L3, x2[s])), cmul(L2, x1[s])));
                }
            }
        }
        for (int s = 0; s < PSV_NSOL; s++) {
            x3[s] = x2[s];
            x2[s] = x1[s];
            x1[s] = bk[s];
        }
        const int row = a.row0 + k - 1;
        if (p < a.nneed && a.need[p] == row) {
            for (int s = 0; s < PSV_NSOL; s++) xo[(size_t)(PSV_NSOL * p + s) * a.lcap] = bk[s];
            p++;
        }
    }
}

// ---------------------------------------------------------------------------
// Coefficient rows.  For a solution Z with interpolated values U, V and
// radial derivatives Ud, Vd at a receiver (interpolate, others.f90:
// 1170-1239), Vl = V/L, Vdl = Vd/L, L = sqrt(l(l+1)), irf = 1/float(r),
// the displacement gradient of order m = sigma*mu (calup +
// locallyCartesianDerivatives, others.f90:310-344, 983-995, with
// caldvecphi0_withplm, trial.f90:97-223; p = (-1)^mu for m < 0, else 1) is
//   U11 = p Ud F0            U12 = p ([m>0] U - Vl) F1 irf
//   U13 = p i m (U F5 - Vl F4) irf      U21 = p Vdl F1
//   U22 = p (Vl F2 + U F0) irf          U23 = p i m Vl (F6 - F7) irf
//   U31 = p i m Vdl F4       U32 = p i m Vl F3 irf
//   U33 = p (-m^2 Vl F8 + U F0 + Vl F9) irf
// with the real theta functions F0..F9 of legendre_lane.h (the [m>0] is
// the reference's dvecdt(1,-m), never assigned and zero).  Summing the
// strain maps udertotsgt/udertorsgt (others.f90:204-285) over m and over
// the multiples of the six solutions gives the rows below; the slots built
// from them are listed in psv_jobs.h.  In the fluid the strain is
// K U ylm on the diagonal (calupfluid, others.f90:1009-1031).
// ---------------------------------------------------------------------------
struct PsvVals {
    double2 U, V, Ud, Vd;
};

// The rows built from solution S alone (z[S] = a); in the fluid, S = 0
// also emits the rows that are zero there.  psv_rows emits all of them.
template <int S, class Put>
__host__ __device__ inline void psv_rows_one(const PsvVals &a, bool fluid, double2 K, double il, double irf,
                                             bool has_mu1, bool has_mu2, Put &&put)
{
    const double2 Z = cz();
    if (fluid) {
        if (S == 0) {
            const double2 u1 = cmul(K, a.U);
            for (int q = 0; q < PSV_NQ; q++) {
                switch (q) {
                case 0: put(q, u1); break;
                case 2: put(q, cneg(u1)); break;
                case 1: case 5: case 14: case 15: case 29: case 30: case 42: case 46: case 49: case 50: break;
                default: put(q, Z); break;
                }
            }
        } else if (S == 1) {
            const double2 u2 = cmul(K, a.U);
            put(1, cneg(u2));
            put(5, u2);
        } else if (S == 4) {
            const double2 ub = cmul(K, a.U);
            put(42, ub);
            put(46, cneg(ub));
        } else if (S == 2 || S == 5) {
            const double2 ux = has_mu1 ? cmul(K, a.U) : Z;
            put(S == 2 ? 14 : 49, cscale(ux, -2.0));
            put(S == 2 ? 15 : 50, cscale(ux, 2.0));
        } else {  // S == 3
            const double2 u2x = has_mu2 ? cmul(K, a.U) : Z;
            put(29, cscale(u2x, 2.0));
            put(30, cscale(u2x, -2.0));
        }
        return;
    }
    if (S == 0) {  // mu = 0: Y1 (imt 1)
        const double2 Vdl = cscale(a.Vd, il);
        const double2 Ui = cscale(a.U, irf), Vli = cscale(cscale(a.V, il), irf);
        put(0, a.Ud);
        put(2, cneg(Ui));
        put(3, cscale(Vli, -0.5));
        put(4, cscale(Vli, -0.5));
        put(8, cscale(cadd(csub(Ui, Vli), Vdl), -0.5));
        put(9, cscale(Vli, 0.5));
        put(10, cscale(Vli, -0.5));
    } else if (S == 1) {  // Y2 (imt 2 = imt 3)
        const double2 Vdl = cscale(a.Vd, il);
        const double2 Ui = cscale(a.U, irf), Vli = cscale(cscale(a.V, il), irf);
        put(1, cneg(a.Ud));
        put(5, Ui);
        put(6, cscale(Vli, 0.5));
        put(7, cscale(Vli, 0.5));
        put(11, cscale(Vli, -0.5));
        put(12, cscale(Vli, 0.5));
        put(13, cscale(cadd(csub(Ui, Vli), Vdl), 0.25));
    } else if (S == 4) {  // YB (back-propagated icomp 1)
        const double2 Vdl = cscale(a.Vd, il);
        const double2 Ui = cscale(a.U, irf), Vli = cscale(cscale(a.V, il), irf);
        put(42, a.Ud);
        put(43, cscale(cadd(csub(Ui, Vli), Vdl), -0.5));
        put(44, cscale(Vli, 0.5));
        put(45, cscale(Vli, -0.5));
        put(46, cneg(Ui));
        put(47, cscale(Vli, -0.5));
        put(48, cscale(Vli, -0.5));
    } else if (S == 2 || S == 5) {  // mu = 1: imt 4 / 5 (X1) and the back-propagated icomp 2 / 3 (S)
        const double2 Vdl = has_mu1 ? cscale(a.Vd, il) : Z;
        const double2 Ui = has_mu1 ? cscale(a.U, irf) : Z;
        const double2 Vli = has_mu1 ? cscale(cscale(a.V, il), irf) : Z;
        const double2 Ud = has_mu1 ? a.Ud : Z;
        const int o = (S == 2) ? 14 : 49;
        put(o, cscale(Ud, -2.0));
        put(o + 1, cscale(Ui, 2.0));
        put(o + 2, Vli);
        put(o + 3, cneg(Vli));
        put(o + 4, Vli);
        const double2 ud = csub(csub(cscale(Vli, 2.0), Ui), cscale(Vdl, 2.0));  // (2Vl-U)/r - 2Vdl
        if (S == 2) {
            put(19, cscale(ud, 0.25));
            put(20, cscale(Vli, -0.5));
            put(21, cscale(Vli, -0.5));
            put(22, cscale(Vli, 0.5));
            put(23, cscale(csub(Vli, Vdl), 0.5));
            put(24, cscale(Ui, -0.5));
            put(25, cscale(Vli, -0.5));
            put(26, cscale(Vli, -0.5));
            put(27, cscale(Vli, 0.5));
        } else {
            put(54, Vli);
            put(55, Vli);
            put(56, cneg(Vli));
            put(57, cscale(ud, -0.5));
            put(58, cneg(Vli));
            put(59, cneg(Vli));
            put(60, Vli);
            put(61, csub(Vdl, Vli));
            put(62, Ui);
        }
    } else {  // S == 3, mu = 2: imt 2 / 3 / 6 (X2)
        const double2 Vdl = has_mu2 ? cscale(a.Vd, il) : Z;
        const double2 Ui = has_mu2 ? cscale(a.U, irf) : Z;
        const double2 Vli = has_mu2 ? cscale(cscale(a.V, il), irf) : Z;
        const double2 Ud = has_mu2 ? a.Ud : Z;
        put(28, cscale(cadd(csub(Ui, cscale(Vli, 2.0)), cscale(Vdl, 2.0)), 0.5));
        put(29, cscale(Ud, 2.0));
        put(30, cscale(Ui, -2.0));
        put(31, cneg(Vli));
        put(32, cscale(Vli, 4.0));
        put(33, cneg(Vli));
        put(34, cscale(Vli, 0.5));
        put(35, cscale(Vli, 2.0));
        put(36, cscale(Vli, -0.5));
        put(37, cscale(csub(Vdl, Vli), 2.0));
        put(38, cscale(Ui, 2.0));
        put(39, cscale(Vli, 2.0));
        put(40, cscale(Vli, 2.0));
        put(41, cscale(Vli, -2.0));
    }
}

template <class Put>
__host__ __device__ inline void psv_rows(const PsvVals z[PSV_NSOL], bool fluid, double2 K, double il, double irf,
                                         bool has_mu1, bool has_mu2, Put &&put)
{
    psv_rows_one<0>(z[0], fluid, K, il, irf, has_mu1, has_mu2, put);
    psv_rows_one<1>(z[1], fluid, K, il, irf, has_mu1, has_mu2, put);
    psv_rows_one<2>(z[2], fluid, K, il, irf, has_mu1, has_mu2, put);
    psv_rows_one<3>(z[3], fluid, K, il, irf, has_mu1, has_mu2, put);
    psv_rows_one<4>(z[4], fluid, K, il, irf, has_mu1, has_mu2, put);
    psv_rows_one<5>(z[5], fluid, K, il, irf, has_mu1, has_mu2, put);
}

// SYNN rows from the surface values (Us, Vs) = g0(nn-1), g0(nn) of the
// solutions (calu + utosynn, others.f90:161-198, 968-978)
__host__ __device__ inline void psv_synn_rows(const double2 us[PSV_NSOL], const double2 vs[PSV_NSOL], double il,
                                              bool has_mu1, bool has_mu2, double2 es[PSV_NQS])
{
    for (int q = 0; q < PSV_NQS; q++) es[q] = cz();
    es[0] = us[0];
    es[3] = cneg(us[1]);
    es[4] = cscale(vs[0], il);
    es[5] = cneg(cscale(vs[1], il));
    if (has_mu1) {
        es[1] = cscale(us[2], -2.0);
        es[8] = cscale(cscale(vs[2], il), -2.0);
        es[9] = cscale(cscale(vs[2], il), -2.0);
    }
    if (has_mu2) {
        es[2] = cscale(us[3], 2.0);
        es[6] = cscale(cscale(vs[3], il), -2.0);
        es[7] = cscale(cscale(vs[3], il), -8.0);
    }
}

__host__ __device__ inline void psv_coef_lane(const PsvCoefArgs &a, int b, int tx)
{
    const int sl = a.slot_list[b / a.nlb];
    const int l = a.l0 + (b % a.nlb) * 64 + tx;
    if (l > a.lend || l == 0) return;  // l = 0: psv_l0_lane
    const int R = a.r_n, LC = a.lcap;
    const double il = 1.0 / sqrt((double)l * (double)(l + 1));
    const bool mu1 = true, mu2 = (l >= 2);
    const unsigned char *rneed = a.rneed ? a.rneed + (size_t)((l - a.l0) >> 8) * R : nullptr;
    const double2 *X = a.xn + (size_t)sl * a.nneed * PSV_NSOL * LC + l;
    auto xv = [&](int p, int s) { return X[(size_t)(PSV_NSOL * p + s) * LC]; };
    // amplitude rows of the top receiver (calamp on imt 1, m = 0)
    a.ampuv[((size_t)sl * LC + l) * 2 + 0] = xv(a.amp_i0, 0);
    a.ampuv[((size_t)sl * LC + l) * 2 + 1] = xv(a.amp_i1, 0);
    {
        double2 us[PSV_NSOL], vs[PSV_NSOL], es[PSV_NQS];
        for (int s = 0; s < PSV_NSOL; s++) {
            us[s] = xv(a.syn_i0, s);
            vs[s] = xv(a.syn_i1, s);
        }
        psv_synn_rows(us, vs, il, mu1, mu2, es);
        if (a.Es) {
            double2 *Es = a.Es + (size_t)sl * PSV_NQS * LC + l;
            for (int q = 0; q < PSV_NQS; q++) Es[(size_t)q * LC] = es[q];
        } else {
            for (int q = 0; q < PSV_NQS; q++) a.esb.put(q, 0, sl, l, es[q]);
        }
    }
    double2 *E = a.E + (size_t)sl * PSV_NQ * R * LC + l;
    const int kst = a.kstart ? a.kstart[b / a.nlb * a.nblk + ((l - a.l0) >> 8)] : 1;
    const int nq = a.nq ? a.nq : PSV_NQ;
    for (int r = 0; r < R; r++) {
        if (rneed && !rneed[r]) continue;  // below the solved rows of every slot the tile holds
        if (a.kstart && kst > a.rlow[r]) {  // below this block's solved rows: zero rows
            if (a.eb.Eb)
                for (int q = 0; q < nq; q++) a.eb.put(q, r, sl, l, cz());
            else
                for (int q = 0; q < nq; q++) E[(size_t)(q * R + r) * LC] = cz();
            continue;
        }
        const int *rr = a.rec_row + 6 * r;
        const double *w0 = a.w0 + 3 * r, *w1 = a.w1 + 3 * r;
        PsvVals z[PSV_NSOL];
        for (int s = 0; s < PSV_NSOL; s++) {
            const double2 g0 = xv(rr[0], s), g1 = xv(rr[1], s), g2 = xv(rr[2], s), g3 = xv(rr[3], s),
                          g4 = xv(rr[4], s), g5 = xv(rr[5], s);
            z[s].U = cadd(cadd(cscale(g0, w0[0]), cscale(g2, w0[1])), cscale(g4, w0[2]));
            z[s].V = cadd(cadd(cscale(g1, w0[0]), cscale(g3, w0[1])), cscale(g5, w0[2]));
            z[s].Ud = cadd(cadd(cscale(g0, w1[0]), cscale(g2, w1[1])), cscale(g4, w1[2]));
            z[s].Vd = cadd(cadd(cscale(g1, w1[0]), cscale(g3, w1[1])), cscale(g5, w1[2]));
        }
        if (a.eb.Eb)
            psv_rows(z, a.fluid[r] != 0, a.fluidK[(size_t)sl * R + r], il, a.irf[r], mu1, mu2,
                     [&](int q, double2 v) {
                         if (q < nq) a.eb.put(q, r, sl, l, v);
                     });
        else
            psv_rows(z, a.fluid[r] != 0, a.fluidK[(size_t)sl * R + r], il, a.irf[r], mu1, mu2,
                     [&](int q, double2 v) {
                         if (q < nq) E[(size_t)(q * R + r) * LC] = v;
                     });
    }
}

// ---------------------------------------------------------------------------
// l = 0 (SGTpsv-MPI.f90:698-855): the U-only system of rea2, tridiagonal
// LDL^T (dcsymbdl0 / dcsbdlv0 with m = 1), three right-hand sides, and
// its coefficient rows at l = 0.  Four phases per frequency t, the lanes
// of one workgroup sharing t (psv_kernels.hip):
//   assemble (nodes in parallel), factor (one lane), substitute (one lane
//   per right-hand side), rows (one lane per receiver, lane R: SYNN).
// The serial sweeps carry the running pivot / solution in registers and
// read the band in groups of PSV_L0_G rows issued ahead of their use.
// work[t]: c(1,i) [N0], c(2,i) [N0], z [3][N0].
// ---------------------------------------------------------------------------
constexpr int PSV_L0_G = 16;

__host__ __device__ inline void psv_l0_assemble(const PsvL0Args &a, int t, int i)
{
    const int sl = a.slot_list[t];
    const int N0 = a.nn0;
    const double2 *band = a.band + (size_t)sl * a.nn * 16;
    double2 *c1 = a.work + (size_t)t * 5 * N0, *c2 = c1 + N0;
    const int row = a.node_row[i];
    const double2 *col = band + (size_t)(row - 1) * 16;
    const int kb = a.kind[row - 1];
    c2[i] = psv_entry(col, kb, 3, 0.0, 0.0);
    c1[i] = a.node_c1[i] ? psv_entry(col, kb, a.node_c1[i] - 1, 0.0, 0.0) : cz();
}

// dcsymbdl0, m = 1: for k = 0..n-2 (local nodes from row0_0)
//   w = c1(k+1) / c2(k);  c2(k+1) = c2(k+1) + w (-c1(k+1));  c1(k+1) = w
__host__ __device__ inline void psv_l0_factor(const PsvL0Args &a, int t)
{
    const int N0 = a.nn0, i0 = a.row0_0 - 1, n = N0 - i0;
    double2 *C1 = a.work + (size_t)t * 5 * N0 + i0, *C2 = C1 + N0;
    double2 d = C2[0];
    double2 p1[PSV_L0_G], p2[PSV_L0_G];
    for (int u = 0; u < PSV_L0_G; u++)
        if (1 + u < n) {
            p1[u] = C1[1 + u];
            p2[u] = C2[1 + u];
        }
    for (int k0 = 0; k0 < n - 1; k0 += PSV_L0_G) {
        double2 c1v[PSV_L0_G], c2v[PSV_L0_G];
        for (int u = 0; u < PSV_L0_G; u++) {
            c1v[u] = p1[u];
            c2v[u] = p2[u];
        }
        for (int u = 0; u < PSV_L0_G; u++)  // next group, in flight during this one
            if (k0 + PSV_L0_G + 1 + u < n) {
                p1[u] = C1[k0 + PSV_L0_G + 1 + u];
                p2[u] = C2[k0 + PSV_L0_G + 1 + u];
            }
        for (int u = 0; u < PSV_L0_G && k0 + u < n - 1; u++) {
            const double2 piv = crecip(d);
            const double2 zz = cneg(c1v[u]);
            const double2 w = cmul(c1v[u], piv);
            d = cadd(c2v[u], cmul(w, zz));
            C1[k0 + u + 1] = w;
            C2[k0 + u + 1] = d;
        }
    }
}

// dcsbdlv0, m = 1, right-hand side s: 0, 1 the source vectors at the
// nodes js, js+1; 2 the back-propagated force -c0 at the surface node.
// Forward z(j) -= c1(j) z(j-1), then z(j) /= c2(j), then backward
// z(j) -= c1(j+1) z(j+1); the division is done as each forward value is
// final (same operations).
__host__ __device__ inline void psv_l0_subst(const PsvL0Args &a, int t, int s)
{
    const int sl = a.slot_list[t];
    const int N0 = a.nn0, i0 = a.row0_0 - 1, n = N0 - i0;
    const double2 *C1 = a.work + (size_t)t * 5 * N0 + i0, *C2 = C1 + N0;
    double2 *zs = a.work + (size_t)t * 5 * N0 + (size_t)(2 + s) * N0;
    for (int i = 0; i < i0; i++) zs[i] = cz();
    double2 *Z = zs + i0;
    int ja = 0, jb = -1;
    double2 va = cz(), vb = cz();
    if (s < 2) {
        double2 g[2][2];
        psv_excitation_l0(a.src[sl], a.sc, g);
        ja = a.js - i0;
        jb = ja + 1;
        va = g[s][0];
        vb = g[s][1];
    } else {
        const double c0f = (double)(float)sqrt(1.0 / (4.0 * DSM_PI) / 1.0);
        ja = n - 1;
        va = make_double2(-c0f, 0.0);
    }
    // forward: zero below the first nonzero entry
    for (int j = 0; j < ja; j++) Z[j] = cz();
    double2 zprev = cz();
    for (int j0 = ja; j0 < n; j0 += PSV_L0_G) {
        double2 c1v[PSV_L0_G], c2v[PSV_L0_G];
        for (int u = 0; u < PSV_L0_G; u++)
            if (j0 + u < n) {
                c1v[u] = C1[j0 + u];
                c2v[u] = C2[j0 + u];
            }
        for (int u = 0; u < PSV_L0_G && j0 + u < n; u++) {
            const int j = j0 + u;
            const double2 b = (j == ja) ? va : (j == jb) ? vb : cz();
            const double2 f = (j == 0) ? b : csub(b, cmul(c1v[u], zprev));
            zprev = f;
            Z[j] = cdiv(f, c2v[u]);
        }
    }
    // backward
    double2 x = Z[n - 1];
    for (int j1 = n - 2; j1 >= 0; j1 -= PSV_L0_G) {
        double2 c1v[PSV_L0_G], zv[PSV_L0_G];
        for (int u = 0; u < PSV_L0_G; u++)
            if (j1 - u >= 0) {
                c1v[u] = C1[j1 - u + 1];
                zv[u] = Z[j1 - u];
            }
        for (int u = 0; u < PSV_L0_G && j1 - u >= 0; u++) {
            x = csub(zv[u], cmul(c1v[u], x));
            Z[j1 - u] = x;
        }
    }
}

// coefficient rows at l = 0 (V = 0; F1 vanishes at l = 0) of receiver r;
// r == r_n: the SYNN rows
__host__ __device__ inline void psv_l0_rows(const PsvL0Args &a, int t, int r)
{
    const int sl = a.slot_list[t];
    const int N0 = a.nn0;
    const double2 *z = a.work + (size_t)t * 5 * N0 + (size_t)2 * N0;
    const int R = a.r_n, LC = a.lcap;
    if (r < R) {
        double2 *E = a.E + (size_t)sl * PSV_NQ * R * LC;
        const int j0 = a.jsta0[r] - 1;
        const double *w0 = a.w0 + 3 * r, *w1 = a.w1 + 3 * r;
        PsvVals v[PSV_NSOL];
        for (int s = 0; s < PSV_NSOL; s++) v[s].U = v[s].V = v[s].Ud = v[s].Vd = cz();
        const int map[3] = {0, 1, 4};
        for (int s = 0; s < 3; s++) {
            const double2 *zs = z + (size_t)s * N0 + j0;
            v[map[s]].U = cadd(cadd(cscale(zs[0], w0[0]), cscale(zs[1], w0[1])), cscale(zs[2], w0[2]));
            v[map[s]].Ud = cadd(cadd(cscale(zs[0], w1[0]), cscale(zs[1], w1[1])), cscale(zs[2], w1[2]));
        }
        if (a.eb.Eb)
            psv_rows(v, a.fluid[r] != 0, a.fluidK[(size_t)sl * R + r], 0.0, a.irf[r], false, false,
                     [&](int q, double2 val) { a.eb.put(q, r, sl, 0, val); });
        else
            psv_rows(v, a.fluid[r] != 0, a.fluidK[(size_t)sl * R + r], 0.0, a.irf[r], false, false,
                     [&](int q, double2 val) { E[(size_t)(q * R + r) * LC] = val; });
    } else if (r == R) {
        double2 us[PSV_NSOL], vs[PSV_NSOL], es[PSV_NQS];
        for (int s = 0; s < PSV_NSOL; s++) us[s] = vs[s] = cz();
        us[0] = z[N0 - 1];
        us[1] = z[(size_t)N0 + N0 - 1];
        psv_synn_rows(us, vs, 0.0, false, false, es);
        if (a.Es) {
            double2 *Es = a.Es + (size_t)sl * PSV_NQS * LC;
            for (int q = 0; q < PSV_NQS; q++) Es[(size_t)q * LC] = es[q];
        } else {
            for (int q = 0; q < PSV_NQS; q++) a.esb.put(q, 0, sl, 0, es[q]);
        }
    }
}

// all phases for frequency t in one thread (host build)
__host__ __device__ inline void psv_l0_serial(const PsvL0Args &a, int t)
{
    for (int i = 0; i < a.nn0; i++) psv_l0_assemble(a, t, i);
    psv_l0_factor(a, t);
    for (int s = 0; s < 3; s++) psv_l0_subst(a, t, s);
    for (int r = 0; r <= a.r_n; r++) psv_l0_rows(a, t, r);
}

}  // namespace dsm
