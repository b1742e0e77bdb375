// sh_lane.h -- per-lane bodies of the SH (omega,l) kernels, shared by the
// HIP kernels (sh_kernels.hip) and a host build used only to test the
// formulation on the CPU (hostcheck.cpp).
#pragma once

#include "dev_common.h"

namespace dsm {

// ---- SH band solve --------------------------------------------------------
// One lane per (omega,l) system: assemble A = a0 + l(l+1) a2 (tridiagonal,
// complex symmetric), factor LDL^T without pivoting in the reference's order
// (src/SGTsh/dclib.f90:1-137, m<3 branch), and run the substitutions of the
// reference's right-hand sides sharing that factorization: the 10 nonzero
// moment-tensor sources (calg2 excitation at the two source nodes) and the
// surface force (one solve; the four back-propagated right-hand sides are
// exact multiples of it by +-1, +-i).  Zero right-hand sides of the
// reference's imt/m loops are skipped (their solutions are exactly zero).
// Writes the 11 solutions at the receiver stencil nodes.
struct ShSrcConst {
    double t4[4], h1[4], h2[4], h3[4], h4[4];          // layer integrals at the source node
    double gt[8], gh1[8], gh2[8], gh3[8], gh4[8];      // two-element source grid integrals
    double r0, mu0;
};

struct ShSolveArgs {
    const double2 *band;   // [slot][4][nn]: a0 offdiag, a0 diag, a2 offdiag, a2 diag
    const int *slot_list;  // slots of this launch
    int nlb;               // 64-wide l blocks per slot
    int l0, lend;          // l range [l0, lend]
    int nn, k;             // band rows, first source node (1-based)
    int nneed;
    const int *need;       // stencil nodes, descending, 1-based
    double2 *scratch;      // [2 ceil(nn/SH_SEG) + 20][nthreads]: segment checkpoints, the sources at rows k, k+1
    long sstride;          // nthreads of the launch
    double2 *xn;           // [slot][nneed][SH_NSOL][lcap]
    int lcap;
    const double2 *om;     // [slot][2]: comega2, coef(spn)
    const ShSrcConst *src;
    const int *jstart;     // mixed forms, depth cut: first row solved per block of the launch (null: row 1)
    unsigned long long *small_pivots;  // count of pivots with |d| <= 1e-14 (may be null)
    const int *lstop;      // [slot] degrees l >= lstop are not solved (the slot's l cut-off; null: none)
};
// Depth cut of the SH systems (mixed forms): per block of 64 degrees the
// first row whose decay below the source, integrated from the source down,
// stays under dcut (sh_depth_kernel).
struct ShDepthArgs {
    const double *rrow, *vrow;  // [nn] node radius, smaller shear speed
    const double *omega;        // [slot] real angular frequency
    const int *slot_list;
    int nslots, nlb, l0, k;     // k: first source node (1-based)
    double dcut, dcap;
    int *jstart;                // [active slot][nlb]
    const int *lstop;           // [slot] blocks from lstop on are skipped (null: none)
};
constexpr int SH_NSOL = 11;  // 10 sources + surface force
constexpr int SH_SEG = 8;    // rows per checkpoint segment of the SH solve

// ---- SH coefficients --------------------------------------------------------
// One lane per (omega,l): source excitation (calga/calg2,
// src/SGTsh/calmat.f90:329-479), back-propagated forces (calbveczero,
// trial.f90:120-157), receiver interpolation (others.f90:843-891) and the
// strain maps (udertotsgtSH/udertorsgtSH/utosynnSH) folded into the rows of
// the synthesis contraction.  Also the l cut-off amplitude |g0(nn)|.
struct ShCoefArgs {
    const double2 *xn;     // from sh_solve: [slot][nneed][SH_NSOL][lcap]
    const int *slot_list;
    int nlb, l0, lend, lcap;
    int nneed, r_n;
    const int *rec_slot;   // [r][3] -> index into need
    const double *w0, *w1, *rinv;  // [r][3], [r][3], [r]
    double2 *E;            // [slot][36][r_n][lcap]
    double2 *Es;           // [slot][5][lcap]
    double *amp;           // [slot][lcap]
    const int *lstop;      // [slot] degrees l >= lstop are skipped (null: none)
};
// The l cut-off of SH replayed on the device, one chunk [l0, lend] of
// degrees at a time, its state carried per slot from chunk to chunk
// (sh_cut_kernel).  lstop stays INT_MAX until the slot's cut-off is known,
// then holds it (maxlmax + 1 when the series runs to lmax).
struct ShCutArgs {
    const double *amp;   // [slot][lcap] |g0(nn)| from the coefficient kernel
    const int *lsuf;     // [slot]
    int lcap, l0, lend, maxlmax;
    double ratl;
    double *carry;       // [slot] running maximum of the amplitudes before l0
    int *count;          // [slot] count of small entries before l0
    int *lstop;          // [slot]
};


#ifndef DSM_PI_DEFINED
#define DSM_PI_DEFINED
static constexpr double DSM_PI = 3.1415926535897932;
#endif

struct Exc {
    double2 a, b;  // g0(k), g0(k+1)
};

__host__ __device__ inline void sh_excitation(int l, double2 comega2, double2 coef, const ShSrcConst &S,
                                              Exc ex[10], double2 &gam)
{
    // The two-element source problem has an element of 0.01 layer (calspo
    // clamps the source) and its solution is formed by cancellation, so this
    // function keeps the reference's rounding: no FMA contraction, the
    // reference's operation order and complex quotient.
#pragma clang fp contract(off)
    // calga (calmat.f90:329-353) for the source layer and the 2-element grid
    const double xl2m2 = (double)l * (double)(l + 1) - 2.0;
    double2 aa[4], ga[8];
    for (int i = 0; i < 4; i++) {
        const double h = S.h1[i] - S.h2[i] + S.h3[i] + xl2m2 * S.h4[i];
        aa[i] = csub(cscale(comega2, S.t4[i]), cscale(coef, h));
    }
    for (int i = 0; i < 8; i++) {
        const double h = S.gh1[i] - S.gh2[i] + S.gh3[i] + xl2m2 * S.gh4[i];
        ga[i] = csub(cscale(comega2, S.gt[i]), cscale(coef, h));
    }
    // overlap (calmat.f90:386-423) of 2 layers -> 3x3 tridiagonal, LDL^T
    double2 d1 = ga[0], d2 = cadd(ga[3], ga[4]), d3 = ga[7];
    double2 o2 = ga[1], o3 = ga[5];
    const double2 dr1 = crecip(d1);
    const double2 L2 = cmul(o2, dr1);
    const double2 dr2 = crecip(csub(d2, cmul(o2, L2)));
    const double2 L3 = cmul(o3, dr2);
    const double2 dr3 = crecip(csub(d3, cmul(o3, L3)));
    const double r0 = S.r0, mu0 = S.mu0;
    // calg2 (calmat.f90:427-479) for the nonzero (m, imt):
    //  ex[0..3]: m=-1 imt4, m=-1 imt5, m=+1 imt4, m=+1 imt5
    //  ex[4..9]: m=-2 imt2, imt3, imt6, m=+2 imt2, imt3, imt6
    const double b11 = sqrt((double)(2 * l + 1) / (16.0 * DSM_PI));
    const double2 den = cscale(coef, r0 * r0 * mu0);
    for (int e = 0; e < 4; e++) {
        const double sgn = (e < 2) ? -1.0 : 1.0;
        const int imt = (e & 1) ? 5 : 4;
        // dd = b1*(sgn*mt13 + i mt12)/(r0^2 mu0 coef)
        const double2 num = (imt == 5) ? make_double2(b11 * sgn, 0.0) : make_double2(0.0, b11);
        const double2 dd = cdiv(num, den);
        double2 c1 = cz();
        double2 c2 = cneg(cmul(dd, cadd(ga[4], ga[5])));
        double2 c3 = cneg(cmul(dd, cadd(ga[6], ga[7])));
        // forward / diagonal / backward (dclib.f90:93-104)
        double2 z1 = c1, z2 = csub(c2, cmul(L2, z1)), z3 = csub(c3, cmul(L3, z2));
        z1 = cmul(z1, dr1);
        z2 = cmul(z2, dr2);
        z3 = cmul(z3, dr3);
        const double2 x3 = z3, x2 = csub(z2, cmul(L3, x3)), x1 = csub(z1, cmul(L2, x2));
        const double2 y3 = cadd(x3, dd);
        ex[e].a = cadd(cmul(aa[0], x1), cmul(aa[1], y3));
        ex[e].b = cadd(cmul(aa[2], x1), cmul(aa[3], y3));
    }
    const double b12 = (l >= 2) ? sqrt((double)(2 * l + 1) * (double)(l - 1) * (double)(l + 2) / (64.0 * DSM_PI)) : 0.0;
    for (int e = 0; e < 6; e++) {
        const double sgn = (e < 3) ? -1.0 : 1.0;
        const int w = e % 3;  // 0: imt2 (mt22), 1: imt3 (mt33), 2: imt6 (mt23)
        const double2 v = (w == 2) ? make_double2(2.0, 0.0) : make_double2(0.0, (w == 0 ? sgn : -sgn));
        const double2 c2 = cscale(v, b12 / r0);
        double2 z1 = cz(), z2 = c2, z3 = cneg(cmul(L3, z2));
        z1 = cmul(z1, dr1);
        z2 = cmul(z2, dr2);
        z3 = cmul(z3, dr3);
        const double2 x3 = z3, x2 = csub(z2, cmul(L3, x3)), x1 = csub(z1, cmul(L2, x2));
        ex[4 + e].a = cadd(cmul(aa[0], x1), cmul(aa[1], x3));
        ex[4 + e].b = cadd(cmul(aa[2], x1), cmul(aa[3], x3));
    }
    // back-propagated surface forces (calbveczero, trial.f90:120-157):
    // g0(nn) = -conj(rvec(icomp,m))/lsq for m = +-1 with
    // rvec(2,+-1) = i hv, rvec(3,+1) = -hv, rvec(3,-1) = hv, hv = l(l+1) c/2:
    // icomp 2 -> i*gam (both m), icomp 3 -> gam (m=+1), -gam (m=-1).
    const double xl2 = (double)l * (double)(l + 1);
    const double lsq = sqrt(xl2);
    const double c1 = sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / ((double)l * (double)(l + 1)));
    const double hv = xl2 * c1 / 2.0;
    gam = cdiv(make_double2(hv, 0.0), make_double2(lsq, 0.0));
    if (l < 2)
        for (int e = 4; e < 10; e++) ex[e].a = ex[e].b = cz();
}

// ---------------------------------------------------------------------------
// Band solve.  Lane layout: 64 consecutive l of one frequency per
// wavefront, so the band pieces a0/a2 of that frequency are read at the
// same address by every lane (scalar loads) and each lane assembles its own
// A(omega,l) on the fly.  Per-lane factor rows and forward-substitution
// values go to a scratch buffer laid out [row][lane] (coalesced), read back
// once by the backward pass.
// ---------------------------------------------------------------------------
// Returns the lane's number of pivots with |d| <= 1e-14 (or not finite),
// the test of dclisb0 (dclib.f90:37-86), which reports and continues.
__host__ __device__ inline int sh_solve_lane(const ShSolveArgs &a, int b, int tx)
{
    const int sl = a.slot_list[b / a.nlb];
    const int l = a.l0 + (b % a.nlb) * 64 + tx;
    if (l > a.lend || (a.lstop && l >= a.lstop[sl])) return 0;
    const long gid = (long)b * 64 + tx;
    const double l2 = (double)l * (double)(l + 1);
    const int nn = a.nn, k = a.k;
    const double2 *a0o = a.band + (size_t)(sl * 4 + 0) * nn;
    const double2 *a0d = a.band + (size_t)(sl * 4 + 1) * nn;
    const double2 *a2o = a.band + (size_t)(sl * 4 + 2) * nn;
    const double2 *a2d = a.band + (size_t)(sl * 4 + 3) * nn;
    // scratch: per segment of SH_SEG rows from jlo, a checkpoint (dr and c
    // of the row below the segment); then the scaled forward values of the
    // 10 sources at rows k and k+1 (z dr, 20 entries).  Above row k+1 the
    // sources' right-hand sides are zero and their forward values follow one
    // recurrence, z(j) = -L_j z(j-1), so z_s(j) = c_j z_s(k+1) with the
    // common c_j = prod_{i=k+2..j} (-L_i).  The backward pass recomputes each
    // segment's L_j and c_j dr_j from its checkpoint (the forward pass's
    // operations on the same values, so the same bits) instead of reading a
    // value per row and source back.
    double2 *Ck = a.scratch;
    double2 *Zs = a.scratch + (size_t)2 * ((nn + SH_SEG - 1) / SH_SEG) * a.sstride;
    const long ss = a.sstride;
    double2 *xo = a.xn + (size_t)sl * a.nneed * SH_NSOL * a.lcap + l;
    if (l == 0) {  // no m with 0 < |m| <= l: nothing is solved (SGTsh-MPI.f90:453)
        for (int p = 0; p < a.nneed; p++)
            for (int q = 0; q < SH_NSOL; q++) xo[(size_t)(SH_NSOL * p + q) * a.lcap] = cz();
        return 0;
    }
    Exc ex[10];
    double2 gam;
    sh_excitation(l, a.om[2 * sl], a.om[2 * sl + 1], *a.src, ex, gam);

    // pass 1: factor (dclib.f90:36-61) + forward substitution of the sources
    // (dclib.f90:93-97), whose right-hand sides sit at rows k, k+1; with the
    // depth cut the band starts at row jlo (rows below carry a decay of more
    // than e^-dcut from the source)
    const int jlo = a.jstart ? a.jstart[b] : 1;
    // one factor step at row j from dr(j-1): L_j, the pivot and dr(j)
    auto step = [&](int j, double2 drm, double2 &Lj, double2 &pv) {
        const double2 dg = caxpy(a0d[j - 1], l2, a2d[j - 1]);
        if (j == jlo) {
            pv = dg;
            Lj = cz();
        } else {
            const double2 xx = caxpy(a0o[j - 1], l2, a2o[j - 1]);
            Lj = cmul(xx, drm);
            pv = csub(dg, cmul(xx, Lj));
        }
        return crecip(pv);
    };
    double2 drp = cz();
    double2 z[10];                       // z_s(k), then z_s(k+1)
    double2 c = make_double2(1.0, 0.0);  // c_j (1 up to row k+1)
    int small = 0;
    for (int j = jlo; j <= nn; j++) {
        if ((j - jlo) % SH_SEG == 0) {
            const int t = (j - jlo) / SH_SEG;
            Ck[(size_t)(2 * t) * ss + gid] = drp;
            Ck[(size_t)(2 * t + 1) * ss + gid] = c;
        }
        double2 Lj, pv;
        const double2 drj = step(j, drp, Lj, pv);
        small += !(pv.x * pv.x + pv.y * pv.y > 1e-28);
        if (j == k || j == k + 1) {
            for (int s = 0; s < 10; s++) {
                z[s] = (j == k) ? ex[s].a : csub(ex[s].b, cmul(Lj, z[s]));
                Zs[(size_t)(10 * (j - k) + s) * ss + gid] = cmul(z[s], drj);
            }
        } else if (j > k + 1) {
            c = cneg(cmul(Lj, c));
        }
        drp = drj;
    }
    // pass 2: backward substitution (dclib.f90:98-104) segment by segment
    // from the top, recording the stencil nodes
    double2 x[10], xs = cz();
    double2 Lup = cz();  // L of the row above the segment
    int p = 0;
    bool fin = false;
    for (int t = (nn - jlo) / SH_SEG; t >= 0 && !fin; t--) {
        const int j0 = jlo + SH_SEG * t;
        double2 L[SH_SEG], q[SH_SEG];
        {
            double2 dr = Ck[(size_t)(2 * t) * ss + gid], cc = Ck[(size_t)(2 * t + 1) * ss + gid];
#pragma unroll
            for (int u = 0; u < SH_SEG; u++) {
                L[u] = q[u] = cz();
                const int j = j0 + u;
                if (j <= nn) {
                    double2 pv;
                    const double2 drj = step(j, dr, L[u], pv);
                    if (j > k + 1) {
                        cc = cneg(cmul(L[u], cc));
                        q[u] = cmul(cc, drj);  // c_j dr_j
                    }
                    dr = drj;
                }
            }
        }
#pragma unroll
        for (int u = SH_SEG - 1; u >= 0; u--) {
            const int j = j0 + u;
            if (j > nn || fin) continue;
            // scaled forward value of source s at row j (z[] holds z_s(k+1))
            auto zz = [&](int s) {
                return (j > k + 1) ? cmul(q[u], z[s]) : (j >= k) ? Zs[(size_t)(10 * (j - k) + s) * ss + gid] : cz();
            };
            if (j == nn) {
                for (int s = 0; s < 10; s++) x[s] = zz(s);
                xs = cmul(gam, drp);  // surface force: z(nn) = gam * dr(nn)
            } else {
                const double2 Ln = (u == SH_SEG - 1) ? Lup : L[u + 1];  // L_{j+1}
                for (int s = 0; s < 10; s++) x[s] = csub(zz(s), cmul(Ln, x[s]));
                xs = cneg(cmul(Ln, xs));
            }
            if (p < a.nneed && a.need[p] == j) {
                for (int s = 0; s < 10; s++) xo[(size_t)(SH_NSOL * p + s) * a.lcap] = x[s];
                xo[(size_t)(SH_NSOL * p + 10) * a.lcap] = xs;
                if (++p == a.nneed) fin = true;  // (need is descending: no row below is recorded)
            }
        }
        Lup = L[0];
    }
    for (; p < a.nneed; p++)  // stencil rows below the cut
        for (int q = 0; q < SH_NSOL; q++) xo[(size_t)(SH_NSOL * p + q) * a.lcap] = cz();
    return small;
}

// ---------------------------------------------------------------------------
// Coefficient rows.  Row q of the contraction for receiver r is
//   E[q](r) = sum over right-hand sides s of (strain-map weight) *
//             D_s(ij),  uder_s(i,j) = D_s(ij) * T_b(ij)(l, theta),
// with the real theta functions T_b (Y0..Y3, Z23, Z33, see legendre.hip).
// For signed order m = sigma*mu:  p = 1 (m>0) or (-1)^mu (m<0),
// kappa = i*sigma*mu,  G = g/lsq,  Gd = g'/lsq,  rho = 1/r:
//   D12 = -rho G p kappa (Y0)   D21 = Gd p kappa (Y0)
//   D13 =  rho G p (Y1)         D31 = -Gd p (Y1)
//   D22 =  rho G p kappa (Y2)   D32 = rho G p (Y3)
//   D23 =  rho G p (Z23)        D33 = rho G p kappa (Z33)
// (locallyCartesianDerivatives others.f90:336-367 with calu 829-838 and
//  calbvecphi0 trial.f90:1-81, where b(-m) = (-1)^m conj(b(m))).
// Row layout (q): tsgt t0..t9 = slots 7,12..20 of udertotsgtSH
//   (others.f90:249-332), rsgt r0..r4 = slots 6..10 of udertorsgtSH
//   (others.f90:196-245); see sh_synth_jobs() in engine.cpp for the (mu,b)
//   of every row.
// ---------------------------------------------------------------------------
__host__ __device__ inline void sh_coef_lane(const ShCoefArgs &a, int b, int tx)
{
    const int sl = a.slot_list[b / a.nlb];
    const int l = a.l0 + (b % a.nlb) * 64 + tx;
    if (l > a.lend || (a.lstop && l >= a.lstop[sl])) return;
    const int R = a.r_n, LC = a.lcap;
    double2 *E = a.E + (size_t)sl * 36 * R * LC + l;
    double2 *Es = a.Es + (size_t)sl * 5 * LC + l;
    if (l == 0) {  // no m satisfies 0 < |m| <= l
        for (int q = 0; q < 36; q++)
            for (int r = 0; r < R; r++) E[(size_t)(q * R + r) * LC] = cz();
        for (int q = 0; q < 5; q++) Es[(size_t)q * LC] = cz();
        a.amp[(size_t)sl * LC + l] = 0.0;
        return;
    }
    const bool mu2 = (l >= 2);
    const double lsq = sqrt((double)l * (double)(l + 1));
    const double il = 1.0 / lsq;
    const double2 *X = a.xn + (size_t)sl * a.nneed * SH_NSOL * LC + l;

    // surface node (need[0] == nn): l cut-off amplitude and SYNN rows
    {
        double2 gs[10];
        for (int e = 0; e < 10; e++) gs[e] = cscale(X[(size_t)e * LC], il);
        // calamp on g0(nn) of imt=3 at m=+-2 (SGTsh-MPI.f90:476-478)
        a.amp[(size_t)sl * LC + l] = mu2 ? cabs2(X[(size_t)5 * LC]) : 0.0;
        // utosynnSH (others.f90:132-186): u2 = G p kappa Y0, u3 = -G p Y1
        // mu=2: m=-2 (p=+1, kappa=-2i) -> gs[4..6]; m=+2 (p=+1, kappa=+2i) -> gs[7..9]
        // mu=1: m=-1 (p=-1, kappa=-i) -> gs[0..1]; m=+1 (p=+1, kappa=+i) -> gs[2..3]
        double2 s0 = cz(), s1 = cz(), s2 = cz(), s3 = cz(), s4 = cz();
        for (int sg = 0; sg < 2; sg++) {
            const double sig = sg ? 1.0 : -1.0;
            const double2 G2 = gs[4 + 3 * sg], G3 = gs[5 + 3 * sg], G6 = gs[6 + 3 * sg];
            // u2 = G * kappa (p=1), kappa = i*sig*2
            s0 = csub(s0, cmul_i(cadd(G2, G3), 0.5 * sig * 2.0));
            s1 = cadd(s1, cmul_i(csub(G3, G2), 0.5 * sig * 2.0));
            s2 = cadd(s2, G6);  // -u3 = G p Y1
            const double p1 = sg ? 1.0 : -1.0;
            const double2 G4 = gs[0 + 2 * sg], G5 = gs[1 + 2 * sg];
            s3 = cadd(s3, cmul_i(G4, p1 * sig * 1.0));  // u2 = G p kappa
            s4 = csub(s4, cscale(G5, p1));               // u3 = -G p
        }
        Es[0] = s0;
        Es[(size_t)1 * LC] = s1;
        Es[(size_t)2 * LC] = s2;
        Es[(size_t)3 * LC] = s3;
        Es[(size_t)4 * LC] = s4;
    }

    for (int r = 0; r < R; r++) {
        const int p0 = a.rec_slot[3 * r], p1 = a.rec_slot[3 * r + 1], p2 = a.rec_slot[3 * r + 2];
        const double w00 = a.w0[3 * r], w01 = a.w0[3 * r + 1], w02 = a.w0[3 * r + 2];
        const double w10 = a.w1[3 * r], w11 = a.w1[3 * r + 1], w12 = a.w1[3 * r + 2];
        const double rho = a.rinv[r];
        double2 v[SH_NSOL], d[SH_NSOL];
        for (int bb = 0; bb < SH_NSOL; bb++) {
            const double2 g0 = X[(size_t)(SH_NSOL * p0 + bb) * LC], g1 = X[(size_t)(SH_NSOL * p1 + bb) * LC],
                          g2 = X[(size_t)(SH_NSOL * p2 + bb) * LC];
            v[bb] = cadd(cadd(cscale(g0, w00), cscale(g1, w01)), cscale(g2, w02));
            d[bb] = cadd(cadd(cscale(g0, w10), cscale(g1, w11)), cscale(g2, w12));
        }
        double2 e[36];
        for (int q = 0; q < 36; q++) e[q] = cz();
        // source right-hand sides
        for (int s = 0; s < 10; s++) {
            const double2 G = cscale(v[s], il);
            const double2 Gd = cscale(d[s], il);
            const int mu = (s < 4) ? 1 : 2;
            const double sig = (s < 4) ? ((s < 2) ? -1.0 : 1.0) : ((s < 7) ? -1.0 : 1.0);
            const double p = (sig > 0) ? 1.0 : ((mu & 1) ? -1.0 : 1.0);
            const double km = sig * (double)mu;                 // kappa = i*km
            const double2 rG = cscale(G, rho * p);              // rho G p
            const double2 D12p21 = cmul_i(csub(cscale(Gd, p), rG), km);  // p kappa (Gd - rho G)
            const double2 D13p31 = csub(rG, cscale(Gd, p));              // p (rho G - Gd)
            const double2 D22 = cmul_i(rG, km), D33 = D22;
            const double2 D23 = rG, D32 = rG;
            int imt;
            if (s < 4) imt = (s & 1) ? 5 : 4;
            else imt = ((s - 4) % 3 == 0) ? 2 : (((s - 4) % 3 == 1) ? 3 : 6);
            switch (imt) {
            case 4:
                e[0] = caxpy(e[0], -0.5, D22);
                e[1] = caxpy(e[1], -0.5, D33);
                e[4] = caxpy(e[4], -0.25, D12p21);
                e[6] = caxpy(e[6], 0.25, D12p21);
                e[8] = caxpy(e[8], 0.25, D22);
                e[9] = caxpy(e[9], -0.25, D33);
                e[12] = caxpy(e[12], -0.25, D22);
                e[13] = caxpy(e[13], 0.25, D33);
                break;
            case 5:
                e[5] = caxpy(e[5], 0.25, D13p31);
                e[7] = caxpy(e[7], 0.25, D13p31);
                e[10] = caxpy(e[10], -0.25, D23);
                e[11] = caxpy(e[11], -0.25, D32);
                e[14] = caxpy(e[14], -0.25, D23);
                e[15] = caxpy(e[15], -0.25, D32);
                break;
            case 2:
                e[2] = caxpy(e[2], -0.25, D22);
                e[3] = caxpy(e[3], -0.25, D33);
                e[16] = caxpy(e[16], -0.125, D12p21);
                e[18] = caxpy(e[18], 0.125, D12p21);
                e[20] = caxpy(e[20], 0.125, D22);
                e[21] = caxpy(e[21], -0.125, D33);
                e[24] = caxpy(e[24], 0.125, D22);
                e[25] = caxpy(e[25], -0.125, D33);
                break;
            case 3:
                e[2] = caxpy(e[2], 0.25, D22);
                e[3] = caxpy(e[3], 0.25, D33);
                e[16] = caxpy(e[16], -0.125, D12p21);
                e[18] = caxpy(e[18], 0.125, D12p21);
                e[20] = caxpy(e[20], -0.125, D22);
                e[21] = caxpy(e[21], 0.125, D33);
                e[24] = caxpy(e[24], -0.125, D22);
                e[25] = caxpy(e[25], 0.125, D33);
                break;
            default:  // 6
                e[17] = caxpy(e[17], 0.25, D13p31);
                e[19] = caxpy(e[19], 0.25, D13p31);
                e[22] = caxpy(e[22], -0.25, D23);
                e[23] = caxpy(e[23], -0.25, D32);
                e[26] = caxpy(e[26], 0.25, D23);
                e[27] = caxpy(e[27], 0.25, D32);
                break;
            }
        }
        // back-propagated right-hand sides (mu = 1): icomp 2, 3 at m = -1, +1
        for (int s = 0; s < 4; s++) {
            const int icomp = (s < 2) ? 2 : 3;
            const double sig = (s & 1) ? 1.0 : -1.0;
            // solution of the surface force times i (icomp 2) or +-1 (icomp 3)
            double2 G = cscale(v[10], il), Gd = cscale(d[10], il);
            if (icomp == 2) {
                G = cmul_i(G, 1.0);
                Gd = cmul_i(Gd, 1.0);
            } else if (sig < 0) {
                G = cneg(G);
                Gd = cneg(Gd);
            }
            const double p = (sig > 0) ? 1.0 : -1.0;
            const double km = sig;
            const double2 rG = cscale(G, rho * p);
            if (icomp == 2) {
                const double2 D12p21 = cmul_i(csub(cscale(Gd, p), rG), km);
                const double2 D22 = cmul_i(rG, km);
                e[28] = caxpy(e[28], -0.5, D22);
                e[29] = caxpy(e[29], -0.5, D22);
                e[30] = caxpy(e[30], -0.5, D22);
                e[31] = caxpy(e[31], 0.5, D22);
                e[34] = caxpy(e[34], -0.5, D12p21);
            } else {
                const double2 D13p31 = csub(rG, cscale(Gd, p));
                e[32] = caxpy(e[32], -0.5, rG);
                e[33] = caxpy(e[33], -0.5, rG);
                e[35] = caxpy(e[35], 0.5, D13p31);
            }
        }
        for (int q = 0; q < 36; q++) E[(size_t)(q * R + r) * LC] = e[q];
    }
}

}  // namespace dsm

