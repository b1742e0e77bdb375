/*
 * psv_oracle.c -- CPU restatement of the SGTpsv engine (spheroidal P-SV,
 * solid and fluid zones) of IPGP/DSM-Kernel, FP64, loop by loop.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle and the timed CPU baseline.
 * The product library never links or calls this file.
 *
 * As in sh_oracle.c, every p_<name> routine restates the reference routine
 * <name> with the reference's argument list, and the driver orc_psv_omega()
 * calls them through a function table that orc_psv_load_ref() can fill with
 * the reference's routines compiled from the src/SGTpsv sources
 * (oracle/_ref/libref_psv.so).
 */
#include "dsm_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;
static const double PI = 3.1415926535897932;
#define F1(a, i) (a)[(i) - 1]

/* ======================================================================= */
/* src/SGTpsv/calmat.f90: element integrals                                  */
/* ======================================================================= */
static void p_pmulti(int n, const double *a, int m, const double *b, int l, double *c)
{
    for (int i = 1; i <= l; i++) c[i - 1] = 0.0;
    for (int i = 1; i <= n; i++)
        for (int j = 1; j <= m; j++) c[i + j - 2] = c[i + j - 2] + a[i - 1] * b[j - 1];
}

static void p_polint(int n, const double *p, double x1, double x2, double *pint)
{
    double a[6], b[6];
    a[0] = 1.0;
    b[0] = 1.0;
    for (int i = 2; i <= n; i++) {
        a[i - 1] = a[i - 2] * x1;
        b[i - 1] = b[i - 2] * x2;
    }
    double dx = x2 - x1;
    *pint = 0.0;
    for (int i = 1; i <= n; i++) {
        double xx = 0.0;
        for (int j = 1; j <= i; j++) xx = xx + a[j - 1] * b[i - j] / (double)i;
        *pint = *pint + p[i - 1] * dx * xx;
    }
}

/* pinteg calmat.f90:86-124 */
static void p_pinteg(int *snp, int n, const double *p, double from, double to, int vnp, const double *vra,
                     const double *con, double *pint)
{
    double q[2], pq[6], psint, x1 = from, x2;
    if (*snp >= vnp) *snp = 1;
    *pint = 0.0;
    for (;;) {
        if (vra[*snp - 1] <= x1 && vra[*snp] > x1) {
            x2 = fmin(to, vra[*snp]);
        } else {
            *snp = *snp + 1;
            if (*snp == vnp) *snp = 1;
            continue;
        }
        q[1] = (con[*snp] - con[*snp - 1]) / (vra[*snp] - vra[*snp - 1]);
        q[0] = -q[1] * vra[*snp - 1] + con[*snp - 1];
        p_pmulti(n, p, 2, q, n + 1, pq);
        p_polint(n + 1, pq, x1, x2, &psint);
        *pint = *pint + psint;
        if (x2 == to) break;
        x1 = x2;
    }
}

/* calmatc calmat.f90:1-62 (P-SV form: no symmetrisation) */
static void p_calmatc(const int *nlayer, const int *vnp, const double *vra, const double *con, const int *rpow,
                      const int *w1dn, const int *w2dn, const double *ra, double *m)
{
    int snp = 1;
    double a[2][2], b[2][2], c[5];
    for (int i = 1; i <= *nlayer; i++) {
        double rh = ra[i] - ra[i - 1];
        if (*w1dn == 0) {
            a[0][1] = -1.0 / rh;
            a[0][0] = ra[i] / rh;
            a[1][1] = 1.0 / rh;
            a[1][0] = -ra[i - 1] / rh;
        } else {
            a[0][1] = 0.0;
            a[0][0] = -1.0 / rh;
            a[1][1] = 0.0;
            a[1][0] = 1.0 / rh;
        }
        if (*w2dn == 0) {
            b[0][1] = -1.0 / rh;
            b[0][0] = ra[i] / rh;
            b[1][1] = 1.0 / rh;
            b[1][0] = -ra[i - 1] / rh;
        } else {
            b[0][1] = 0.0;
            b[0][0] = -1.0 / rh;
            b[1][1] = 0.0;
            b[1][0] = 1.0 / rh;
        }
        for (int j = 1; j <= 2; j++)
            for (int k = 1; k <= 2; k++) {
                for (int kk = 0; kk < 5; kk++) c[kk] = 0.0;
                p_pmulti(2, a[j - 1], 2, b[k - 1], 3, c);
                for (int l = 3; l >= 1; l--) {
                    c[l + *rpow - 1] = c[l - 1];
                    if (*rpow > 0) c[l - 1] = 0.0;
                }
                int nn = 4 * (i - 1) + 2 * (j - 1) + k;
                p_pinteg(&snp, 5, c, ra[i - 1], ra[i], *vnp, vra, con, &m[nn - 1]);
            }
    }
}

static void p_caltl(const int *nlayer, const int *vnp, const double *vra, const double *rho, const double *ra,
                    double *tl)
{
    int snp = 1;
    double c[3] = {0.0, 0.0, 1.0};
    for (int i = 1; i <= *nlayer; i++) {
        double from = ra[i - 1], to = (ra[i - 1] + ra[i]) / 2.0;
        int nn = 4 * (i - 1);
        p_pinteg(&snp, 3, c, from, to, *vnp, vra, rho, &tl[nn]);
        tl[nn + 1] = 0.0;
        tl[nn + 2] = 0.0;
        p_pinteg(&snp, 3, c, to, ra[i], *vnp, vra, rho, &tl[nn + 3]);
    }
}

static void p_calhl(const int *nlayer, const int *vnp, const double *vra, const double *mu, const double *ra,
                    double *hl)
{
    int snp = 1;
    double c[1] = {1.0};
    for (int i = 1; i <= *nlayer; i++) {
        double from = ra[i - 1], to = (ra[i - 1] + ra[i]) / 2.0;
        int nn = 4 * (i - 1);
        p_pinteg(&snp, 1, c, from, to, *vnp, vra, mu, &hl[nn]);
        hl[nn + 1] = 0.0;
        hl[nn + 2] = 0.0;
        p_pinteg(&snp, 1, c, to, ra[i], *vnp, vra, mu, &hl[nn + 3]);
    }
}

static void p_calt(const int *nlayer, const double *tl, const double *tc, double *t)
{
    for (int i = 1; i <= 4 * *nlayer; i++) t[i - 1] = (tl[i - 1] + tc[i - 1]) / 2.0;
}

/* calh5 calmat.f90:246-277 */
static void p_calh5(const int *nlayer, const double *vra, const double *con, const double *ra, double *h5)
{
    for (int i = 0; i < 4 * *nlayer; i++) h5[i] = 0.0;
    int itmp = 0;
    for (;;) {
        itmp++;
        if (F1(vra, itmp) == F1(ra, 1)) {
            if (F1(vra, itmp + 1) == F1(vra, itmp)) itmp++;
            break;
        }
    }
    itmp--;
    for (int i = 1; i <= *nlayer; i++) {
        itmp++;
        F1(h5, 4 * i - 3) = -3.0 / 8.0 * F1(con, itmp) * F1(ra, i) - 1.0 / 8.0 * F1(con, itmp + 1) * F1(ra, i + 1);
        F1(h5, 4 * i - 2) = -F1(h5, 4 * i - 3);
        F1(h5, 4 * i - 1) = -1.0 / 8.0 * F1(con, itmp) * F1(ra, i) - 3.0 / 8.0 * F1(con, itmp + 1) * F1(ra, i + 1);
        F1(h5, 4 * i) = -F1(h5, 4 * i - 1);
    }
}

/* calhm1 calmat.f90:282-318; hm1(-1:2,*) */
#define HM1(a, i, j) (a)[((i) + 1) + 4 * ((j) - 1)]
static void p_calhm1(const int *nlayer, const double *vra, const double *con, const double *ra, double *hm1)
{
    int itmp = 0, N = *nlayer;
    for (;;) {
        itmp++;
        if (F1(vra, itmp) == F1(ra, 1)) {
            if (F1(vra, itmp + 1) == F1(vra, itmp)) itmp++;
            break;
        }
    }
    HM1(hm1, 0, 1) = HM1(hm1, 0, 1) - 7.0 / 12.0 * F1(con, itmp) * F1(ra, 1);
    HM1(hm1, 1, 1) = HM1(hm1, 1, 1) + 8.0 / 12.0 * F1(con, itmp) * F1(ra, 1);
    HM1(hm1, 2, 1) = HM1(hm1, 2, 1) - 1.0 / 12.0 * F1(con, itmp) * F1(ra, 1);
    for (int i = 2; i <= N - 1; i++) {
        itmp++;
        HM1(hm1, -1, i) = HM1(hm1, -1, i) - 5.0 / 12.0 * F1(con, itmp) * F1(ra, i);
        HM1(hm1, 0, i) = HM1(hm1, 0, i) - 3.0 / 12.0 * F1(con, itmp) * F1(ra, i);
        HM1(hm1, 1, i) = HM1(hm1, 1, i) + 9.0 / 12.0 * F1(con, itmp) * F1(ra, i);
        HM1(hm1, 2, i) = HM1(hm1, 2, i) - 1.0 / 12.0 * F1(con, itmp) * F1(ra, i);
    }
    itmp++;
    HM1(hm1, -1, N) = HM1(hm1, -1, N) - 5.0 / 12.0 * F1(con, itmp) * F1(ra, N);
    HM1(hm1, 0, N) = HM1(hm1, 0, N) - 3.0 / 12.0 * F1(con, itmp) * F1(ra, N);
    HM1(hm1, 1, N) = HM1(hm1, 1, N) + 8.0 / 12.0 * F1(con, itmp) * F1(ra, N);
    itmp++;
    HM1(hm1, -1, N + 1) = HM1(hm1, -1, N + 1) - 5.0 / 12.0 * F1(con, itmp) * F1(ra, N + 1);
    HM1(hm1, 0, N + 1) = HM1(hm1, 0, N + 1) + 5.0 / 12.0 * F1(con, itmp) * F1(ra, N + 1);
}

/* calhm2 calmat.f90:322-358; hm2(-2:1,*) */
#define HM2(a, i, j) (a)[((i) + 2) + 4 * ((j) - 1)]
static void p_calhm2(const int *nlayer, const double *vra, const double *con, const double *ra, double *hm2)
{
    int itmp = 0, N = *nlayer;
    for (;;) {
        itmp++;
        if (F1(vra, itmp) == F1(ra, 1)) {
            if (F1(vra, itmp + 1) == F1(vra, itmp)) itmp++;
            break;
        }
    }
    HM2(hm2, 0, 1) = HM2(hm2, 0, 1) - 5.0 / 12.0 * F1(con, itmp) * F1(ra, 1);
    HM2(hm2, 1, 1) = HM2(hm2, 1, 1) + 5.0 / 12.0 * F1(con, itmp) * F1(ra, 1);
    itmp++;
    HM2(hm2, -1, 2) = HM2(hm2, -1, 2) - 8.0 / 12.0 * F1(con, itmp) * F1(ra, 2);
    HM2(hm2, 0, 2) = HM2(hm2, 0, 2) + 3.0 / 12.0 * F1(con, itmp) * F1(ra, 2);
    HM2(hm2, 1, 2) = HM2(hm2, 1, 2) + 5.0 / 12.0 * F1(con, itmp) * F1(ra, 2);
    for (int i = 3; i <= N; i++) {
        itmp++;
        HM2(hm2, -2, i) = HM2(hm2, -2, i) + 1.0 / 12.0 * F1(con, itmp) * F1(ra, i);
        HM2(hm2, -1, i) = HM2(hm2, -1, i) - 9.0 / 12.0 * F1(con, itmp) * F1(ra, i);
        HM2(hm2, 0, i) = HM2(hm2, 0, i) + 3.0 / 12.0 * F1(con, itmp) * F1(ra, i);
        HM2(hm2, 1, i) = HM2(hm2, 1, i) + 5.0 / 12.0 * F1(con, itmp) * F1(ra, i);
    }
    itmp++;
    HM2(hm2, -2, N + 1) = HM2(hm2, -2, N + 1) + 1.0 / 12.0 * F1(con, itmp) * F1(ra, N + 1);
    HM2(hm2, -1, N + 1) = HM2(hm2, -1, N + 1) - 8.0 / 12.0 * F1(con, itmp) * F1(ra, N + 1);
    HM2(hm2, 0, N + 1) = HM2(hm2, 0, N + 1) + 7.0 / 12.0 * F1(con, itmp) * F1(ra, N + 1);
}

/* mtrnp calmat.f90:363-386 */
static void p_mtrnp(const int *nlayer, const double *a1, double *a2)
{
    int nn = 4 * *nlayer;
    for (int i = 1; i <= nn - 3; i += 4) F1(a2, i) = F1(a1, i);
    for (int i = 2; i <= nn - 2; i += 4) F1(a2, i) = F1(a1, i + 1);
    for (int i = 3; i <= nn - 1; i += 4) F1(a2, i) = F1(a1, i - 1);
    for (int i = 4; i <= nn; i += 4) F1(a2, i) = F1(a1, i);
}

/* mtrnp2 calmat.f90:390-404: hm1(-p:q,*) -> hm2(-q:p,*) */
static void p_mtrnp2(const int *nlayer, const int *p, const int *q, const double *hm1, double *hm2)
{
    const int w = *p + *q + 1;
    for (int i = -*q; i <= *p; i++)
        for (int j = 1; j <= *nlayer + 1; j++) {
            int m = -i, n = i + j;
            if (n >= 1 && n <= *nlayer + 1) hm2[(i + *q) + w * (j - 1)] = hm1[(m + *p) + w * (n - 1)];
        }
}

/* submat others.f90:852-864 */
static void p_submat(const int *nlayer, const double *ha, const double *hb, double *h)
{
    for (int i = 0; i < 4 * *nlayer; i++) h[i] = ha[i] - hb[i];
}

/* ======================================================================= */
/* per-omega assembly: calmat.f90:409-809                                    */
/* ======================================================================= */
static void p_calcoef(const int *nzone, const double *omega, const double *qmu, const double *qkappa, zc *coef1,
                      zc *coef2, zc *coef)
{
    for (int i = 0; i < *nzone; i++) {
        if (qmu[i] <= 0.0) coef1[i] = 1.0;
        else {
            double aa = (*omega == 0.0) ? 1.0 : 1.0 + log(*omega / (2.0 * PI)) / (PI * qmu[i]);
            double bb = 1.0 / (2.0 * qmu[i]);
            zc c = CMPLX(aa, bb);
            coef1[i] = c * c;
        }
        if (qkappa[i] <= 0.0) {
            coef2[i] = 1.0;
            coef[i] = 1.0 / coef2[i];
        } else {
            double aa = (*omega == 0.0) ? 1.0 : 1.0 + log(*omega / (2.0 * PI)) / (PI * qkappa[i]);
            double bb = 1.0 / (2.0 * qkappa[i]);
            zc c = CMPLX(aa, bb);
            coef2[i] = c * c;
            coef[i] = 1.0 / coef2[i];
        }
    }
}

/* cala0 calmat.f90:452-494 */
static void p_cala0(const int *nlayer, const double *omega, const double *omegai, const double *t, const double *h1x,
                    const double *h1y, const double *h1z, const double *h2L, const double *h2N, const double *h3ax,
                    const double *h3ay, const double *h3az, const double *h4aL, const double *h4aN,
                    const double *h5ax, const double *h5ay, const double *h5az, const double *h6aL,
                    const double *h6aN, const double *h7x, const double *h7y, const double *h7z, const double *h8L,
                    const double *h8N, const zc *coef1, const zc *coef2, zc *a0)
{
    (void)h1y;
    (void)h1z;
    (void)h2L;
    (void)h3ax;
    (void)h3az;
    (void)h5ax;
    (void)h5az;
    (void)h8L;
    for (int i = 0; i < 16 * *nlayer; i++) a0[i] = 0.0;
    zc w = CMPLX(*omega, -*omegai);
    zc comega2 = w * w;
    for (int i = 1; i <= 16 * *nlayer - 3; i += 4) {
        int nnn = (i + 3) / 4;
        if (nnn % 4 == 3) continue;
        double tt = F1(t, nnn);
        zc hh0 = *coef2 * (4.0 * F1(h1x, nnn)) + *coef1 * (16.0 / 3.0 * F1(h2N, nnn)) +
                 *coef1 * (-4.0 * F1(h2N, nnn)) + *coef2 * (2.0 * (F1(h3ay, nnn) + F1(h5ay, nnn))) -
                 *coef1 * (4.0 / 3.0 * (F1(h4aN, nnn) + F1(h6aN, nnn))) +
                 *coef2 * (3.0 * F1(h7z, nnn) - 2.0 * F1(h7y, nnn)) + *coef1 * (4.0 / 3.0 * F1(h8N, nnn));
        F1(a0, i) = comega2 * tt - hh0;
    }
    for (int i = 4; i <= 16 * *nlayer; i += 4) {
        int nnn = i / 4;
        if (nnn % 4 == 3) continue;
        double tt = F1(t, nnn);
        zc hh0 = *coef1 * F1(h2L, nnn) + *coef1 * (-2.0 * F1(h2N, nnn)) -
                 *coef1 * (F1(h4aL, nnn) + F1(h6aL, nnn)) + *coef1 * F1(h8L, nnn);
        F1(a0, i) = comega2 * tt - hh0;
    }
}

/* cala1 calmat.f90:499-537 */
static void p_cala1(const int *nlayer, const double *h1x, const double *h1y, const double *h1z, const double *h2L,
                    const double *h2N, const double *h3x, const double *h3y, const double *h3z, const double *h4L,
                    const double *h4N, const double *h5x, const double *h5y, const double *h5z, const double *h6L,
                    const double *h6N, const zc *coef1, const zc *coef2, zc *a1)
{
    (void)h1y;
    (void)h1z;
    (void)h3x;
    (void)h3z;
    (void)h5x;
    (void)h5z;
    for (int i = 0; i < 16 * *nlayer; i++) a1[i] = 0.0;
    for (int i = 2; i <= 16 * *nlayer - 2; i += 4) {
        int nnn = (i + 2) / 4;
        if (nnn % 4 == 3) continue;
        zc hh1 = -*coef2 * (2.0 * F1(h1x, nnn)) - *coef1 * (8.0 / 3.0 * F1(h2N, nnn)) - *coef1 * F1(h2L, nnn) -
                 *coef1 * (-2.0 * F1(h2N, nnn)) - *coef2 * F1(h3y, nnn) - *coef1 * (-2.0 / 3.0 * F1(h4N, nnn)) +
                 *coef1 * F1(h6L, nnn);
        F1(a1, i) = -hh1;
    }
    for (int i = 3; i <= 16 * *nlayer - 1; i += 4) {
        int nnn = (i + 1) / 4;
        if (nnn % 4 != 2) continue;
        zc hh1 = -*coef2 * (2.0 * F1(h1x, nnn)) - *coef1 * (8.0 / 3.0 * F1(h2N, nnn)) -
                 *coef1 * (-2.0 * F1(h2N, nnn)) - *coef1 * F1(h2L, nnn) + *coef1 * F1(h4L, nnn) -
                 *coef2 * F1(h5y, nnn) + *coef1 * (2.0 / 3.0 * F1(h6N, nnn));
        F1(a1, i) = -hh1;
    }
}

/* cala2 calmat.f90:543-567 */
static void p_cala2(const int *nlayer, const double *h1x, const double *h1y, const double *h1z, const double *h2L,
                    const double *h2N, const zc *coef1, const zc *coef2, zc *a2)
{
    (void)h1y;
    (void)h1z;
    for (int i = 0; i < 16 * *nlayer; i++) a2[i] = 0.0;
    for (int i = 1; i <= 16 * *nlayer - 3; i += 4) {
        int nnn = (i + 3) / 4;
        if (nnn % 4 == 3) continue;
        zc hh2 = *coef1 * F1(h2L, nnn);
        F1(a2, i) = -hh2;
    }
    for (int i = 4; i <= 16 * *nlayer; i += 4) {
        int nnn = i / 4;
        if (nnn % 4 == 3) continue;
        zc hh2 = *coef2 * F1(h1x, nnn) + *coef1 * (4.0 / 3.0 * F1(h2N, nnn));
        F1(a2, i) = -hh2;
    }
}

/* calb0 / calb2 calmat.f90:572-622 */
static void p_calb0(const int *nlayer, const double *omega, const double *omegai, const double *p1, const double *p3,
                    const zc *coef, zc *b0)
{
    for (int i = 0; i < 4 * *nlayer; i++) b0[i] = 0.0;
    zc w = CMPLX(*omega, -*omegai);
    zc comega2 = w * w;
    for (int i = 1; i <= 4 * *nlayer - 3; i += 4) F1(b0, i) = -F1(p1, i) / comega2 + *coef * F1(p3, i);
    for (int i = 2; i <= 4 * *nlayer - 2; i += 4) F1(b0, i) = -F1(p1, i) / comega2 + *coef * F1(p3, i);
    for (int i = 4; i <= 4 * *nlayer; i += 4) F1(b0, i) = -F1(p1, i) / comega2 + *coef * F1(p3, i);
}

static void p_calb2(const int *nlayer, const double *omega, const double *omegai, const double *p2, const zc *coef,
                    zc *b2)
{
    (void)coef;
    for (int i = 0; i < 4 * *nlayer; i++) b2[i] = 0.0;
    zc w = CMPLX(*omega, -*omegai);
    zc comega2 = w * w;
    for (int i = 1; i <= 4 * *nlayer - 3; i += 4) F1(b2, i) = -F1(p2, i) / comega2;
    for (int i = 2; i <= 4 * *nlayer - 2; i += 4) F1(b2, i) = -F1(p2, i) / comega2;
    for (int i = 4; i <= 4 * *nlayer; i += 4) F1(b2, i) = -F1(p2, i) / comega2;
}

/* calhml calmat.f90:629-664; a1(4,*) */
static void p_calhml(const int *nlayer, const zc *coef1, const zc *coef2, const double *h3mx, const double *h3my,
                     const double *h3mz, const double *h5mx, const double *h5my, const double *h5mz,
                     const double *h4m1L, const double *h4m1N, const double *h4m2L, const double *h4m2N,
                     const double *h6m1L, const double *h6m1N, const double *h6m2L, const double *h6m2N, zc *a1)
{
    (void)h3mx;
    (void)h3mz;
    (void)h5mx;
    (void)h5mz;
    (void)h4m1N;
    (void)h4m2L;
    (void)h6m1L;
    (void)h6m2N;
#define A4(i, j) a1[((i) - 1) + 4 * ((j) - 1)]
    for (int j = 2; j <= 2 * (*nlayer + 1); j += 2)
        for (int i = 1; i <= 3; i += 2) {
            if (i == 1 && j == 2) continue;
            int m = (-i + 3) / 2, n = (i + j - 3) / 2;
            A4(i, j) = A4(i, j) - (-*coef2 * HM2(h3my, m, n) + *coef1 * (2.0 / 3.0 * HM2(h4m2N, m, n)) +
                                   *coef1 * HM2(h6m2L, m, n));
        }
    for (int j = 3; j <= 2 * *nlayer + 1; j += 2)
        for (int i = 1; i <= 3; i += 2) {
            if (i == 1 && j == 3) continue;
            int m = (-i + 5) / 2, n = (i + j - 4) / 2;
            A4(i, j) = A4(i, j) - (*coef1 * HM1(h4m1L, m, n) - *coef2 * HM1(h5my, m, n) +
                                   *coef1 * (2.0 / 3.0 * HM1(h6m1N, m, n)));
        }
#undef A4
}

/* overlapa calmat.f90:670-714 */
static void p_overlapa(const int *nlayer, const zc *a, zc *a2)
{
    int nsize = 2 * (*nlayer + 1), msize = 16 * *nlayer;
#define A4(i, j) a2[((i) - 1) + 4 * ((j) - 1)]
    for (int i = 0; i < 4 * nsize; i++) a2[i] = 0.0;
    A4(4, 1) = F1(a, 1);
    A4(3, 2) = F1(a, 2);
    A4(4, 2) = F1(a, 4);
    A4(4, nsize - 1) = F1(a, msize - 3);
    A4(3, nsize) = F1(a, msize - 2);
    A4(4, nsize) = F1(a, msize);
    for (int i = 3; i <= nsize - 1; i += 2) {
        int n1 = 16 * (i - 3) / 2 + 5;
        A4(2, i) = F1(a, n1);
        n1 = n1 + 2;
        A4(3, i) = F1(a, n1);
        if (i != nsize - 1) {
            n1 = n1 + 6;
            int n2 = n1 + 4;
            A4(4, i) = F1(a, n1) + F1(a, n2);
        }
    }
    for (int i = 4; i <= nsize; i += 2) {
        int n1 = 16 * (i - 4) / 2 + 6;
        A4(1, i) = F1(a, n1);
        n1 = n1 + 2;
        A4(2, i) = F1(a, n1);
        if (i != nsize) {
            n1 = n1 + 6;
            int n2 = n1 + 4;
            A4(3, i) = F1(a, n1) + F1(a, n2);
            n1 = n1 + 2;
            n2 = n1 + 4;
            A4(4, i) = F1(a, n1) + F1(a, n2);
        }
    }
#undef A4
}

/* overlapb calmat.f90:720-742 */
static void p_overlapb(const int *nlayer, const zc *b, zc *b2)
{
#define B4(i, j) b2[((i) - 1) + 4 * ((j) - 1)]
    for (int i = 0; i < 4 * (*nlayer + 1); i++) b2[i] = 0.0;
    B4(4, 1) = F1(b, 1);
    for (int i = 2; i <= *nlayer; i++) {
        int n1 = 4 * (i - 1), n2 = n1 + 1;
        B4(4, i) = F1(b, n1) + F1(b, n2);
    }
    B4(4, *nlayer + 1) = F1(b, 4 * *nlayer);
    for (int i = 2; i <= *nlayer + 1; i++) {
        int n1 = 4 * (i - 2) + 2;
        B4(3, i) = F1(b, n1);
    }
#undef B4
}

/* cala calmat.f90:747-789 */
static void p_cala(const int *maxnzone, const int *ndc, const int *iphase, const int *nlayer, const int *kkdr,
                   const int *kdr, const int *ksp, const double *l2, const double *lsq, const int *nn, const zc *a0,
                   const zc *a1, const zc *a2, zc *a)
{
    (void)maxnzone;
#define A4(x, i, j) (x)[((i) - 1) + 4 * ((j) - 1)]
    for (int i = 0; i < 4 * *nn; i++) a[i] = 0.0;
    for (int j = 1; j <= *ndc + 1; j++) {
        if (F1(iphase, j) == 1) {
            int i1 = F1(kkdr, j), i2 = F1(kkdr, j) + 2 * F1(nlayer, j) + 1;
            for (int jj = i1; jj <= i2; jj++) {
                int mtmp = *kdr + F1(ksp, j) + jj - F1(kkdr, j);
                for (int ii = 2; ii <= 4; ii += 2)
                    A4(a, ii, jj) = A4(a, ii, jj) + CMPLX(creal(A4(a0, ii, mtmp)) + *l2 * creal(A4(a2, ii, mtmp)),
                                                          cimag(A4(a0, ii, mtmp)) + *l2 * cimag(A4(a2, ii, mtmp)));
                for (int ii = 1; ii <= 3; ii += 2)
                    A4(a, ii, jj) =
                        A4(a, ii, jj) + CMPLX(*lsq * creal(A4(a1, ii, mtmp)), *lsq * cimag(A4(a1, ii, mtmp)));
            }
        } else {
            int i1 = F1(kkdr, j), i2 = F1(kkdr, j) + F1(nlayer, j);
            for (int jj = i1; jj <= i2; jj++) {
                int mtmp = *kdr + F1(ksp, j) + jj - F1(kkdr, j);
                for (int ii = 3; ii <= 4; ii++)
                    A4(a, ii, jj) = A4(a, ii, jj) + CMPLX(creal(A4(a0, ii, mtmp)) + *l2 * creal(A4(a2, ii, mtmp)),
                                                          cimag(A4(a0, ii, mtmp)) + *l2 * cimag(A4(a2, ii, mtmp)));
            }
        }
    }
#undef A4
}

/* calbc calmat.f90:795-809 */
static void p_calbc(const int *maxnzone, const int *ndc, const double *rdc, const int *iphase, const int *kkdr, zc *a)
{
    (void)maxnzone;
    for (int i = 1; i <= *ndc; i++) {
        if (F1(iphase, i) == 1 && F1(iphase, i + 1) == 2)
            a[1 + 4 * (F1(kkdr, i + 1) - 1)] = F1(rdc, i) * F1(rdc, i);
        if (F1(iphase, i) == 2 && F1(iphase, i + 1) == 1)
            a[2 + 4 * (F1(kkdr, i + 1) - 1)] = -(F1(rdc, i) * F1(rdc, i));
    }
}

/* ======================================================================= */
/* source: calmat.f90:815-1243, rk3.f90, glu2.f90                            */
/* ======================================================================= */
#define ANUM(x, i, j, k) (x)[((i) - 1) + 4 * ((j) - 1) + 16 * ((k) - 1)]
static void p_calabnum(const double *omega, const double *omegai, const double *rmax, const double *rrho,
                       const double *vpv, const double *vph, const double *vsv, const double *vsh, const double *eta,
                       const double *ra, const double *r0, const zc *coef1, const zc *coef2, zc *anum, zc *bnum)
{
    for (int k = 0; k < 160; k++) anum[k] = bnum[k] = 0.0;
    for (int i = 1; i <= 10; i++) {
        double r;
        if (i <= 5) r = ra[0] + (double)(i - 1) / 4.0 * (*r0 - ra[0]);
        else r = ra[1] - (double)(10 - i) / 4.0 * (ra[1] - *r0);
        double trho = 0, tvpv = 0, tvph = 0, tvsv = 0, tvsh = 0, teta = 0, coef = 1.0;
        for (int j = 1; j <= 4; j++) {
            if (j == 1) coef = 1.0;
            else coef = coef * (r / *rmax);
            trho = trho + rrho[j - 1] * coef;
            tvpv = tvpv + vpv[j - 1] * coef;
            tvph = tvph + vph[j - 1] * coef;
            tvsv = tvsv + vsv[j - 1] * coef;
            tvsh = tvsh + vsh[j - 1] * coef;
            teta = teta + eta[j - 1] * coef;
        }
        double tecA = trho * tvph * tvph, tecC = trho * tvpv * tvpv, tecL = trho * tvsv * tvsv;
        double tecN = trho * tvsh * tvsh, tmu = trho * tvsv * tvsv;
        double tAkappa = tecA - tmu * 4.0 / 3.0, tCkappa = tecC - tmu * 4.0 / 3.0;
        zc xAkappa = tAkappa * *coef2, xCkappa = tCkappa * *coef2;
        zc xrho = trho, xmu = tmu * *coef1, xecL = tecL * *coef1, xecN = tecN * *coef1;
        zc xecA = xAkappa + xmu * (4.0 / 3.0), xecC = xCkappa + xmu * (4.0 / 3.0);
        zc xecF = teta * (xecA - 2.0 * xecL);
        zc xFdC = xecF / xecC;
        zc xAmF2dC = xecA - xecF * xFdC;
        double r2 = r * r;
        zc w = CMPLX(*omega, -*omegai);
        zc comega2 = w * w;
        ANUM(anum, 1, 1, i) = -(2.0 / r) * xFdC;
        ANUM(anum, 1, 2, i) = 1.0 / xecC;
        ANUM(anum, 2, 1, i) = -xrho * comega2 + (4.0 / r2) * (xAmF2dC - xecN);
        ANUM(anum, 2, 2, i) = (2.0 / r) * (xFdC - 1.0);
        ANUM(anum, 3, 1, i) = -(1.0 / r);
        ANUM(anum, 3, 3, i) = 1.0 / r;
        ANUM(anum, 3, 4, i) = 1.0 / xecL;
        ANUM(anum, 4, 1, i) = -(2.0 / r2) * (xAmF2dC - xecN);
        ANUM(anum, 4, 2, i) = -(1.0 / r) * xFdC;
        ANUM(anum, 4, 3, i) = -xrho * comega2 - (2.0 / r2) * xecN;
        ANUM(anum, 4, 4, i) = -(3.0 / r);
        ANUM(bnum, 1, 3, i) = (1.0 / r) * xFdC;
        ANUM(bnum, 2, 3, i) = -(2.0 / r2) * (xAmF2dC - xecN);
        ANUM(bnum, 2, 4, i) = 1.0 / r;
        ANUM(bnum, 4, 3, i) = xAmF2dC / r2;
    }
}

/* calya + rk3/rkstep + eqmotion1 (blank COMMON a,b,l2,itmp): four
 * fundamental solutions by one RK4 step on each half of [ra1,r0] and
 * [r0,ra2]; the coefficient matrices at the RK points are the 10 of
 * calabnum, picked by the call counter of eqmotion1. */
typedef struct {
    zc a[64], b[64];
    double l2;
    int itmp;
} rk_common_t;

static void p_eqmotion1(rk_common_t *cm, const zc *y, zc *f)
{
    cm->itmp++;
    int mtmp = 16 * (cm->itmp - 1);
    zc c[16];
    for (int j = 1; j <= 4; j++)
        for (int i = 1; i <= 4; i++) {
            c[(i - 1) + 4 * (j - 1)] = cm->a[mtmp] + CMPLX(cm->l2 * creal(cm->b[mtmp]), cm->l2 * cimag(cm->b[mtmp]));
            mtmp++;
        }
    for (int i = 0; i < 4; i++) {
        f[i] = 0.0;
        for (int j = 0; j < 4; j++) f[i] = f[i] + c[i + 4 * j] * y[j];
    }
}

static void p_rkstep(rk_common_t *cm, double h, const zc *y0, zc *yn)
{
    const double c1 = 1.0 / 6.0, c2 = 1.0 / 3.0, c3 = c2, c4 = c1, b2 = 0.5, b3 = b2;
    zc ak[4], w[4];
    p_eqmotion1(cm, y0, ak);
    for (int i = 0; i < 4; i++) yn[i] = y0[i] + CMPLX(h * c1, 0.0) * ak[i];
    for (int i = 0; i < 4; i++) w[i] = y0[i] + CMPLX(h * b2, 0.0) * ak[i];
    p_eqmotion1(cm, w, ak);
    for (int i = 0; i < 4; i++) yn[i] = yn[i] + CMPLX(h * c2, 0.0) * ak[i];
    for (int i = 0; i < 4; i++) w[i] = y0[i] + CMPLX(h * b3, 0.0) * ak[i];
    p_eqmotion1(cm, w, ak);
    for (int i = 0; i < 4; i++) yn[i] = yn[i] + CMPLX(h * c3, 0.0) * ak[i];
    for (int i = 0; i < 4; i++) w[i] = y0[i] + CMPLX(h, 0.0) * ak[i];
    p_eqmotion1(cm, w, ak);
    for (int i = 0; i < 4; i++) yn[i] = yn[i] + CMPLX(h * c4, 0.0) * ak[i];
}

/* rk3 with n = 1 step; x0 is set to xe on return */
static void p_rk3(rk_common_t *cm, double *x0, double xe, zc *y0)
{
    zc yn[4];
    double h = (xe - *x0) / 1.0;
    p_rkstep(cm, h, y0, yn);
    *x0 = *x0 + h;
    for (int j = 0; j < 4; j++) y0[j] = yn[j];
    *x0 = xe;
}

static void p_calya(const zc *aa, const zc *bb, const double *l2, const double *ra, const double *r0, zc *ya, zc *yb,
                    zc *yc, zc *yd)
{
    rk_common_t cm;
    cm.l2 = *l2;
    ya[0] = 0.0, ya[1] = 1.0, ya[2] = 0.0, ya[3] = 0.0;
    yb[0] = 0.0, yb[1] = 0.0, yb[2] = 0.0, yb[3] = 1.0;
    double dr = (*r0 - ra[0]) / 2.0;
    if (dr > 0.0) {
        double xs = ra[0], xe;
        for (int k = 1; k <= 2; k++) {
            int ktmp1 = 32 * (k - 1), ktmp2 = ktmp1 - 16;
            for (int i = 1; i <= 64; i++) {
                if (i <= 32) {
                    cm.a[i - 1] = aa[i + ktmp1 - 1];
                    cm.b[i - 1] = bb[i + ktmp1 - 1];
                } else {
                    cm.a[i - 1] = aa[i + ktmp2 - 1];
                    cm.b[i - 1] = bb[i + ktmp2 - 1];
                }
            }
            cm.itmp = 0;
            xe = xs + dr;
            p_rk3(&cm, &xs, xe, ya);
            cm.itmp = 0;
            xs = xe - dr;
            p_rk3(&cm, &xs, xe, yb);
        }
    }
    yc[0] = 0.0, yc[1] = 1.0, yc[2] = 0.0, yc[3] = 0.0;
    yd[0] = 0.0, yd[1] = 0.0, yd[2] = 0.0, yd[3] = 1.0;
    dr = (ra[1] - *r0) / 2.0;
    if (dr > 0.0) {
        double xs = ra[1], xe;
        for (int k = 1; k <= 2; k++) {
            int ktmp1 = 144 - 32 * (k - 1), ktmp2 = ktmp1 - 32, ktmp3 = ktmp1 - 48, ktmp4 = ktmp1 - 80;
            for (int i = 1; i <= 64; i++) {
                int o = (i <= 16) ? ktmp1 : (i <= 32) ? ktmp2 : (i <= 48) ? ktmp3 : ktmp4;
                cm.a[i - 1] = aa[i + o - 1];
                cm.b[i - 1] = bb[i + o - 1];
            }
            cm.itmp = 0;
            xe = xs - dr;
            p_rk3(&cm, &xs, xe, yc);
            cm.itmp = 0;
            xs = xe + dr;
            p_rk3(&cm, &xs, xe, yd);
        }
    }
}

/* glu (P-SV copy) glu2.f90:2-137 */
static void p_glu(zc *a, int n, int n1, zc *b, double *eps, zc *wk, int *ip, int *ier)
{
#define A_(i, j) a[((i) - 1) + n1 * ((j) - 1)]
    if (*eps < 0.0) *eps = 3.52e-15;
    for (int i = 1; i <= n; i++) wk[i - 1] = cabs(A_(i, 1));
    for (int j = 2; j <= n; j++)
        for (int i = 1; i <= n; i++) wk[i - 1] = fmax(cabs(wk[i - 1]), cabs(A_(i, j)));
    for (int i = 1; i <= n; i++)
        if (cabs(wk[i - 1]) < *eps) {
            *ier = 2;
            return;
        }
    *ier = 0;
    for (int k = 1; k <= n; k++) {
        zc amax = cabs(A_(k, k));
        int ipk = k;
        for (int i = k + 1; i <= n; i++) {
            zc aik = cabs(A_(i, k));
            if (cabs(aik) > cabs(amax)) {
                ipk = i;
                amax = aik;
            }
        }
        ip[k - 1] = ipk;
        if (cabs(amax) > *eps) {
            if (ipk != k) {
                zc w = A_(ipk, k);
                A_(ipk, k) = A_(k, k);
                A_(k, k) = w;
            }
            for (int i = k + 1; i <= n; i++) {
                A_(i, k) = -A_(i, k) / A_(k, k);
                wk[i - 1] = A_(i, k);
            }
            for (int j = k + 1; j <= n; j++) {
                if (ipk != k) {
                    zc w = A_(ipk, j);
                    A_(ipk, j) = A_(k, j);
                    A_(k, j) = w;
                }
                zc t = A_(k, j);
                for (int i = k + 1; i <= n; i++) A_(i, j) = A_(i, j) + wk[i - 1] * t;
            }
        } else {
            *ier = 1;
            ip[k - 1] = k;
            for (int i = k + 1; i <= n; i++) A_(i, k) = 0.0;
            return;
        }
    }
    for (int k = 1; k <= n; k++) {
        if (ip[k - 1] != k) {
            zc w = b[ip[k - 1] - 1];
            b[ip[k - 1] - 1] = b[k - 1];
            b[k - 1] = w;
        }
        zc t = b[k - 1];
        for (int i = k + 1; i <= n; i++) b[i - 1] = b[i - 1] + A_(i, k) * t;
    }
    b[n - 1] = b[n - 1] / A_(n, n);
    for (int k = n - 1; k >= 1; k--) {
        zc t = b[k];
        for (int i = 1; i <= k; i++) b[i - 1] = b[i - 1] - A_(i, k + 1) * t;
        b[k - 1] = b[k - 1] / A_(k, k);
    }
#undef A_
}

/* calg calmat.f90:1075-1171 (with sab1 1210-1224, sab2 1231-1243) */
static void p_calg(const int *lp, const int *mp, const zc *coef1, const zc *coef2, const double *lsqp,
                   const double *ecC0, const double *ecF0, const double *ecL0, const zc *ya, const zc *yb, const zc *yc,
                   const zc *yd, const double *ra, const double *r0, const double *mt, zc *g)
{
#define MT(i, j) mt[((i) - 1) + 3 * ((j) - 1)]
    int l = *lp, m = *mp;
    double lsq = *lsqp;
    double eps = -1.0, sgn = (m >= 0) ? 1.0 : -1.0;
    double b1 = sqrt((double)(2 * l + 1) / (16.0 * PI)), b2 = 0.0, r03;
    zc dd = 0.0, ee = 0.0, s1 = 0.0, s2 = 0.0, s[4], a[16], b[4], wk[4], xtmp;
    int ip[4], ier;
    if (l != 0) b2 = sqrt((double)(2 * l + 1) * (double)(l - 1) * (double)(l + 2) / (64.0 * PI));
    if (abs(m) == 2) {
        dd = 0.0;
        ee = 0.0;
    }
    if (abs(m) == 1) {
        dd = 0.0;
        ee = CMPLX(-b1 * sgn * MT(1, 2), b1 * MT(1, 3)) / (CMPLX(*r0 * *r0 * *ecL0, 0.0) * *coef1);
    }
    if (abs(m) == 0) {
        dd = CMPLX(2.0 * b1 * MT(1, 1) / (*r0 * *r0), 0.0) /
             (CMPLX(*ecC0 - 4.0 / 3.0 * *ecL0, 0.0) * *coef2 + CMPLX(4.0 / 3.0 * *ecL0, 0.0) * *coef1);
        ee = 0.0;
    }
    if (abs(m) == 0) {
        r03 = *r0 * *r0 * *r0;
        xtmp = ((*ecF0 + 2.0 / 3.0 * *ecL0) * *coef2 - 2.0 / 3.0 * *ecL0 * *coef1) /
               ((*ecC0 - 4.0 / 3.0 * *ecL0) * *coef2 + 4.0 / 3.0 * *ecL0 * *coef1);
        s1 = CMPLX(-2.0 * b1 * (MT(2, 2) + MT(3, 3)) / r03, 0.0) + CMPLX(4.0 * b1 * MT(1, 1) / r03, 0.0) * xtmp;
        s2 = CMPLX(b1 * lsq * (MT(2, 2) + MT(3, 3)) / r03, 0.0) - CMPLX(2.0 * b1 * lsq * MT(1, 1) / r03, 0.0) * xtmp;
    }
    if (abs(m) == 1) {
        s1 = 0.0;
        s2 = 0.0;
    }
    if (abs(m) == 2) {
        r03 = *r0 * *r0 * *r0;
        s1 = 0.0;
        s2 = CMPLX(-b2 * (MT(2, 2) - MT(3, 3)) / r03, sgn * 2.0 * b2 * MT(2, 3) / r03);
    }
#undef MT
    s[0] = dd;
    s[1] = s1;
    if (l != 0) {
        s[2] = CMPLX(creal(ee) / lsq, cimag(ee) / lsq);
        s[3] = CMPLX(creal(s2) / lsq, cimag(s2) / lsq);
    } else {
        s[2] = 0.0;
        s[3] = 0.0;
    }
    if (l != 0) {
        for (int i = 0; i < 4; i++) {  /* sab1 */
            a[i + 0] = -ya[i];
            a[i + 4] = -yb[i];
            a[i + 8] = yc[i];
            a[i + 12] = yd[i];
            b[i] = s[i];
        }
        p_glu(a, 4, 4, b, &eps, wk, ip, &ier);
    } else {
        for (int i = 0; i < 2; i++) {  /* sab2 */
            a[i + 0] = -ya[i];
            a[i + 4] = yc[i];
            b[i] = s[i];
        }
        p_glu(a, 2, 4, b, &eps, wk, ip, &ier);
        b[2] = b[1];
        b[1] = 0.0;
        b[3] = 0.0;
    }
    double dtmp[4];
    dtmp[0] = -ra[0] * ra[0];
    dtmp[1] = dtmp[0] * lsq;
    dtmp[2] = ra[1] * ra[1];
    dtmp[3] = dtmp[2] * lsq;
    for (int i = 0; i < 4; i++) {
        xtmp = b[i];
        g[i] = CMPLX(creal(xtmp) * dtmp[i], cimag(xtmp) * dtmp[i]);
    }
}

/* ======================================================================= */
/* band LDL^T: dcsymbdl.f90, dcsymbdl3.f90                                   */
/* ======================================================================= */
static void p_dcsymbdl0(zc *a, const int *mp, const int *np, const int *nnp, double *eps, zc *z, zc *w, int *l,
                        int *li, int *lj, int *ier)
{
    (void)nnp;
    int m = *mp, n = *np;
    *ier = 0;
    int ij = 0;
    for (int i = 1; i <= m; i++) {
        ij++;
        F1(li, ij) = i + 1;
        F1(lj, ij) = 2;
        F1(l, ij) = m * (i - 1) + 1;
        for (int j = i + 1; j <= m; j++) {
            ij++;
            F1(l, ij) = F1(l, ij - 1) + m + 1;
            F1(li, ij) = F1(li, ij - 1) + 1;
            F1(lj, ij) = F1(lj, ij - 1) + 1;
        }
    }
    int mm = (m + 1) * m / 2;
    for (int k = 1; k <= n - m; k++) {
        int nk = (m + 1) * (k - 1) + 1;
        if (cabs(F1(a, nk + m)) < *eps) {
            *ier = 1;
            return;
        }
        zc piv = 1.0 / F1(a, nk + m);
        for (int j = 2; j <= m + 1; j++) {
            F1(z, j) = -F1(a, nk + m * j);
            F1(w, j) = F1(a, nk + m * j) * piv;
            F1(a, nk + m * j) = F1(w, j);
        }
        int kk = nk + m + m;
        for (int i = 1; i <= mm; i++) F1(a, kk + F1(l, i)) = F1(a, kk + F1(l, i)) + F1(w, F1(lj, i)) * F1(z, F1(li, i));
    }
    for (int k = n - m + 1; k <= n - 1; k++) {
        int nk = (m + 1) * (k - 1) + 1, nkk = (m + 1) * k - 1;
        if (cabs(F1(a, nk + m)) < *eps) {
            *ier = 1;
            return;
        }
        zc piv = 1.0 / F1(a, nk + m);
        for (int j = 2; j <= n - k + 1; j++) {
            F1(z, j) = -F1(a, nk + m * j);
            F1(w, j) = F1(a, nk + m * j) * piv;
            F1(a, nk + m * j) = F1(w, j);
            int nki = nkk + m * (j - 1);
            for (int i = 2; i <= j; i++) F1(a, nki + i) = F1(a, nki + i) + F1(w, j) * F1(z, i);
        }
    }
}

static void p_dcsbdlv0(zc *a, zc *b, const int *mp, const int *np, double *eps, zc *z, int *ier)
{
    (void)eps;
    (void)ier;
    int m = *mp, n = *np, mm = m + 1;
#define A_(i, j) a[((i) - 1) + (m + 1) * ((j) - 1)]
    if (mm < 3) {
        z[0] = b[0];
        for (int j = 2; j <= n; j++) F1(z, j) = F1(b, j) - A_(1, j) * F1(z, j - 1);
        for (int j = 1; j <= n; j++) F1(z, j) = F1(z, j) / A_(m + 1, j);
        F1(b, n) = F1(z, n);
        for (int j = 1; j <= n - 1; j++) F1(b, n - j) = F1(z, n - j) - A_(1, n - j + 1) * F1(b, n - j + 1);
    } else {
        z[0] = b[0];
        z[1] = b[1] - A_(m - 1, 2) * z[0];
        for (int j = 3; j <= n; j++) {
            int i1 = (j > mm) ? 1 : mm - j + 1;
            zc sum = 0.0;
            for (int k = i1; k <= mm - 1; k++) sum = sum + A_(k, j) * F1(z, j - mm + k);
            F1(z, j) = F1(b, j) - sum;
        }
        for (int j = 1; j <= n; j++) F1(z, j) = F1(z, j) / A_(m + 1, j);
        F1(b, n) = F1(z, n);
        F1(b, n - 1) = F1(z, n - 1) - A_(mm - 1, n) * F1(z, n);
        for (int j = 3; j <= n; j++) {
            int j1 = n - j + 1, i1 = 1;
            if (j < mm) i1 = mm - j + 1;
            zc sum = 0.0;
            for (int k = i1; k <= mm - 1; k++) sum = sum + A_(k, mm - k + j1) * F1(b, mm - k + j1);
            F1(b, j1) = F1(z, j1) - sum;
        }
    }
#undef A_
}

/* mydcsbdlv3 dcsymbdl3.f90:151-191: dcsbdlv0 with m = 3 */
static void p_mydcsbdlv3(zc *a, zc *b, const int *n, zc *z)
{
    int three = 3, ier;
    double eps = -1.0;
    p_dcsbdlv0(a, b, &three, n, &eps, z, &ier);
}

/* rea2 calmat.f90:1247-1315 (and rea2_back 1320-1389, which fills only d) */
static void p_rea2(const int *nn, const zc *a, const zc *g, zc *c, zc *d, const int *nzone, const int *iphase,
                   const int *kkdr, const int *spn, int *kkdr0, int *nn0, const int *maxnstack, const int *nsta,
                   const int *istazone, const int *iista, int *jsta)
{
    (void)maxnstack;
#define A4(i, j) a[((i) - 1) + 4 * ((j) - 1)]
#define C2(i, j) c[((i) - 1) + 2 * ((j) - 1)]
    int itmp = 0;
    for (int iz = 1; iz <= *nzone; iz++) {
        for (int ista = 1; ista <= *nsta; ista++)
            if (iz == F1(istazone, ista)) F1(jsta, ista) = itmp + iista[3 * (ista - 1)];
        int i1, i2;
        if (F1(iphase, iz) == 1) {
            if (iz == *spn) *kkdr0 = itmp + 1;
            i1 = itmp + 1;
            if (iz != *nzone) i2 = i1 + (F1(kkdr, iz + 1) - F1(kkdr, iz)) / 2 - 1;
            else i2 = i1 + (*nn + 1 - F1(kkdr, iz)) / 2 - 1;
            F1(d, i1) = F1(g, F1(kkdr, iz));
            if (iz != 1) {
                if (F1(iphase, iz - 1) == 2) C2(1, i1) = A4(3, F1(kkdr, iz));
                else C2(1, i1) = A4(2, F1(kkdr, iz));
            }
            C2(2, i1) = A4(4, F1(kkdr, iz));
            for (int i = i1 + 1; i <= i2; i++) {
                int mtmp = F1(kkdr, iz) + 2 * (i - i1);
                F1(d, i) = F1(g, mtmp);
                C2(1, i) = A4(2, mtmp);
                C2(2, i) = A4(4, mtmp);
            }
            itmp = i2;
        } else {
            i1 = itmp + 1;
            i2 = i1 + (F1(kkdr, iz + 1) - F1(kkdr, iz)) - 1;
            F1(d, i1) = F1(g, F1(kkdr, iz));
            if (iz != 1) {
                if (F1(iphase, iz - 1) == 1) C2(1, i1) = A4(2, F1(kkdr, iz));
                else C2(1, i1) = A4(3, F1(kkdr, iz));
            }
            C2(2, i1) = A4(4, F1(kkdr, iz));
            for (int i = i1 + 1; i <= i2; i++) {
                int mtmp = F1(kkdr, iz) + (i - i1);
                F1(d, i) = F1(g, mtmp);
                C2(1, i) = A4(3, mtmp);
                C2(2, i) = A4(4, mtmp);
            }
            itmp = i2;
        }
    }
    *nn0 = itmp;
#undef A4
#undef C2
}

/* rea2_back: the reference writes its matrix into a local array, so only d
 * (and jsta, kkdr0, nn0) reach the caller. */
static void p_rea2_back(const int *nn, const zc *a, const zc *g, zc *ctmp, zc *d, const int *nzone,
                        const int *iphase, const int *kkdr, const int *spn, int *kkdr0, int *nn0,
                        const int *maxnstack, const int *nsta, const int *istazone, const int *iista, int *jsta)
{
    (void)ctmp;
    zc *cl = calloc(2 * (size_t)*nn + 2, sizeof(zc));
    p_rea2(nn, a, g, cl, d, nzone, iphase, kkdr, spn, kkdr0, nn0, maxnstack, nsta, istazone, iista, jsta);
    free(cl);
}

/* ======================================================================= */
/* src/SGTpsv/others.f90                                                     */
/* ======================================================================= */
static void p_calnl(const int *nzone, const double *vs, int *iphase, int *nsl, int *nll)
{
    *nsl = 0;
    *nll = 0;
    for (int i = 1; i <= *nzone; i++) {
        const double *v = &vs[4 * (i - 1)];
        if (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0 && v[3] == 0.0) {
            (*nll)++;
            F1(iphase, i) = 2;
        } else {
            (*nsl)++;
            F1(iphase, i) = 1;
        }
    }
}

/* calgrid (P-SV: Vp in fluid zones) others.f90:374-450 */
static void p_calgrid(const int *nzone, const double *vrmin, const double *vrmax, const double *vp, const double *vs,
                      const double *rmin, const double *rmax, const int *imax, const int *lmin, const double *tlen,
                      double *vmin, double *gridpar, double *dzpar)
{
    for (int iz = 1; iz <= *nzone; iz++) {
        double v[4], coef1 = 1, coef2 = 1, vs1 = 0, vs2 = 0;
        const double *src = (vs[4 * (iz - 1)] == 0.0) ? &vp[4 * (iz - 1)] : &vs[4 * (iz - 1)];
        for (int i = 0; i < 4; i++) v[i] = src[i];
        for (int j = 1; j <= 4; j++) {
            if (j == 1) coef1 = 1.0;
            else coef1 = coef1 * (F1(vrmin, iz) / *rmax);
            if (j == 1) coef2 = 1.0;
            else coef2 = coef2 * (F1(vrmax, iz) / *rmax);
            vs1 = vs1 + v[j - 1] * coef1;
            vs2 = vs2 + v[j - 1] * coef2;
        }
        double rh = F1(vrmax, iz) - F1(vrmin, iz);
        double omega = 2.0 * PI * (double)*imax / *tlen;
        F1(vmin, iz) = (vs1 >= vs2) ? vs2 : vs1;
        double amax = F1(vrmax, iz);
        double gtmp = (omega * omega) / (F1(vmin, iz) * F1(vmin, iz)) -
                      (((double)*lmin + 0.5) * ((double)*lmin + 0.5)) / (amax * amax);
        if (gtmp > 0.0) {
            F1(dzpar, iz) = sqrt(1.0 / gtmp);
            F1(gridpar, iz) = rh / F1(dzpar, iz);
        } else {
            F1(dzpar, iz) = 0.0;
            F1(gridpar, iz) = 0.0;
        }
    }
    double gtmp = 0.0;
    for (int iz = 1; iz <= *nzone; iz++) gtmp = gtmp + F1(gridpar, iz);
    for (int iz = 1; iz <= *nzone; iz++) {
        if (F1(gridpar, iz) > 0.0) F1(gridpar, iz) = F1(gridpar, iz) / gtmp;
        else F1(gridpar, iz) = (F1(vrmax, iz) - F1(vrmin, iz)) / (*rmax - *rmin) * 0.1;
    }
    gtmp = 0.0;
    for (int iz = 1; iz <= *nzone; iz++) gtmp = gtmp + F1(gridpar, iz);
    for (int iz = 1; iz <= *nzone; iz++) F1(gridpar, iz) = F1(gridpar, iz) / gtmp;
}

/* calra_psv others.f90:1292-1340 */
static void p_calra_psv(int *nlayer, int *inlayer, int *jnlayer, int *jnslay, int *jnllay, const double *gridpar,
                        const double *dzpar, const int *nzone, const double *vrmin, const double *vrmax,
                        const int *iphase, const double *rmin, const double *rmax, int *nslay, int *nllay, int *nnl,
                        const double *re)
{
    (void)gridpar;
    (void)rmin;
    (void)rmax;
    *nslay = 0;
    *nllay = 0;
    *inlayer = 0;
    for (int i = 0; i < *nzone; i++) nnl[i] = 0;
    *jnlayer = 0;
    *jnslay = 0;
    *jnllay = 0;
    for (int iz = 1; iz <= *nzone; iz++) {
        double rh = F1(vrmax, iz) - F1(vrmin, iz);
        int ntmp;
        if (F1(dzpar, iz) == 0.0) ntmp = 1;
        else ntmp = (int)(sqrt(3.3 / *re) * rh / F1(dzpar, iz) / 2.0 / PI / 7.0e-1 + 1);
        F1(nnl, iz) = ntmp;
        if (F1(nnl, iz) < 5) F1(nnl, iz) = 5;
        if (F1(iphase, iz) == 1) *nslay = *nslay + F1(nnl, iz);
        if (F1(iphase, iz) == 2) *nllay = *nllay + F1(nnl, iz);
    }
    *inlayer = 0;
    for (int iz = 1; iz <= *nzone; iz++) *inlayer = *inlayer + F1(nnl, iz);
    *nlayer = *inlayer;
    *jnlayer = *jnlayer + *inlayer;
    *jnslay = *jnslay + *nslay;
    *jnllay = *jnllay + *nllay;
}

/* calra2_psv others.f90:1344-1419 (log_solid_liquid is a Fortran logical: 4 bytes) */
static void p_calra2_psv(const int *nlayer, const double *gridpar, const double *dzpar, const int *nzone,
                         const double *vrmin, const double *vrmax, const double *rmin, const double *rmax,
                         const int *nnl, double *ra, const double *re, const int *nsta, const double *rsta,
                         double *rrsta, int *iista, int *log_solid_liquid, const double *rs, int *cista,
                         const int *iphase, int *istazone, int *ciista)
{
    (void)gridpar;
    (void)dzpar;
    (void)rmax;
    (void)re;
    double chikasa = 0.0, ctmp = 6371.0;
    for (int i = 0; i < *nlayer + *nzone + 1; i++) ra[i] = 0.0;
    ra[0] = *rmin;
    *ciista = 0;
    int itmp = 1;
    for (int iz = 1; iz <= *nzone; iz++) {
        double rh = F1(vrmax, iz) - F1(vrmin, iz);
        for (int i = 1; i <= F1(nnl, iz); i++) {
            itmp++;
            F1(ra, itmp) = F1(vrmin, iz) + rh * (double)i / (double)F1(nnl, iz);
        }
    }
    itmp = 1;
    for (int iz = 1; iz <= *nzone; iz++) {
        for (int i = 1; i <= F1(nnl, iz); i++) {
            for (int ista = 1; ista <= *nsta; ista++) {
                if (F1(ra, itmp) < F1(rsta, ista) && F1(rsta, ista) <= F1(ra, itmp + 1)) {
                    F1(istazone, ista) = iz;
                    F1(log_solid_liquid, ista) = (F1(iphase, iz) == 2) ? 0 : 1;
                    int b0 = (i != F1(nnl, iz)) ? 0 : -1;
                    for (int k = 0; k < 3; k++) {
                        rrsta[k + 3 * (ista - 1)] = F1(ra, itmp + b0 + k);
                        iista[k + 3 * (ista - 1)] = i + b0 + k;
                    }
                    if (fabs(*rs - F1(rsta, ista)) < ctmp && fabs(*rs - F1(rsta, ista)) >= chikasa) {
                        *cista = ista;
                        *ciista = itmp;
                        ctmp = fabs(*rs - F1(rsta, ista));
                    }
                }
            }
            itmp++;
        }
    }
    if (*cista == 0) *cista = 1;
}

/* calsp (P-SV) others.f90:455-522 */
static void p_calsp(const int *maxnzone, const int *ndc, const int *nsl, const int *nll, const int *iphase,
                    const int *nlayer, const int *nslay, const int *nllay, int *isp, int *jsp, int *ksp, int *issp,
                    int *ilsp, int *lsp, int *jssp, int *isdr, int *jsdr, int *ildr, int *jdr, int *kdr)
{
    (void)nslay;
    for (int i = 0; i < *maxnzone; i++) isp[i] = jsp[i] = ksp[i] = issp[i] = ilsp[i] = lsp[i] = jssp[i] = 0;
    *isdr = *jsdr = *ildr = *jdr = *kdr = 0;
    isp[0] = jsp[0] = ksp[0] = issp[0] = ilsp[0] = lsp[0] = jssp[0] = 1;
    int isl = 0, ill = 0;
    for (int i = 1; i <= *ndc; i++) {
        F1(isp, i + 1) = F1(isp, i) + F1(nlayer, i);
        if (F1(iphase, i) == 1) {
            F1(jsp, i + 1) = F1(jsp, i) + 16 * F1(nlayer, i);
            F1(ksp, i + 1) = F1(ksp, i) + 2 * (F1(nlayer, i) + 1);
            F1(lsp, i + 1) = F1(lsp, i) + 4 * F1(nlayer, i);
            isl++;
            if (isl != *nsl) {
                F1(issp, isl + 1) = F1(issp, isl) + 4 * F1(nlayer, i);
                F1(jssp, isl + 1) = F1(jssp, isl) + F1(nlayer, i) + 1;
            }
        } else {
            F1(jsp, i + 1) = F1(jsp, i) + 4 * F1(nlayer, i);
            F1(ksp, i + 1) = F1(ksp, i) + (F1(nlayer, i) + 1);
            F1(lsp, i + 1) = F1(lsp, i) + 2 * F1(nlayer, i);
            ill++;
            if (ill != *nll) F1(ilsp, ill + 1) = F1(ilsp, ill) + 4 * F1(nlayer, i);
        }
    }
    *isdr = 0;
    *jsdr = 0;
    *ildr = 0;
    *jdr = 0;
    *isdr = *isdr + F1(issp, *nsl) - 1 + 4 * F1(nlayer, *ndc + 1);
    *jsdr = *jsdr + F1(jssp, *nsl) - 1 + F1(nlayer, *ndc + 1) + 1;
    *ildr = *ildr + 4 * *nllay;
    *jdr = *jdr + F1(jsp, *ndc + 1) - 1 + 16 * F1(nlayer, *ndc + 1);
    *kdr = *kdr + F1(ksp, *ndc + 1) - 1 + 2 * (F1(nlayer, *ndc + 1) + 1);
}

/* calspo (P-SV) others.f90:527-584 */
static void p_calspo(const int *maxnlay, const int *maxnzone, const int *ndc, const double *rdc, const int *iphase,
                     const int *inlayer, double *r0, const double *rmin, const double *rmax, const double *ra,
                     const int *isp, double *spo, int *spn)
{
    (void)maxnlay;
    (void)maxnzone;
    (void)ndc;
    (void)rmin;
    *spo = 0;
    if (*r0 == *rmax) {
        *spo = (double)*inlayer - 0.01;
        *r0 = F1(ra, *inlayer) + (*spo - (double)(*inlayer - 1)) * (F1(ra, *inlayer + 1) - F1(ra, *inlayer));
    } else {
        int itmp = 2;
        while (!(*r0 < F1(ra, itmp))) itmp++;
        *spo = (double)(itmp - 2) + (*r0 - F1(ra, itmp - 1)) / (F1(ra, itmp) - F1(ra, itmp - 1));
        if (*spo - (double)(itmp - 2) < 0.01) {
            *spo = (double)(itmp - 2) + 0.01;
            *r0 = F1(ra, itmp - 1) + (*spo - (double)(itmp - 2)) * (F1(ra, itmp) - F1(ra, itmp - 1));
        }
        if (*spo - (double)(itmp - 2) > 0.99) {
            *spo = (double)(itmp - 2) + 0.99;
            *r0 = F1(ra, itmp - 1) + (*spo - (double)(itmp - 2)) * (F1(ra, itmp) - F1(ra, itmp - 1));
        }
    }
    *spn = 0;
    int itmp = 1;
    for (;;) {
        if (F1(iphase, itmp) == 1) {
            (*spn)++;
            if (*r0 <= F1(rdc, itmp)) break;
            itmp++;
        } else {
            (*spn)++;
            if (*r0 <= F1(rdc, itmp)) return; /* the reference pauses: source in the liquid */
            itmp++;
        }
    }
    *spo = *spo - (double)(F1(isp, *spn) - 1);
}

/* calstg (P-SV) others.f90:590-682 */
static void p_calstg(const int *maxnlay, const int *maxnzone, const int *nzone, const int *iphase, const double *rrho,
                     const double *vpv, const double *vph, const double *vsv, const double *vsh, const double *eta,
                     const int *nnl, const double *ra, const double *rmax, int *vnp, double *vra, double *rho,
                     double *kappa, double *ecKx, double *ecKy, double *ecKz, double *mu, double *ecL, double *ecN,
                     const double *r0, const int *spn, double *ecC0, double *ecF0, double *ecL0)
{
    (void)iphase;
    int NT = *maxnlay + 2 * *maxnzone + 1;
    for (int i = 0; i < NT; i++) vra[i] = rho[i] = kappa[i] = ecKx[i] = ecKy[i] = ecKz[i] = mu[i] = ecL[i] = ecN[i] = 0;
    int itmp = 0, jtmp = 0;
    for (int iz = 1; iz <= *nzone; iz++) {
        for (int i = 1; i <= F1(nnl, iz) + 1; i++) {
            itmp++;
            jtmp++;
            F1(vra, itmp) = F1(ra, jtmp);
            double trho = 0, tvpv = 0, tvph = 0, tvsv = 0, tvsh = 0, teta = 0, coef = 1.0;
            for (int j = 1; j <= 4; j++) {
                if (j == 1) coef = 1.0;
                else coef = coef * (F1(vra, itmp) / *rmax);
                int o = (j - 1) + 4 * (iz - 1);
                trho = trho + rrho[o] * coef;
                tvpv = tvpv + vpv[o] * coef;
                tvph = tvph + vph[o] * coef;
                tvsv = tvsv + vsv[o] * coef;
                tvsh = tvsh + vsh[o] * coef;
                teta = teta + eta[o] * coef;
            }
            F1(rho, itmp) = trho;
            F1(ecL, itmp) = F1(rho, itmp) * tvsv * tvsv;
            F1(ecN, itmp) = F1(rho, itmp) * tvsh * tvsh;
            double ecA = trho * tvph * tvph, ecC = trho * tvpv * tvpv;
            double ecF = teta * (ecA - 2.0 * F1(ecL, itmp));
            F1(kappa, itmp) = (4.0 * ecA + ecC + 4.0 * ecF - 4.0 * F1(ecN, itmp)) / 9.0;
            F1(ecKx, itmp) = ecA - 4.0 / 3.0 * F1(ecN, itmp);
            F1(ecKy, itmp) = ecF + 2.0 / 3.0 * F1(ecN, itmp);
            F1(ecKz, itmp) = (ecC + 2.0 * ecF) / 3.0;
        }
        jtmp--;
    }
    *vnp = itmp;
    double trho = 0, tvpv = 0, tvph = 0, tvsv = 0, tvsh = 0, teta = 0, coef = 1.0;
    for (int j = 1; j <= 4; j++) {
        if (j == 1) coef = 1.0;
        else coef = coef * (*r0 / *rmax);
        int o = (j - 1) + 4 * (*spn - 1);
        trho = trho + rrho[o] * coef;
        tvpv = tvpv + vpv[o] * coef;
        tvph = tvph + vph[o] * coef;
        tvsv = tvsv + vsv[o] * coef;
        tvsh = tvsh + vsh[o] * coef;
        teta = teta + eta[o] * coef;
    }
    *ecL0 = trho * tvsv * tvsv;
    double ecA0 = trho * tvph * tvph;
    *ecC0 = trho * tvpv * tvpv;
    *ecF0 = teta * (ecA0 - 2.0 * *ecL0);
}

/* calstg_for_stackpoints others.f90:687-762 */
static void p_calstg_sp(const double *r, const int *nzone, const double *vrmin, const double *vrmax,
                        const double *rrho, const double *vpv, const double *vph, const double *vsv, const double *vsh,
                        const double *eta, const double *qmu, const double *qkappa, double *lambda, double *qkp)
{
    (void)qmu;
    double rmax = F1(vrmax, *nzone);
    for (int iz = 1; iz <= *nzone; iz++) {
        if (*r > F1(vrmin, iz) && *r <= F1(vrmax, iz)) {
            double trho = 0, tvpv = 0, tvph = 0, tvsv = 0, tvsh = 0, teta = 0, coef = 1.0;
            for (int j = 1; j <= 4; j++) {
                if (j == 1) coef = 1.0;
                else coef = coef * (*r / rmax);
                int o = (j - 1) + 4 * (iz - 1);
                trho = trho + rrho[o] * coef;
                tvpv = tvpv + vpv[o] * coef;
                tvph = tvph + vph[o] * coef;
                tvsv = tvsv + vsv[o] * coef;
                tvsh = tvsh + vsh[o] * coef;
                teta = teta + eta[o] * coef;
            }
            *lambda = trho * tvph * tvph;
            *qkp = F1(qkappa, iz);
        }
    }
}

/* caltstg others.f90:767-831 (tmu is never assigned: left zero) */
static void p_caltstg(const int *maxnlay, const int *maxnzone, const int *nzone, const double *rrho,
                      const double *vpv, const double *vph, const double *vsv, const double *vsh, const double *eta,
                      const int *nnl, const double *ra, const double *rmax, double *tvra, double *tkappa,
                      double *tecKx, double *tecKy, double *tecKz, double *tmu, double *tecL, double *tecN)
{
    int NT = *maxnlay + 2 * *maxnzone + 1;
    for (int i = 0; i < NT; i++) tvra[i] = tkappa[i] = tecKx[i] = tecKy[i] = tecKz[i] = tmu[i] = tecL[i] = tecN[i] = 0;
    int itmp = 0, jtmp = 0;
    for (int iz = 1; iz <= *nzone; iz++) {
        for (int i = 1; i <= F1(nnl, iz) + 1; i++) {
            itmp++;
            jtmp++;
            F1(tvra, itmp) = F1(ra, jtmp);
            double trho = 0, tvpv = 0, tvph = 0, tvsv = 0, tvsh = 0, teta = 0, coef = 1.0;
            for (int j = 1; j <= 4; j++) {
                if (j == 1) coef = 1.0;
                else coef = coef * (F1(tvra, itmp) / *rmax);
                int o = (j - 1) + 4 * (iz - 1);
                trho = trho + rrho[o] * coef;
                tvpv = tvpv + vpv[o] * coef;
                tvph = tvph + vph[o] * coef;
                tvsv = tvsv + vsv[o] * coef;
                tvsh = tvsh + vsh[o] * coef;
                teta = teta + eta[o] * coef;
            }
            F1(tecL, itmp) = trho * tvsv * tvsv;
            F1(tecN, itmp) = trho * tvsh * tvsh;
            double ecA = trho * tvph * tvph, ecC = trho * tvpv * tvpv;
            double ecF = teta * (ecA - 2.0 * F1(tecL, itmp));
            F1(tkappa, itmp) = (4.0 * ecA + ecC + 4.0 * ecF - 4.0 * F1(tecN, itmp)) / 9.0;
            F1(tecKx, itmp) = ecA - 4.0 / 3.0 * F1(tecN, itmp);
            F1(tecKy, itmp) = ecF + 2.0 / 3.0 * F1(tecN, itmp);
            F1(tecKz, itmp) = (ecC + 2.0 * ecF) / 3.0;
        }
        jtmp--;
    }
}

static void p_calinv(const int *vnp, const double *rho, const double *kappa, double *rhoinv, double *kappainv)
{
    for (int i = 0; i < *vnp; i++) {
        rhoinv[i] = 1.0 / rho[i];
        kappainv[i] = 1.0 / kappa[i];
    }
}

/* calspdr others.f90:870-898 */
static void p_calspdr(const int *maxnzone, const int *nzone, const int *iphase, const int *nlayer, int *jjdr,
                      int *kkdr)
{
    (void)maxnzone;
    jjdr[0] = 1;
    kkdr[0] = 1;
    for (int iz = 1; iz <= *nzone - 1; iz++) {
        if (F1(iphase, iz) == 1) {
            F1(jjdr, iz + 1) = F1(jjdr, iz) + 16 * F1(nlayer, iz);
            if (F1(iphase, iz + 1) == 1) F1(kkdr, iz + 1) = F1(kkdr, iz) + 2 * F1(nlayer, iz);
            else F1(kkdr, iz + 1) = F1(kkdr, iz) + 2 * (F1(nlayer, iz) + 1);
        } else {
            F1(jjdr, iz + 1) = F1(jjdr, iz) + 4 * F1(nlayer, iz);
            if (F1(iphase, iz + 1) == 1) F1(kkdr, iz + 1) = F1(kkdr, iz) + (F1(nlayer, iz) + 1);
            else F1(kkdr, iz + 1) = F1(kkdr, iz) + F1(nlayer, iz);
        }
    }
}

static void p_callsuf(const double *omega, const int *nzone, const double *vrmax, const double *vsv, int *lsuf)
{
    double tvs = 0.0, coef = 1.0;
    for (int i = 1; i <= 4; i++) {
        if (i == 1) coef = 1.0;
        tvs = tvs + vsv[(i - 1) + 4 * (*nzone - 1)] * coef;
    }
    *lsuf = (int)(*omega * F1(vrmax, *nzone) / tvs - 0.5) + 1;
}

/* calcutd (P-SV) others.f90:1421-1499 -- defined but never called by the
 * reference's SGTpsv.  |U| of every node from the solution tmpc (solid rows
 * alternate U, V from kkdr; fluid rows are U; a node shared by two zones of
 * different phase counted once), threshold rat * max, nc the LAST node above
 * it (the loop's `cycle`), its zone, and kc the band row of node nc. */
void orc_psv_calcutd(const int *nzone, const int *nnlayer, const int *nnl, const zc *tmpc, const double *rat,
                     const int *nn, const int *iphase, const double *spo, const int *spn, const double *ra,
                     const int *kkdr, int *kc)
{
    (void)nnlayer;
    (void)spo;
    (void)spn;
    const int n = *nn, nz = *nzone;
    double *cU = calloc((size_t)n + 1, sizeof(double)), *amp = calloc((size_t)n + 1, sizeof(double));
    int *ml = calloc((size_t)nz + 1, sizeof(int));
    int iz = 2, jz = 1;
    for (int jj = 1; jj <= n; jj++) {
        if (iz <= nz) {
            if (jj == kkdr[iz - 1]) {
                if (iphase[iz - 1] != iphase[iz - 2]) jz = jz - 1;
                iz = iz + 1;
            }
        }
        if (iphase[iz - 2] == 1) {
            if ((jj - kkdr[iz - 2]) % 2 == 1) {
                cU[jz - 1] = cabs(tmpc[jj - 1]);
                jz = jz + 1;
            }
        } else {
            cU[jz - 1] = cabs(tmpc[jj - 1]);
            jz = jz + 1;
        }
    }
    double maxamp = -1.0;
    for (int i = 1; i <= jz - 1; i++) {
        amp[i - 1] = cU[i - 1];
        if (maxamp < amp[i - 1]) maxamp = amp[i - 1];
    }
    maxamp = maxamp * *rat;
    int nc = 1;
    for (int i = 1; i <= jz - 1; i++)
        if (amp[i - 1] > maxamp) nc = i;
    int i = 1, tzone = 0;
    for (int jj = 1; jj <= nz; jj++) {
        i = i + nnl[jj - 1];
        ml[jj - 1] = i;
    }
    for (int jj = nz; jj >= 1; jj--)
        if (ml[jj - 1] > nc) tzone = jj;
    if (tzone == 0) {  // the reference reads an unassigned tzone here: kc unchanged
        free(cU);
        free(amp);
        free(ml);
        return;
    }
    const double rc = ra[nc - 1];
    for (int ii = 1; ii <= jz - 1; ii++) {
        if (ra[ii - 1] <= rc && rc < ra[ii]) {
            nc = ii;
            if (tzone == 1) {
                if (iphase[tzone - 1] == 1) *kc = 1 + 2 * nc;
                if (iphase[tzone - 1] == 2) *kc = 1 + nc;
            } else {
                if (iphase[tzone - 1] == 1) *kc = kkdr[tzone - 1] + 2 * (nc - ml[tzone - 2]);
                if (iphase[tzone - 1] == 2) *kc = kkdr[tzone - 1] + nc - ml[tzone - 2];
            }
        }
    }
    free(cU);
    free(amp);
    free(ml);
}

/* Test hook: the solution g0 (nn values) of the (l, m, imt) right-hand side
 * of the next orc_psv_omega calls, copied after its band solve. */
static int CAP_L = -1, CAP_M = 0, CAP_IMT = 0, CAP_N = 0;
static zc *CAP_BUF = NULL;

int orc_psv_capture(int32_t l, int32_t m, int32_t imt, double complex *buf, int32_t cap)
{
    CAP_L = buf ? l : -1;
    CAP_M = m;
    CAP_IMT = imt;
    CAP_BUF = buf;
    CAP_N = cap;
    return 0;
}

static void psv_capture(int l, int m, int imt, const zc *g0, int nn)
{
    if (l != CAP_L || m != CAP_M || imt != CAP_IMT || !CAP_BUF) return;
    for (int k = 0; k < nn && k < CAP_N; k++) CAP_BUF[k] = g0[k];
}

/* calamp (P-SV) others.f90:1246-1264 */
static void p_calamp(const zc *g, const int *l, const int *lsuf, double *maxamp, int *ismall, const double *ratl)
{
    double ampratio = 0.0;
    double a1 = cabs(g[0]), a2 = cabs(g[1]);
    double amp = sqrt(a1 * a1 + a2 * a2);
    if (amp > *maxamp) *maxamp = amp;
    if (amp != 0.0 && *maxamp != 0.0) ampratio = amp / *maxamp;
    if (ampratio < *ratl && *l >= *lsuf) *ismall = *ismall + 1;
    else *ismall = 0;
}

/* calu / calup / calup0 others.f90:968-1004 */
static void p_calu(const zc *c0, const double *lsq, const zc *bvec, zc *u)
{
    u[0] = u[0] + c0[0] * bvec[0];
    u[1] = u[1] + c0[1] * bvec[1] / CMPLX(*lsq, 0.0);
    u[2] = u[2] + c0[1] * bvec[2] / CMPLX(*lsq, 0.0);
}
static void p_calup(const zc *c1, const zc *c2, const double *lsq, const zc *bvec, zc *u)
{
    u[0] = u[0] + *c1 * bvec[0];
    u[1] = u[1] + *c2 * bvec[1] / CMPLX(*lsq, 0.0);
    u[2] = u[2] + *c2 * bvec[2] / CMPLX(*lsq, 0.0);
}
static void p_calup0(const zc *c1, const zc *bvec, zc *u) { u[0] = u[0] + *c1 * bvec[0]; }

/* calupfluid others.f90:1009-1031 */
static void p_calupfluid(const zc *c1, const zc *omega, const double *lambda, const double *qkp, const zc *ylm,
                         zc *strain)
{
    const double deuxsurpi = 0.63661977236, unsurdeuxpi = 0.15915494309, unsurtrois = 0.33333333333333333333;
    zc fl = CMPLX(*lambda, 0.0) * CMPLX(1.0, 1.0 / *qkp);
    fl = fl * CMPLX(1 + deuxsurpi / *qkp * log(unsurdeuxpi * creal(*omega)), 0.0);
    strain[0] = *omega * unsurtrois / fl * *c1 * *ylm;
    strain[4] = strain[0];
    strain[8] = strain[0];
}

/* interpolate (P-SV copy) others.f90:1170-1239 */
static void p_interpolate(const int *ncomp, const int *nderiv, const double *rsta, const double *rrsta, const zc *g,
                          zc *u)
{
    double eps = -1.0, dh[3];
    for (int ic = 0; ic < *ncomp; ic++) u[ic] = 0.0;
    for (int i = 0; i < 3; i++) dh[i] = rrsta[i] - *rsta;
    if (dh[1] == 0.0 && *nderiv == 0) {
        int itmp = *ncomp + 1;
        for (int ic = 1; ic <= *ncomp; ic++) {
            u[ic - 1] = g[itmp - 1];
            itmp++;
        }
        return;
    }
    zc a[9], b[3] = {0, 0, 0}, wk[3];
    int ip[3], ier;
    for (int i = 1; i <= 3; i++) {
        a[0 + 3 * (i - 1)] = 1.0;
        a[1 + 3 * (i - 1)] = dh[i - 1];
        a[2 + 3 * (i - 1)] = dh[i - 1] * dh[i - 1] / 2.0;
    }
    b[*nderiv] = 1.0;
    p_glu(a, 3, 3, b, &eps, wk, ip, &ier);
    for (int ic = 1; ic <= *ncomp; ic++)
        for (int i = 1; i <= 3; i++) u[ic - 1] = u[ic - 1] + b[i - 1] * g[*ncomp * (i - 1) + ic - 1];
}

/* locallyCartesianDerivatives (P-SV) others.f90:310-344: cmplx() of a
 * real(8) is a default-kind (single precision) complex, so r, sin(theta)
 * and cot(theta) enter at float precision. */
#define U_(i, j) uder[((i) - 1) + 3 * ((j) - 1)]
static void p_lcd(const zc *u, const zc *udr, const zc *udt, const zc *udp, zc *uder, const double *r,
                  const double *theta)
{
    const double rf = (double)(float)*r;
    if (*theta == 0.0 || *theta == PI) {
        U_(1, 1) = udr[0];
        U_(2, 2) = u[0] / rf;
        U_(3, 3) = u[0] / rf;
        return;
    }
    double thetasin = sin(*theta);
    double thetacot = cos(*theta) / thetasin;
    const double sf = (double)(float)thetasin, cf = (double)(float)thetacot;
    U_(1, 1) = udr[0];
    U_(1, 2) = (udt[0] - u[1]) / rf;
    U_(1, 3) = (udp[0] / sf - u[2]) / rf;
    U_(2, 1) = udr[1];
    U_(2, 2) = (udt[1] + u[0]) / rf;
    U_(2, 3) = (udp[1] / sf - u[2] * cf) / rf;
    U_(3, 1) = udr[2];
    U_(3, 2) = udt[2] / rf;
    U_(3, 3) = (udp[2] / sf + u[0] + u[1] * cf) / rf;
}

/* udertorsgt others.f90:204-227 */
static void p_udertorsgt(const int *icomp, const zc *uder, zc *r)
{
    if (*icomp == 1) {
        r[0] = r[0] + U_(1, 1);
        r[1] = r[1] - 0.5 * U_(2, 1) - 0.5 * U_(1, 2);
        r[2] = r[2] + 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
        r[3] = r[3] - 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
    }
    if (*icomp == 2) {
        r[4] = r[4] + U_(1, 1);
        r[5] = r[5] - 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
        r[6] = r[6] - 0.5 * U_(2, 2) + 0.5 * U_(3, 3);
        r[8] = r[8] - 0.5 * U_(1, 2) - 0.5 * U_(2, 1);
    }
    if (*icomp == 3) {
        r[7] = r[7] - 0.5 * U_(2, 3) - 0.5 * U_(3, 2);
        r[9] = r[9] + 0.5 * U_(1, 3) + 0.5 * U_(3, 1);
    }
}

/* udertotsgt others.f90:231-285 */
static void p_udertotsgt(const int *imt, const zc *uder, zc *t)
{
    switch (*imt) {
    case 1:
        t[0] = t[0] + U_(1, 1);
        t[2] = t[2] - 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
        t[5] = t[5] - 0.5 * U_(1, 2) - 0.5 * U_(2, 1);
        t[9] = t[9] + 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
        break;
    case 2:
        t[1] = t[1] - 0.5 * U_(1, 1);
        t[3] = t[3] + 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
        t[7] = t[7] + 0.25 * U_(1, 2) + 0.25 * U_(2, 1);
        t[8] = t[8] + 0.5 * U_(1, 1);
        t[10] = t[10] - 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
        t[11] = t[11] - 0.25 * U_(2, 2) - 0.25 * U_(3, 3);
        t[16] = t[16] - 0.125 * U_(1, 2) - 0.125 * U_(2, 1);
        t[17] = t[17] + 0.125 * U_(1, 2) + 0.125 * U_(2, 1);
        t[18] = t[18] + 0.125 * U_(2, 2) - 0.125 * U_(3, 3);
        t[19] = t[19] + 0.125 * U_(2, 2) - 0.125 * U_(3, 3);
        break;
    case 3:
        t[1] = t[1] - 0.5 * U_(1, 1);
        t[3] = t[3] + 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
        t[7] = t[7] - 0.25 * U_(1, 2) - 0.25 * U_(2, 1);
        t[8] = t[8] - 0.5 * U_(1, 1);
        t[10] = t[10] - 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
        t[11] = t[11] + 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
        t[16] = t[16] - 0.125 * U_(1, 2) - 0.125 * U_(2, 1);
        t[17] = t[17] + 0.125 * U_(1, 2) + 0.125 * U_(2, 1);
        t[18] = t[18] - 0.125 * U_(2, 2) + 0.125 * U_(3, 3);
        t[19] = t[19] - 0.125 * U_(2, 2) + 0.125 * U_(3, 3);
        break;
    case 4:
        t[4] = t[4] + U_(1, 1);
        t[6] = t[6] - 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
        t[12] = t[12] - 0.25 * U_(1, 2) - 0.25 * U_(2, 1);
        t[13] = t[13] + 0.25 * U_(1, 2) + 0.25 * U_(2, 1);
        t[14] = t[14] + 0.25 * U_(2, 2) - 0.25 * U_(3, 3);
        t[15] = t[15] - 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
        break;
    case 5:
        t[12] = t[12] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[13] = t[13] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[14] = t[14] - 0.25 * U_(2, 3) - 0.25 * U_(3, 2);
        t[15] = t[15] - 0.25 * U_(2, 3) - 0.25 * U_(3, 2);
        break;
    case 6:
        t[16] = t[16] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[17] = t[17] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[18] = t[18] - 0.25 * U_(2, 3) - 0.25 * U_(3, 2);
        t[19] = t[19] + 0.25 * U_(2, 3) + 0.25 * U_(3, 2);
        break;
    }
}
#undef U_

/* utosynn others.f90:161-198 */
static void p_utosynn(const int *imt, const zc *u, zc *s)
{
    if (*imt == 1) {
        s[0] = s[0] + u[0];
        s[4] = s[4] + u[1];
    }
    if (*imt == 2) {
        s[2] = s[2] + 0.5 * u[0];
        s[3] = s[3] - 0.5 * u[0];
        s[5] = s[5] - 0.5 * u[1];
        s[6] = s[6] - 0.5 * u[1];
    }
    if (*imt == 3) {
        s[2] = s[2] - 0.5 * u[0];
        s[3] = s[3] - 0.5 * u[0];
        s[5] = s[5] - 0.5 * u[1];
        s[6] = s[6] + 0.5 * u[1];
    }
    if (*imt == 4) {
        s[1] = s[1] + u[0];
        s[8] = s[8] + u[1];
    }
    if (*imt == 5) s[9] = s[9] + u[2];
    if (*imt == 6) s[7] = s[7] - u[2];
}

/* ======================================================================= */
/* src/SGTpsv/trial.f90: Legendre recurrence and spheroidal harmonics       */
/* ======================================================================= */
static void p_calplm(int l, int m, double x, double *plm)
{
    if (l == m) {
        double pmm = 1.0;
        if (m > 0) {
            double somx2 = sqrt((1.0 - x) * (1.0 + x)), fact = 1.0;
            for (int i = 1; i <= m; i++) {
                pmm = -pmm * fact * somx2;
                fact = fact + 2.0;
            }
        }
        plm[2] = 0.0;
        plm[1] = 0.0;
        plm[0] = pmm;
    } else {
        plm[2] = plm[1];
        plm[1] = plm[0];
        if (l == m + 1) plm[0] = x * (double)(2 * m + 1) * plm[1];
        else plm[0] = (x * (double)(2 * l - 1) * plm[1] - (double)(l + m - 1) * plm[2]) / (double)(l - m);
    }
}

#define BV(a, i, m) (a)[((i) - 1) + 3 * ((m) + 2)]
/* caldvecphi0_withplm trial.f90:97-223.  bvecdt(1,-m) is never assigned
 * by the reference; it keeps the zero of a fresh array. */
static void p_caldvecphi0(const int *lp, const double *thetap, double *plm, zc *bvec, zc *bvecdt, zc *bvecdp)
{
    int l = *lp;
    double theta = *thetap;
    if (theta == 0.0 || theta == PI) {
        double sign1 = 1.0;
        if (theta == PI) sign1 = (l % 2 == 0) ? 1.0 : -1.0;
        for (int i = 0; i < 15; i++) bvec[i] = bvecdt[i] = bvecdp[i] = 0.0;
        double coeff = sqrt((double)(2 * l + 1) / (4.0 * PI));
        BV(bvec, 1, 0) = (double)(float)coeff * sign1; /* cmplx(coeff,0.d0): default kind */
        return;
    }
    double x = cos(theta), xl2 = (double)l * (double)(l + 1);
    for (int m = 0; m <= (l < 3 ? l : 3); m++) p_calplm(l, m, x, &plm[3 * m]);
    for (int m = 0; m <= (l < 2 ? l : 2); m++) {
        double fact = 1.0;
        if (m != 0)
            for (int i = l - m + 1; i <= l + m; i++) fact = fact * (double)i;
        double coef = sqrt((double)(2 * l + 1) / (4.0 * PI) / fact);
        double P = plm[3 * m], P1 = plm[3 * (m + 1)];
        double plmdt = (double)m * x / sin(theta) * P + P1;
        zc im = CMPLX(0.0, (double)m);
        BV(bvec, 1, m) = coef * P;
        BV(bvec, 1, -m) = conj(BV(bvec, 1, m));
        BV(bvec, 2, m) = coef * plmdt;
        BV(bvec, 2, -m) = conj(BV(bvec, 2, m));
        BV(bvec, 3, m) = im / sin(theta) * coef * P;
        BV(bvec, 3, -m) = conj(BV(bvec, 3, m));
        BV(bvecdt, 1, m) = plmdt * coef;
        BV(bvecdt, 1, m) = conj(BV(bvecdt, 1, m));
        BV(bvecdt, 2, m) = (-x / sin(theta) * plmdt + (double)m * (double)m / (1 - x * x) * P - xl2 * P) * coef;
        BV(bvecdt, 2, -m) = conj(BV(bvecdt, 2, m));
        BV(bvecdt, 3, m) = im * (-x / (1 - x * x) * P + 1.0 / sin(theta) * plmdt) * coef;
        BV(bvecdt, 3, -m) = conj(BV(bvecdt, 3, m));
        BV(bvecdp, 1, m) = im * P * coef;
        BV(bvecdp, 1, -m) = conj(BV(bvecdp, 1, m));
        BV(bvecdp, 2, m) = im * plmdt * coef;
        BV(bvecdp, 2, -m) = conj(BV(bvecdp, 2, m));
        BV(bvecdp, 3, m) = -(double)m * (double)m / sin(theta) * P * coef;
        BV(bvecdp, 3, -m) = conj(BV(bvecdp, 3, m));
        if (m % 2 == 1) {
            for (int i = 1; i <= 3; i++) {
                BV(bvec, i, -m) = -BV(bvec, i, -m);
                BV(bvecdt, i, -m) = -BV(bvecdt, i, -m);
                BV(bvecdp, i, -m) = -BV(bvecdp, i, -m);
            }
        }
    }
}

/* caldveczero trial.f90:326-364 (m=0 coefficient via default-kind cmplx) */
static void p_caldveczero(const int *lp, zc *bvec)
{
    int l = *lp;
    for (int i = 0; i < 15; i++) bvec[i] = 0.0;
    double xl2 = (double)l * (double)(l + 1);
    for (int m = 0; m <= (l < 1 ? l : 1); m++) {
        double fact = 1.0;
        if (m != 0)
            for (int i = l - m + 1; i <= l + m; i++) fact = fact * (double)i;
        double coef = sqrt((double)(2 * l + 1) / (4.0 * PI) / fact);
        if (m == 0) BV(bvec, 1, m) = (double)(float)coef;
        if (m == 1) {
            BV(bvec, 2, m) = CMPLX(-1.0, 0.0) * xl2 * coef / 2.0;
            BV(bvec, 2, -m) = conj(BV(bvec, 2, m));
            BV(bvec, 3, m) = CMPLX(0.0, (double)m) * xl2 * coef / 2.0;
            BV(bvec, 3, -m) = conj(BV(bvec, 3, m));
        }
        if (m % 2 == 1) {
            BV(bvec, 2, -m) = -BV(bvec, 2, -m);
            BV(bvec, 3, -m) = -BV(bvec, 3, -m);
        }
    }
}
#undef BV

/* ======================================================================= */
/* function table                                                            */
/* ======================================================================= */
typedef struct {
    void (*calgrid)();
    void (*calra_psv)();
    void (*calra2_psv)();
    void (*calsp)();
    void (*calspo)();
    void (*calstg)();
    void (*calstg_sp)();
    void (*caltstg)();
    void (*calinv)();
    void (*calmatc)();
    void (*caltl)();
    void (*calhl)();
    void (*calt)();
    void (*calh5)();
    void (*calhm1)();
    void (*calhm2)();
    void (*mtrnp)();
    void (*mtrnp2)();
    void (*submat)();
    void (*callsuf)();
    void (*calcoef)();
    void (*cala0)();
    void (*cala1)();
    void (*cala2)();
    void (*calb0)();
    void (*calb2)();
    void (*calhml)();
    void (*overlapa)();
    void (*overlapb)();
    void (*calabnum)();
    void (*calya)();
    void (*calg)();
    void (*cala)();
    void (*calbc)();
    void (*calspdr)();
    void (*dcsymbdl0)();
    void (*dcsbdlv0)();
    void (*mydcsbdlv3)();
    void (*rea2)();
    void (*rea2_back)();
    void (*caldvecphi0)();
    void (*caldveczero)();
    void (*calamp)();
    void (*calu)();
    void (*calup)();
    void (*calup0)();
    void (*calupfluid)();
    void (*interpolate)();
    void (*lcd)();
    void (*udertotsgt)();
    void (*udertorsgt)();
    void (*utosynn)();
    void (*calnl)();
} psv_fns_t;

static const psv_fns_t PC_FNS = {
    (void (*)())p_calgrid,    (void (*)())p_calra_psv, (void (*)())p_calra2_psv, (void (*)())p_calsp,
    (void (*)())p_calspo,     (void (*)())p_calstg,    (void (*)())p_calstg_sp,  (void (*)())p_caltstg,
    (void (*)())p_calinv,     (void (*)())p_calmatc,   (void (*)())p_caltl,      (void (*)())p_calhl,
    (void (*)())p_calt,       (void (*)())p_calh5,     (void (*)())p_calhm1,     (void (*)())p_calhm2,
    (void (*)())p_mtrnp,      (void (*)())p_mtrnp2,    (void (*)())p_submat,     (void (*)())p_callsuf,
    (void (*)())p_calcoef,    (void (*)())p_cala0,     (void (*)())p_cala1,      (void (*)())p_cala2,
    (void (*)())p_calb0,      (void (*)())p_calb2,     (void (*)())p_calhml,     (void (*)())p_overlapa,
    (void (*)())p_overlapb,   (void (*)())p_calabnum,  (void (*)())p_calya,      (void (*)())p_calg,
    (void (*)())p_cala,       (void (*)())p_calbc,     (void (*)())p_calspdr,    (void (*)())p_dcsymbdl0,
    (void (*)())p_dcsbdlv0,   (void (*)())p_mydcsbdlv3, (void (*)())p_rea2,      (void (*)())p_rea2_back,
    (void (*)())p_caldvecphi0, (void (*)())p_caldveczero, (void (*)())p_calamp,  (void (*)())p_calu,
    (void (*)())p_calup,      (void (*)())p_calup0,    (void (*)())p_calupfluid, (void (*)())p_interpolate,
    (void (*)())p_lcd,        (void (*)())p_udertotsgt, (void (*)())p_udertorsgt, (void (*)())p_utosynn,
    (void (*)())p_calnl};

static psv_fns_t PREF_FNS, PMIX_FNS;
static int PREF_LOADED = 0;

int orc_psv_load_ref(const char *path)
{
    void *h = dlopen(path, RTLD_LAZY | RTLD_LOCAL);
    if (!h) return -1;
    const char *names[] = {"calgrid_",     "calra_psv_",  "calra2_psv_",   "calsp_",       "calspo_",
                           "calstg_",      "calstg_for_stackpoints_",      "caltstg_",     "calinv_",
                           "calmatc_",     "caltl_",      "calhl_",        "calt_",        "calh5_",
                           "calhm1_",      "calhm2_",     "mtrnp_",        "mtrnp2_",      "submat_",
                           "callsuf_",     "calcoef_",    "cala0_",        "cala1_",       "cala2_",
                           "calb0_",       "calb2_",      "calhml_",       "overlapa_",    "overlapb_",
                           "calabnum_",    "calya_",      "calg_",         "cala_",        "calbc_",
                           "calspdr_",     "dcsymbdl0_",  "dcsbdlv0_",     "mydcsbdlv3_",  "rea2_",
                           "rea2_back_",   "caldvecphi0_withplm_",         "caldveczero_", "calamp_",
                           "calu_",        "calup_",      "calup0_",       "calupfluid_",  "interpolate_",
                           "locallycartesianderivatives_", "udertotsgt_",  "udertorsgt_",  "utosynn_",
                           "calnl_"};
    int n = (int)(sizeof(names) / sizeof(names[0]));
    if ((size_t)n * sizeof(void *) != sizeof(PREF_FNS)) return -2;
    void **dst = (void **)&PREF_FNS;
    for (int k = 0; k < n; k++) {
        dst[k] = dlsym(h, names[k]);
        if (!dst[k]) return -3 - k;
    }
    PREF_LOADED = 1;
    return 0;
}

int orc_psv_mix(uint64_t mask)
{
    if (!PREF_LOADED) return -10;
    void **dst = (void **)&PMIX_FNS;
    void *const *c = (void *const *)&PC_FNS, *const *r = (void *const *)&PREF_FNS;
    for (size_t k = 0; k < sizeof(PMIX_FNS) / sizeof(void *); k++) dst[k] = ((mask >> k) & 1) ? r[k] : c[k];
    return 0;
}

#include "psv_driver.inc"
