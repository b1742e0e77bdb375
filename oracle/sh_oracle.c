/*
 * sh_oracle.c -- CPU restatement of the SGTsh engine (toroidal SH,
 * solid only) of IPGP/DSM-Kernel, FP64, loop by loop.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle and the timed CPU baseline.
 * The product library never links or calls this file.
 *
 * Each o_<name> routine restates the reference routine <name> with the
 * reference's argument list (pointers, column-major, 1-based index
 * arithmetic written as [i-1]).  The driver orc_sh_omega() calls them
 * through a function table; orc_sh_load_ref() fills a second table with the
 * reference's own routines compiled from the src/SGTsh sources (oracle/_ref), so
 * the same driver can run on either, which pins every restated routine in
 * the context of the driver.
 */
#include "dsm_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;
static const double PI = 3.1415926535897932;

/* ----------------------------------------------------------------------- */
/* FE element integrals: src/SGTsh/calmat.f90                               */
/* ----------------------------------------------------------------------- */

/* pmulti  calmat.f90:164-182 */
static void o_pmulti(int n, const double *a, int m, const double *b, int l, double *c)
{
    for (int i = 1; i <= l; i++) c[i - 1] = 0.0;
    for (int i = 1; i <= n; i++)
        for (int j = 1; j <= m; j++) c[i + j - 2] = c[i + j - 2] + a[i - 1] * b[j - 1];
}

/* polint  calmat.f90:230-262 */
static void o_polint(int n, const double *p, double x1, double x2, double *pint)
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

/* pinteg  calmat.f90:186-225: integral of p(r)*con(r), con linear between
 * the structure nodes vra (zero-length intervals skipped). */
static void o_pinteg(int *snp, int n, const double *p, double from, double to, int vnp,
                     const double *vra, const double *con, double *pint)
{
    double q[2], pq[6], psint, x1, x2;
    if (*snp >= vnp) *snp = 1;
    *pint = 0.0;
    x1 = from;
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
        o_pmulti(n, p, 2, q, n + 1, pq);
        o_polint(n + 1, pq, x1, x2, &psint);
        *pint = *pint + psint;
        if (x2 == to) break;
        x1 = x2;
    }
}

/* calmatc  calmat.f90:2-83: \int r^rpow con W_p^(w1dn) W_q^(w2dn) dr */
static void o_calmatc(const int *nlayer, const int *vnp, const double *vra, const double *con,
                      const int *rpow, const int *w1dn, const int *w2dn, const double *ra,
                      double *m, double *work)
{
    int snp = 1;
    double a[2][2], b[2][2], c[5];  /* a[col][row] == Fortran a(row,col) */
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
                o_pmulti(2, a[j - 1], 2, b[k - 1], 3, c);
                for (int l = 3; l >= 1; l--) {
                    c[l + *rpow - 1] = c[l - 1];
                    if (*rpow > 0) c[l - 1] = 0.0;
                }
                int nn = 4 * (i - 1) + 2 * (j - 1) + k;
                o_pinteg(&snp, 5, c, ra[i - 1], ra[i], *vnp, vra, con, &work[nn - 1]);
            }
    }
    if (*w1dn != *w2dn) {
        for (int i = 1; i <= 4 * *nlayer; i++) {
            if (i % 4 == 0 || i % 4 == 1) m[i - 1] = 2.0 * work[i - 1];
            else if (i % 4 == 2) m[i - 1] = work[i - 1] + work[i];
            else m[i - 1] = work[i - 2] + work[i - 1];
        }
    } else {
        for (int i = 1; i <= 4 * *nlayer; i++) m[i - 1] = work[i - 1];
    }
}

/* caltl  calmat.f90:87-118: lumped mass */
static void o_caltl(const int *nlayer, const int *vnp, const double *vra, const double *rho,
                    const double *ra, double *tl)
{
    int snp = 1;
    double c[3] = {0.0, 0.0, 1.0};
    for (int i = 1; i <= *nlayer; i++) {
        double from = ra[i - 1], to = (ra[i - 1] + ra[i]) / 2.0;
        int nn = 4 * (i - 1);
        o_pinteg(&snp, 3, c, from, to, *vnp, vra, rho, &tl[nn]);
        tl[nn + 1] = 0.0;
        tl[nn + 2] = 0.0;
        from = to;
        to = ra[i];
        o_pinteg(&snp, 3, c, from, to, *vnp, vra, rho, &tl[nn + 3]);
    }
}

/* calhl  calmat.f90:122-147: lumped rigidity */
static void o_calhl(const int *nlayer, const int *vnp, const double *vra, const double *mu,
                    const double *ra, double *hl)
{
    int snp = 1;
    double c[1] = {1.0};
    for (int i = 1; i <= *nlayer; i++) {
        double from = ra[i - 1], to = (ra[i - 1] + ra[i]) / 2.0;
        int nn = 4 * (i - 1);
        o_pinteg(&snp, 1, c, from, to, *vnp, vra, mu, &hl[nn]);
        hl[nn + 1] = 0.0;
        hl[nn + 2] = 0.0;
        from = to;
        to = ra[i];
        o_pinteg(&snp, 1, c, from, to, *vnp, vra, mu, &hl[nn + 3]);
    }
}

/* calt  calmat.f90:151-159 */
static void o_calt(const int *nlayer, const double *tl, const double *tc, double *t)
{
    for (int i = 1; i <= 4 * *nlayer; i++) t[i - 1] = (tl[i - 1] + tc[i - 1]) / 2.0;
}

/* cala0  calmat.f90:267-285 */
static void o_cala0(const int *nlayer, const double *omega, const double *omegai, const double *t,
                    const double *h1, const double *h2, const double *h3, const double *h4,
                    const zc *coef, zc *a0)
{
    zc w = CMPLX(*omega, -*omegai);
    zc comega2 = w * w;
    for (int i = 1; i <= 4 * *nlayer; i++) {
        double h = h1[i - 1] - h2[i - 1] + h3[i - 1] - 2.0 * h4[i - 1];
        a0[i - 1] = comega2 * t[i - 1] - *coef * h;
    }
}

/* cala2  calmat.f90:290-304 */
static void o_cala2(const int *nlayer, const double *h4, const zc *coef, zc *a2)
{
    for (int i = 1; i <= 4 * *nlayer; i++) a2[i - 1] = -*coef * h4[i - 1];
}

/* cala  calmat.f90:309-324 */
static void o_cala(const int *nn, const int *l, const int *lda, const zc *a0, const zc *a2, zc *a)
{
    double xl2 = (double)*l * (double)(*l + 1);
    for (int j = 1; j <= *nn; j++)
        for (int i = 1; i <= 2; i++)
            a[(i - 1) + *lda * (j - 1)] = a0[(i - 1) + *lda * (j - 1)] + xl2 * a2[(i - 1) + *lda * (j - 1)];
}

/* calga  calmat.f90:329-353 */
static void o_calga(const int *nlayer, const double *omega, const double *omegai, const int *l,
                    const double *t, const double *h1, const double *h2, const double *h3,
                    const double *h4, const zc *coef, zc *a)
{
    for (int i = 1; i <= 4 * *nlayer; i++) a[i - 1] = 0.0;
    zc w = CMPLX(*omega, -*omegai);
    zc comega2 = w * w;
    double xl2m2 = (double)*l * (double)(*l + 1) - 2.0;
    for (int i = 1; i <= 4 * *nlayer; i++) {
        double h = h1[i - 1] - h2[i - 1] + h3[i - 1] + xl2m2 * h4[i - 1];
        a[i - 1] = comega2 * t[i - 1] - *coef * h;
    }
}

/* overlap  calmat.f90:386-423: element 2x2 blocks -> band a2(2,*) */
static void o_overlap(const int *nlayer, const zc *a, zc *a2)
{
    int nsize = *nlayer + 1, mu = 1, m = mu + 1;
    for (int j = 1; j <= nsize; j++) {
        int i1 = (1 > j - mu) ? 1 : j - mu;
        for (int i = i1; i <= j; i++) {
            int k = i - j + m;
            zc *dst = &a2[(k - 1) + 2 * (j - 1)];
            if (i == j) {
                if (i == 1) *dst = *dst + a[0];
                else if (i == nsize) *dst = *dst + a[4 * *nlayer - 1];
                else {
                    int k1 = 4 * i - 4, k2 = k1 + 1;
                    *dst = *dst + a[k1 - 1] + a[k2 - 1];
                }
            }
            if (i + 1 == j) {
                int k1 = 4 * i - 2;
                *dst = *dst + a[k1 - 1];
            }
        }
    }
}

/* ----------------------------------------------------------------------- */
/* Band LDL^T: src/SGTsh/dclib.f90 (dclisb0 + entry dcsbsub0)              */
/* ----------------------------------------------------------------------- */

static void o_dcsbsub0(zc *a, const int *n, const int *nud, const int *n1, zc *b, double *eps,
                       zc *dr, zc *z, int *ier);

/* dclisb0  dclib.f90:1-88: complex-symmetric LDL^T without pivoting;
 * a(n1,n) holds a(m-k,j) = A(j-k,j); reciprocal pivots in dr. */
static void o_dclisb0(zc *a, const int *n, const int *nud, const int *n1, zc *b, double *eps,
                      zc *dr, zc *z, int *ier)
{
#define A_(i, j) a[((i) - 1) + *n1 * ((j) - 1)]
    *ier = 0;
    double eps1 = 1.0e-14;
    int m = *nud + 1;
    if (*eps <= 0.0) *eps = eps1;
    dr[0] = 1.0 / A_(m, 1);
    zc xx = A_(m - 1, 2);
    A_(m - 1, 2) = A_(m - 1, 2) * dr[0];
    zc s = A_(m, 2) - xx * A_(m - 1, 2);
    dr[1] = 1.0 / s;
    if (m < 3) {
        for (int j = 3; j <= *n; j++) {
            xx = A_(1, j);
            A_(1, j) = xx * dr[j - 2];
            s = A_(2, j) - xx * A_(1, j);
            if (cabs(s) <= *eps) *ier = 1;
            dr[j - 1] = 1.0 / s;
        }
    } else {
        for (int j = 3; j <= *n; j++) {
            int k1 = 1;
            if (j >= m) k1 = j - m + 1;
            int mj = m - j;
            for (int i = k1 + 1; i <= j - 1; i++) {
                zc sum = 0.0;
                for (int k = k1; k <= i - 1; k++) sum = sum + A_(m - i + k, i) * A_(mj + k, j);
                A_(mj + i, j) = A_(mj + i, j) - sum;
            }
            zc sum = 0.0;
            for (int i = k1; i <= j - 1; i++) {
                zc x = A_(mj + i, j);
                zc au = x * dr[i - 1];
                sum = sum + x * au;
                A_(mj + i, j) = au;
            }
            zc t = A_(m, j) - sum;
            if (cabs(t) <= *eps) *ier = 1;
            dr[j - 1] = 1.0 / t;
        }
    }
#undef A_
    o_dcsbsub0(a, n, nud, n1, b, eps, dr, z, ier);
}

/* dcsbsub0  dclib.f90:90-137 (entry): forward, diagonal, backward */
static void o_dcsbsub0(zc *a, const int *n, const int *nud, const int *n1, zc *b, double *eps,
                       zc *dr, zc *z, int *ier)
{
#define A_(i, j) a[((i) - 1) + *n1 * ((j) - 1)]
    (void)eps;
    (void)ier;
    int m = *nud + 1, N = *n;
    if (m < 3) {
        z[0] = b[0];
        for (int j = 2; j <= N; j++) z[j - 1] = b[j - 1] - A_(1, j) * z[j - 2];
        for (int j = 1; j <= N; j++) z[j - 1] = z[j - 1] * dr[j - 1];
        b[N - 1] = z[N - 1];
        for (int j = 1; j <= N - 1; j++) b[N - j - 1] = z[N - j - 1] - A_(1, N - j + 1) * b[N - j];
    } else {
        z[0] = b[0];
        z[1] = b[1] - A_(m - 1, 2) * z[0];
        for (int j = 3; j <= N; j++) {
            int i1 = (j > m) ? 1 : m - j + 1;
            zc sum = 0.0;
            for (int k = i1; k <= m - 1; k++) sum = sum + A_(k, j) * z[j - m + k - 1];
            z[j - 1] = b[j - 1] - sum;
        }
        for (int j = 1; j <= N; j++) z[j - 1] = z[j - 1] * dr[j - 1];
        b[N - 1] = z[N - 1];
        b[N - 2] = z[N - 2] - A_(m - 1, N) * z[N - 1];
        for (int j = 3; j <= N; j++) {
            int j1 = N - j + 1;
            int i1 = 1;
            if (j < m) i1 = m - j + 1;
            zc sum = 0.0;
            for (int k = i1; k <= m - 1; k++) sum = sum + A_(k, m - k + j1) * b[m - k + j1 - 1];
            b[j1 - 1] = z[j1 - 1] - sum;
        }
    }
#undef A_
}

/* ----------------------------------------------------------------------- */
/* Dense LU with partial pivoting: src/SGTsh/glu2.f90:2-143                */
/* ----------------------------------------------------------------------- */
static void o_glu(zc *a, int n, int n1, zc *b, double *eps, zc *wk, int *ip, int *ier)
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

/* ----------------------------------------------------------------------- */
/* src/SGTsh/others.f90                                                     */
/* ----------------------------------------------------------------------- */

/* setmt  others.f90:113-127 (rr, tt, pp, rt, rp, tp order), mt(3,3) */
static void o_setmt(int imt, double *mt)
{
    for (int i = 0; i < 9; i++) mt[i] = 0.0;
    if (imt == 1) mt[0 + 3 * 0] = 1.0;
    if (imt == 2) mt[1 + 3 * 1] = 1.0;
    if (imt == 3) mt[2 + 3 * 2] = 1.0;
    if (imt == 4) mt[0 + 3 * 1] = 1.0;
    if (imt == 5) mt[0 + 3 * 2] = 1.0;
    if (imt == 6) mt[1 + 3 * 2] = 1.0;
}

/* utosynnSH  others.f90:132-186 */
static void o_utosynnSH(const int *imt, const zc *u, zc *s)
{
    if (*imt == 2) {
        s[0] = s[0] - 0.5 * u[1];
        s[1] = s[1] - 0.5 * u[1];
    }
    if (*imt == 3) {
        s[0] = s[0] - 0.5 * u[1];
        s[1] = s[1] + 0.5 * u[1];
    }
    if (*imt == 4) s[3] = s[3] + u[1];
    if (*imt == 5) s[4] = s[4] + u[2];
    if (*imt == 6) s[2] = s[2] - u[2];
}

#define U_(i, j) uder[((i) - 1) + 3 * ((j) - 1)]
/* udertorsgtSH  others.f90:196-245 (slots 6..10 of the P-SV layout) */
static void o_udertorsgtSH(const int *icomp, const zc *uder, zc *r)
{
    if (*icomp == 2) {
        r[0] = r[0] - 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
        r[1] = r[1] - 0.5 * U_(2, 2) + 0.5 * U_(3, 3);
        r[3] = r[3] - 0.5 * U_(1, 2) - 0.5 * U_(2, 1);
    }
    if (*icomp == 3) {
        r[2] = r[2] - 0.5 * U_(2, 3) - 0.5 * U_(3, 2);
        r[4] = r[4] + 0.5 * U_(1, 3) + 0.5 * U_(3, 1);
    }
}

/* udertotsgtSH  others.f90:249-332: ttsgt(1..10) = P-SV slots 7,12..20 */
static void o_udertotsgtSH(const int *imt, const zc *uder, zc *t)
{
    if (*imt == 2) {
        t[1] = t[1] - 0.25 * U_(2, 2) - 0.25 * U_(3, 3);
        t[6] = t[6] - 0.125 * U_(1, 2) - 0.125 * U_(2, 1);
        t[7] = t[7] + 0.125 * U_(1, 2) + 0.125 * U_(2, 1);
        t[8] = t[8] + 0.125 * U_(2, 2) - 0.125 * U_(3, 3);
        t[9] = t[9] + 0.125 * U_(2, 2) - 0.125 * U_(3, 3);
    }
    if (*imt == 3) {
        t[1] = t[1] + 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
        t[6] = t[6] - 0.125 * U_(1, 2) - 0.125 * U_(2, 1);
        t[7] = t[7] + 0.125 * U_(1, 2) + 0.125 * U_(2, 1);
        t[8] = t[8] - 0.125 * U_(2, 2) + 0.125 * U_(3, 3);
        t[9] = t[9] - 0.125 * U_(2, 2) + 0.125 * U_(3, 3);
    }
    if (*imt == 4) {
        t[0] = t[0] - 0.5 * U_(2, 2) - 0.5 * U_(3, 3);
        t[2] = t[2] - 0.25 * U_(1, 2) - 0.25 * U_(2, 1);
        t[3] = t[3] + 0.25 * U_(1, 2) + 0.25 * U_(2, 1);
        t[4] = t[4] + 0.25 * U_(2, 2) - 0.25 * U_(3, 3);
        t[5] = t[5] - 0.25 * U_(2, 2) + 0.25 * U_(3, 3);
    }
    if (*imt == 5) {
        t[2] = t[2] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[3] = t[3] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[4] = t[4] - 0.25 * U_(2, 3) - 0.25 * U_(3, 2);
        t[5] = t[5] - 0.25 * U_(2, 3) - 0.25 * U_(3, 2);
    }
    if (*imt == 6) {
        t[6] = t[6] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[7] = t[7] + 0.25 * U_(1, 3) + 0.25 * U_(3, 1);
        t[8] = t[8] - 0.25 * U_(2, 3) - 0.25 * U_(3, 2);
        t[9] = t[9] + 0.25 * U_(2, 3) + 0.25 * U_(3, 2);
    }
}

/* locallyCartesianDerivatives  others.f90:336-367 */
static void o_lcd(const zc *u, const zc *udr, const zc *udt, const zc *udp, zc *uder,
                  const double *r, const double *theta)
{
    if (*theta == 0.0 || *theta == PI) {
        U_(1, 1) = udr[0];
        U_(2, 2) = u[0] / *r;
        U_(3, 3) = u[0] / *r;
        return;
    }
    double thetasin = sin(*theta);
    double thetacot = cos(*theta) / thetasin;
    U_(1, 1) = udr[0];
    U_(1, 2) = (udt[0] - u[1]) / *r;
    U_(1, 3) = (udp[0] / thetasin - u[2]) / *r;
    U_(2, 1) = udr[1];
    U_(2, 2) = (udt[1] + u[0]) / *r;
    U_(2, 3) = (udp[1] / thetasin - u[2] * thetacot) / *r;
    U_(3, 1) = udr[2];
    U_(3, 2) = udt[2] / *r;
    U_(3, 3) = (udp[2] / thetasin + u[0] + u[1] * thetacot) / *r;
}
#undef U_

/* calgrid  others.f90:383-459 (SH: S velocity only) */
static void o_calgrid(const int *nzone, const double *vrmin, const double *vrmax, const double *vs,
                      const double *rmin, const double *rmax, const int *imax, const int *lmin,
                      const double *tlen, double *vmin, double *gridpar, double *dzpar)
{
    for (int iz = 1; iz <= *nzone; iz++) {
        double v[4], coef1 = 1.0, coef2 = 1.0, vs1 = 0.0, vs2 = 0.0;
        for (int i = 1; i <= 4; i++) v[i - 1] = vs[(i - 1) + 4 * (iz - 1)];
        for (int j = 1; j <= 4; j++) {
            if (j == 1) coef1 = 1.0;
            else coef1 = coef1 * (vrmin[iz - 1] / *rmax);
            if (j == 1) coef2 = 1.0;
            else coef2 = coef2 * (vrmax[iz - 1] / *rmax);
            vs1 = vs1 + v[j - 1] * coef1;
            vs2 = vs2 + v[j - 1] * coef2;
        }
        double rh = vrmax[iz - 1] - vrmin[iz - 1];
        double omega = 2.0 * PI * (double)*imax / *tlen;
        vmin[iz - 1] = (vs1 >= vs2) ? vs2 : vs1;
        double amax = vrmax[iz - 1];
        double gtmp = (omega * omega) / (vmin[iz - 1] * vmin[iz - 1]) -
                      (((double)*lmin + 0.5) * ((double)*lmin + 0.5)) / (amax * amax);
        if (gtmp > 0.0) {
            dzpar[iz - 1] = sqrt(1.0 / gtmp);
            gridpar[iz - 1] = rh / dzpar[iz - 1];
        } else {
            dzpar[iz - 1] = 0.0;
            gridpar[iz - 1] = 0.0;
        }
    }
    double gtmp = 0.0;
    for (int iz = 1; iz <= *nzone; iz++) gtmp = gtmp + gridpar[iz - 1];
    for (int iz = 1; iz <= *nzone; iz++) {
        if (gridpar[iz - 1] > 0.0) gridpar[iz - 1] = gridpar[iz - 1] / gtmp;
        else {
            double rh = vrmax[iz - 1] - vrmin[iz - 1];
            gridpar[iz - 1] = rh / (*rmax - *rmin) * 0.1;
        }
    }
    gtmp = 0.0;
    for (int iz = 1; iz <= *nzone; iz++) gtmp = gtmp + gridpar[iz - 1];
    for (int iz = 1; iz <= *nzone; iz++) gridpar[iz - 1] = gridpar[iz - 1] / gtmp;
}

/* calra  others.f90:463-495 */
static void o_calra(int *nlayer, const double *gridpar, const double *dzpar, const int *nzone,
                    const double *vrmin, const double *vrmax, const double *rmin, const double *rmax,
                    int *nnl, const double *re)
{
    (void)gridpar;
    (void)rmin;
    (void)rmax;
    for (int iz = 1; iz <= *nzone; iz++) {
        double rh = vrmax[iz - 1] - vrmin[iz - 1];
        int ntmp;
        if (dzpar[iz - 1] == 0.0) ntmp = 1;
        else ntmp = (int)(sqrt(3.3 / *re) * rh / dzpar[iz - 1] / 2.0 / PI / 7.0e-1 + 1);
        nnl[iz - 1] = ntmp;
        if (nnl[iz - 1] < 5) nnl[iz - 1] = 5;
    }
    int in = 0;
    for (int iz = 1; iz <= *nzone; iz++) in += nnl[iz - 1];
    *nlayer = in;
}

/* calra2  others.f90:500-563: node radii and receiver 3-node stencils */
static void o_calra2(const int *nlayer, const double *gridpar, const double *dzpar,
                     const int *nzone, const double *vrmin, const double *vrmax, const double *rmin,
                     const double *rmax, const int *nnl, double *ra, const double *re,
                     const int *nsta, const double *rsta, double *rrsta, int *iista)
{
    (void)gridpar;
    (void)dzpar;
    (void)rmax;
    (void)re;
    for (int i = 0; i < *nlayer + *nzone + 1; i++) ra[i] = 0.0;
    ra[0] = *rmin;
    int itmp = 1;
    for (int iz = 1; iz <= *nzone; iz++) {
        double rh = vrmax[iz - 1] - vrmin[iz - 1];
        for (int i = 1; i <= nnl[iz - 1]; i++) {
            itmp++;
            ra[itmp - 1] = vrmin[iz - 1] + rh * (double)i / (double)nnl[iz - 1];
        }
    }
    itmp = 1;
    for (int iz = 1; iz <= *nzone; iz++) {
        for (int i = 1; i <= nnl[iz - 1]; i++) {
            for (int ista = 1; ista <= *nsta; ista++) {
                if (ra[itmp - 1] < rsta[ista - 1] && rsta[ista - 1] <= ra[itmp]) {
                    int base = (i != nnl[iz - 1]) ? itmp : itmp - 1;
                    for (int k = 0; k < 3; k++) {
                        rrsta[k + 3 * (ista - 1)] = ra[base + k - 1];
                        iista[k + 3 * (ista - 1)] = base + k;
                    }
                }
            }
            itmp++;
        }
    }
}

/* calsp  others.f90:569-584 */
static void o_calsp(const int *ndc, const int *nlayer, int *isp, int *jsp)
{
    isp[0] = 1;
    jsp[0] = 1;
    for (int i = 1; i <= *ndc; i++) {
        isp[i] = isp[i - 1] + nlayer[i - 1];
        jsp[i] = jsp[i - 1] + 4 * nlayer[i - 1];
    }
}

/* calspo  others.f90:588-631: source layer and offset (r0 is clamped) */
static void o_calspo(const int *ndc, const double *rdc, const int *nlayer, double *r0,
                     const double *rmin, const double *rmax, const double *ra, const int *isp,
                     double *spo, int *spn)
{
    (void)rmin;
    int NL = *nlayer;
    if (*r0 == *rmax) {
        *spo = (double)NL - 0.01;
        *r0 = ra[NL - 1] + (*spo - (double)(NL - 1)) * (ra[NL] - ra[NL - 1]);
    } else {
        int itmp;
        for (itmp = 2; itmp <= *ndc + NL + 2; itmp++)
            if (*r0 < ra[itmp - 1]) break;
        *spo = (double)(itmp - 2) + (*r0 - ra[itmp - 2]) / (ra[itmp - 1] - ra[itmp - 2]);
        if (*spo - (double)(itmp - 2) < 0.01) {
            *spo = (double)(itmp - 2) + 0.01;
            *r0 = ra[itmp - 2] + (*spo - (double)(itmp - 2)) * (ra[itmp - 1] - ra[itmp - 2]);
        }
        if (*spo - (double)(itmp - 2) > 0.99) {
            *spo = (double)(itmp - 2) + 0.99;
            *r0 = ra[itmp - 2] + (*spo - (double)(itmp - 2)) * (ra[itmp - 1] - ra[itmp - 2]);
        }
    }
    *spn = 0;
    for (int itmp = 1; itmp <= *ndc + 1; itmp++) {
        *spn = *spn + 1;
        if (*r0 <= rdc[itmp - 1]) break;
    }
    *spo = *spo - (double)(isp[*spn - 1] - 1);
}

/* calgra  others.f90:637-647 */
static void o_calgra(const int *isp, const double *ra, const double *r0, const int *spn,
                     const double *spo, double *gra)
{
    int itmp = isp[*spn - 1] + (int)trunc(*spo);
    gra[0] = ra[itmp - 1];
    gra[1] = *r0;
    gra[2] = ra[itmp];
}

/* calstg  others.f90:651-694 */
static void o_calstg(const int *nzone, const double *rrho, const double *vsv, const double *vsh,
                     const int *nlayer, const int *nnl, const double *ra, const double *rmax,
                     int *vnp, double *vra, double *rho, double *ecL, double *ecN)
{
    int NT = *nlayer + 2 * *nzone + 1;
    for (int i = 0; i < NT; i++) vra[i] = rho[i] = ecL[i] = ecN[i] = 0.0;
    int itmp = 0, jtmp = 0;
    for (int iz = 1; iz <= *nzone; iz++) {
        for (int i = 1; i <= nnl[iz - 1] + 1; i++) {
            itmp++;
            jtmp++;
            vra[itmp - 1] = ra[jtmp - 1];
            double trho = 0.0, tvsv = 0.0, tvsh = 0.0, coef = 1.0;
            for (int j = 1; j <= 4; j++) {
                if (j == 1) coef = 1.0;
                else coef = coef * (vra[itmp - 1] / *rmax);
                trho = trho + rrho[(j - 1) + 4 * (iz - 1)] * coef;
                tvsv = tvsv + vsv[(j - 1) + 4 * (iz - 1)] * coef;
                tvsh = tvsh + vsh[(j - 1) + 4 * (iz - 1)] * coef;
            }
            rho[itmp - 1] = trho;
            ecL[itmp - 1] = rho[itmp - 1] * tvsv * tvsv;
            ecN[itmp - 1] = rho[itmp - 1] * tvsh * tvsh;
        }
        jtmp--;
    }
    *vnp = itmp;
}

/* calgstg  others.f90:699-742 */
static void o_calgstg(const int *nzone, const int *nlayer, const int *spn, const double *rrho,
                      const double *vsv, const double *vsh, const double *ra, double *vra,
                      const double *rmax, double *rho, double *ecL, double *ecN, const double *r0,
                      double *mu0)
{
    (void)nzone;
    (void)nlayer;
    (void)r0;
    int s = *spn;
    for (int i = 1; i <= 3; i++) {
        vra[i - 1] = ra[i - 1];
        double trho = 0.0, tvsv = 0.0, tvsh = 0.0, coef = 1.0;
        for (int j = 1; j <= 4; j++) {
            if (j == 1) coef = 1.0;
            else coef = coef * (vra[i - 1] / *rmax);
            trho = trho + rrho[(j - 1) + 4 * (s - 1)] * coef;
            tvsv = tvsv + vsv[(j - 1) + 4 * (s - 1)] * coef;
            tvsh = tvsh + vsh[(j - 1) + 4 * (s - 1)] * coef;
        }
        rho[i - 1] = trho;
        ecL[i - 1] = rho[i - 1] * tvsv * tvsv;
        ecN[i - 1] = rho[i - 1] * tvsh * tvsh;
    }
    *mu0 = ecL[1];
}

/* callsuf  others.f90:748-766 (sums the coefficients: the surface value) */
static void o_callsuf(const double *omega, const int *nzone, const double *vrmax, const double *vsv,
                      int *lsuf)
{
    double tvs = 0.0, coef = 1.0;
    for (int i = 1; i <= 4; i++) {
        if (i == 1) coef = 1.0;
        tvs = tvs + vsv[(i - 1) + 4 * (*nzone - 1)] * coef;
    }
    *lsuf = (int)(*omega * vrmax[*nzone - 1] / tvs - 0.5) + 1;
}

/* calcoef  others.f90:772-800 */
static void o_calcoef(const int *nzone, const double *omega, const double *q, zc *coef)
{
    for (int iz = 1; iz <= *nzone; iz++) {
        if (q[iz - 1] <= 0.0) coef[iz - 1] = 1.0;
        else {
            double aa = (*omega == 0.0) ? 1.0 : 1.0 + log(*omega / (2.0 * PI)) / (PI * q[iz - 1]);
            double bb = 1.0 / (2.0 * q[iz - 1]);
            zc c = CMPLX(aa, bb);
            coef[iz - 1] = c * c;
        }
    }
}

/* calamp trace (dsm_oracle.h) */
static int32_t *TR_L = NULL;
static double *TR_A = NULL;
static int32_t TR_CAP = 0, TR_N = 0;

int orc_calamp_trace(int32_t *l, double *amp, int32_t cap)
{
    TR_L = cap > 0 ? l : NULL;
    TR_A = cap > 0 ? amp : NULL;
    TR_CAP = cap > 0 ? cap : 0;
    TR_N = 0;
    return 0;
}

int orc_calamp_count(void) { return TR_N; }

void orc_trace_calamp(int l, const double complex *g, int n)
{
    if (!TR_CAP) return;
    double s = 0.0;
    for (int k = 0; k < n; k++) s += cabs(g[k]) * cabs(g[k]);
    if (TR_N < TR_CAP) {
        TR_L[TR_N] = l;
        TR_A[TR_N] = sqrt(s);
    }
    TR_N++;
}

/* calcutd (SH) others.f90:928-982 -- defined but never called by the
 * reference's SGTsh.  The real part of every value (cU = tmpr assigns the
 * complex value's real part), threshold rat * max, nc the FIRST node above
 * it (`exit`), kc = nc; kc = 1 when the threshold is zero. */
void orc_sh_calcutd(const int *nzone, const int *nnl, const zc *tmpr, const double *rat, const int *nn,
                    const double *ra, int *kc)
{
    (void)nnl;
    (void)ra;
    (void)nzone;
    const int n = *nn;
    double maxamp = -1.0;
    for (int i = 1; i <= n; i++)
        if (maxamp < creal(tmpr[i - 1])) maxamp = creal(tmpr[i - 1]);
    maxamp = maxamp * *rat;
    if (maxamp == 0.0) {
        *kc = 1;
        return;
    }
    int nc = 0;
    for (int i = 1; i <= n; i++)
        if (creal(tmpr[i - 1]) > maxamp) {
            nc = i;
            break;
        }
    *kc = nc;
}

/* calamp  others.f90:805-824 */
static void o_calamp(const zc *g, const int *l, const int *lsuf, double *maxamp, int *ismall,
                     const double *ratl)
{
    double ampratio = 0.0;
    double amp = cabs(*g);
    if (amp > *maxamp) *maxamp = amp;
    if (amp != 0.0 && *maxamp != 0.0) ampratio = amp / *maxamp;
    if (ampratio < *ratl && *l > *lsuf) *ismall = *ismall + 1;
    else *ismall = 0;
}

/* calu  others.f90:829-838 */
static void o_calu(const zc *c0, const double *lsq, const zc *bvec, zc *u)
{
    u[0] = 0.0;
    u[1] = u[1] + *c0 * bvec[1] / *lsq;
    u[2] = u[2] + *c0 * bvec[2] / *lsq;
}

/* interpolate  others.f90:843-891 (+ fillinpb 899-921) */
static void o_interpolate(const int *ncomp, const int *nderiv, const double *rsta,
                          const double *rrsta, const zc *g, zc *u)
{
    static double eps = -1.0;  /* data eps /-1.d0/: set to 3.52e-15 by glu */
    double dh[3];
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
    zc a[9], b[3], wk[3];
    int ip[3], ier;
    for (int i = 1; i <= 3; i++) {
        a[0 + 3 * (i - 1)] = 1.0;
        a[1 + 3 * (i - 1)] = dh[i - 1];
        a[2 + 3 * (i - 1)] = dh[i - 1] * dh[i - 1] / 2.0;
    }
    b[0] = b[1] = b[2] = 0.0;
    b[*nderiv] = 1.0;
    o_glu(a, 3, 3, b, &eps, wk, ip, &ier);
    for (int ic = 1; ic <= *ncomp; ic++)
        for (int i = 1; i <= 3; i++) u[ic - 1] = u[ic - 1] + b[i - 1] * g[*ncomp * (i - 1) + ic - 1];
}

/* ----------------------------------------------------------------------- */
/* src/SGTsh/calmat.f90:427-479 calg2 (excitation)                          */
/* ----------------------------------------------------------------------- */
static void o_calg2(const int *l, const int *m, const double *spo, const double *r0,
                    const double *mt, const double *mu0, const zc *coef, const zc *ga,
                    const zc *a, zc *ga2, zc *dr, zc *g2, int *ig2)
{
    static double eps = -1.0;  /* data eps /-1.d0/ */
#define MT(i, j) mt[((i) - 1) + 3 * ((j) - 1)]
    zc cg2[3] = {0, 0, 0}, z[3];
    zc dd = 0.0;
    double sgn = (*m >= 0) ? 1.0 : -1.0;
    int L = *l, M = *m;
    if (abs(M) == 1) {
        double b1 = sqrt((double)(2 * L + 1) / (16.0 * PI));
        dd = b1 * CMPLX(sgn * MT(1, 3), MT(1, 2)) / (CMPLX(*r0 * *r0 * *mu0, 0.0) * *coef);
        int itmp = 4;
        cg2[0] = cg2[1] = cg2[2] = 0.0;
        for (int i = 2; i <= 3; i++) {
            cg2[i - 1] = -dd * (ga[itmp] + ga[itmp + 1]);
            itmp += 2;
        }
    } else if (abs(M) == 2) {
        double b1 = sqrt((double)(2 * L + 1) * (double)(L - 1) * (double)(L + 2) / (64.0 * PI));
        cg2[1] = (b1 / *r0) * CMPLX(2.0 * MT(2, 3), sgn * (MT(2, 2) - MT(3, 3)));
    }
#undef MT
    int three = 3, one = 1, two = 2, ier;
    if ((M == -2 || M == -L) && *ig2 == 0) {
        o_dclisb0(ga2, &three, &one, &two, cg2, &eps, dr, z, &ier);
        *ig2 = *ig2 + 1;
    } else {
        o_dcsbsub0(ga2, &three, &one, &two, cg2, &eps, dr, z, &ier);
    }
    cg2[2] = cg2[2] + dd;
    int itmp = (int)trunc(*spo);
    g2[itmp] = a[0] * cg2[0] + a[1] * cg2[2];
    g2[itmp + 1] = a[2] * cg2[0] + a[3] * cg2[2];
}

/* ----------------------------------------------------------------------- */
/* src/SGTsh/trial.f90: Legendre recurrence and toroidal harmonics          */
/* ----------------------------------------------------------------------- */

/* calplm  trial.f90:85-117 (state plm(1:3) = P_l, P_{l-1}, P_{l-2}) */
static void o_calplm(int l, int m, double x, double *plm)
{
    if (l == m) {
        double pmm = 1.0;
        if (m > 0) {
            double somx2 = sqrt((1.0 - x) * (1.0 + x));
            double fact = 1.0;
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
/* calbvecphi0  trial.f90:1-81; plm(1:3,0:3), bvec(1:3,-2:2) */
static void o_calbvecphi0(const int *lp, const double *thetap, double *plm, zc *bvec, zc *bvecdt,
                          zc *bvecdp)
{
    int l = *lp;
    double theta = *thetap;
    if (theta == 0.0 || theta == PI) {
        for (int i = 0; i < 15; i++) bvec[i] = bvecdt[i] = bvecdp[i] = 0.0;
        return;
    }
    double x = cos(theta);
    double xl2 = (double)l * (double)(l + 1);
    for (int m = 0; m <= (l < 3 ? l : 3); m++) o_calplm(l, m, x, &plm[3 * m]);
    for (int m = 0; m <= (l < 2 ? l : 2); m++) {
        double fact = 1.0;
        if (m != 0)
            for (int i = l - m + 1; i <= l + m; i++) fact = fact * (double)i;
        double coef = sqrt((double)(2 * l + 1) / (4.0 * PI) / fact);
        double P = plm[3 * m], P1 = plm[3 * (m + 1)];
        double plmdt = (double)m * x / sin(theta) * P + P1;
        zc im = CMPLX(0.0, (double)m);
        BV(bvec, 1, m) = 0.0;
        BV(bvec, 1, -m) = 0.0;
        BV(bvec, 2, m) = im / sin(theta) * coef * P;
        BV(bvec, 2, -m) = conj(BV(bvec, 2, m));
        BV(bvec, 3, m) = -coef * plmdt;
        BV(bvec, 3, -m) = conj(BV(bvec, 3, m));
        BV(bvecdt, 1, m) = 0.0;
        BV(bvecdt, 1, -m) = 0.0;
        BV(bvecdt, 2, m) = im * (plmdt / sin(theta) - x / (1 - x * x) * P) * coef;
        BV(bvecdt, 2, -m) = conj(BV(bvecdt, 2, m));
        BV(bvecdt, 3, m) = (x / sin(theta) * plmdt - (double)m * (double)m / (1 - x * x) * P + xl2 * P) * coef;
        BV(bvecdt, 3, -m) = conj(BV(bvecdt, 3, m));
        BV(bvecdp, 1, m) = 0.0;
        BV(bvecdp, 1, -m) = 0.0;
        BV(bvecdp, 2, m) = -(double)m * (double)m / sin(theta) * P * coef;
        BV(bvecdp, 2, -m) = conj(BV(bvecdp, 2, m));
        BV(bvecdp, 3, m) = -im * plmdt * coef;
        BV(bvecdp, 3, -m) = conj(BV(bvecdp, 3, m));
        if (m % 2 == 1) {
            BV(bvec, 2, -m) = -BV(bvec, 2, -m);
            BV(bvec, 3, -m) = -BV(bvec, 3, -m);
            BV(bvecdt, 2, -m) = -BV(bvecdt, 2, -m);
            BV(bvecdt, 3, -m) = -BV(bvecdt, 3, -m);
            BV(bvecdp, 2, -m) = -BV(bvecdp, 2, -m);
            BV(bvecdp, 3, -m) = -BV(bvecdp, 3, -m);
        }
    }
}

/* calbveczero  trial.f90:120-157 */
static void o_calbveczero(const int *lp, zc *bvec)
{
    int l = *lp;
    double xl2 = (double)l * (double)(l + 1);
    for (int m = 0; m <= (l < 1 ? l : 1); m++) {
        double fact = 1.0;
        if (m != 0)
            for (int i = l - m + 1; i <= l + m; i++) fact = fact * (double)i;
        double coef = sqrt((double)(2 * l + 1) / (4.0 * PI) / fact);
        BV(bvec, 1, m) = 0.0;
        BV(bvec, 1, -m) = 0.0;
        BV(bvec, 2, m) = CMPLX(0.0, (double)m) * xl2 * coef / 2.0;
        BV(bvec, 2, -m) = conj(BV(bvec, 2, m));
        BV(bvec, 3, m) = -CMPLX((double)m, 0.0) * xl2 * coef / 2.0;
        BV(bvec, 3, -m) = conj(BV(bvec, 3, m));
        if (m % 2 == 1) {
            BV(bvec, 2, -m) = -BV(bvec, 2, -m);
            BV(bvec, 3, -m) = -BV(bvec, 3, -m);
        }
    }
}

/* ----------------------------------------------------------------------- */
/* Function table: restated routines or the reference's own (oracle/_ref)  */
/* ----------------------------------------------------------------------- */
typedef struct {
    void (*calgrid)(const int *, const double *, const double *, const double *, const double *,
                    const double *, const int *, const int *, const double *, double *, double *,
                    double *);
    void (*calra)(int *, const double *, const double *, const int *, const double *,
                  const double *, const double *, const double *, int *, const double *);
    void (*calra2)(const int *, const double *, const double *, const int *, const double *,
                   const double *, const double *, const double *, const int *, double *,
                   const double *, const int *, const double *, double *, int *);
    void (*calsp)(const int *, const int *, int *, int *);
    void (*calspo)(const int *, const double *, const int *, double *, const double *,
                   const double *, const double *, const int *, double *, int *);
    void (*calgra)(const int *, const double *, const double *, const int *, const double *,
                   double *);
    void (*calstg)(const int *, const double *, const double *, const double *, const int *,
                   const int *, const double *, const double *, int *, double *, double *,
                   double *, double *);
    void (*calgstg)(const int *, const int *, const int *, const double *, const double *,
                    const double *, const double *, double *, const double *, double *, double *,
                    double *, const double *, double *);
    void (*calmatc)(const int *, const int *, const double *, const double *, const int *,
                    const int *, const int *, const double *, double *, double *);
    void (*caltl)(const int *, const int *, const double *, const double *, const double *,
                  double *);
    void (*calhl)(const int *, const int *, const double *, const double *, const double *,
                  double *);
    void (*calt)(const int *, const double *, const double *, double *);
    void (*callsuf)(const double *, const int *, const double *, const double *, int *);
    void (*calcoef)(const int *, const double *, const double *, zc *);
    void (*cala0)(const int *, const double *, const double *, const double *, const double *,
                  const double *, const double *, const double *, const zc *, zc *);
    void (*cala2)(const int *, const double *, const zc *, zc *);
    void (*overlap)(const int *, const zc *, zc *);
    void (*cala)(const int *, const int *, const int *, const zc *, const zc *, zc *);
    void (*calga)(const int *, const double *, const double *, const int *, const double *,
                  const double *, const double *, const double *, const double *, const zc *,
                  zc *);
    void (*calg2)(const int *, const int *, const double *, const double *, const double *,
                  const double *, const zc *, const zc *, const zc *, zc *, zc *, zc *, int *);
    void (*dclisb0)(zc *, const int *, const int *, const int *, zc *, double *, zc *, zc *,
                    int *);
    void (*dcsbsub0)(zc *, const int *, const int *, const int *, zc *, double *, zc *, zc *,
                     int *);
    void (*calamp)(const zc *, const int *, const int *, double *, int *, const double *);
    void (*calbvecphi0)(const int *, const double *, double *, zc *, zc *, zc *);
    void (*calbveczero)(const int *, zc *);
    void (*calu)(const zc *, const double *, const zc *, zc *);
    void (*interpolate)(const int *, const int *, const double *, const double *, const zc *,
                        zc *);
    void (*lcd)(const zc *, const zc *, const zc *, const zc *, zc *, const double *,
                const double *);
    void (*udertotsgtSH)(const int *, const zc *, zc *);
    void (*udertorsgtSH)(const int *, const zc *, zc *);
    void (*utosynnSH)(const int *, const zc *, zc *);
} sh_fns_t;

static const sh_fns_t C_FNS = {
    o_calgrid, o_calra,  o_calra2,      o_calsp,        o_calspo,  o_calgra,  o_calstg,
    o_calgstg, o_calmatc, o_caltl,      o_calhl,        o_calt,    o_callsuf, o_calcoef,
    o_cala0,   o_cala2,  o_overlap,     o_cala,         o_calga,   o_calg2,   o_dclisb0,
    o_dcsbsub0, o_calamp, o_calbvecphi0, o_calbveczero, o_calu,    o_interpolate, o_lcd,
    o_udertotsgtSH, o_udertorsgtSH, o_utosynnSH};

static sh_fns_t REF_FNS;
static int REF_LOADED = 0;

int orc_sh_load_ref(const char *path)
{
    void *h = dlopen(path, RTLD_LAZY | RTLD_LOCAL);
    if (!h) return -1;
    void **dst = (void **)&REF_FNS;
    const char *names[] = {"calgrid_",  "calra_",      "calra2_",       "calsp_",
                           "calspo_",   "calgra_",     "calstg_",       "calgstg_",
                           "calmatc_",  "caltl_",      "calhl_",        "calt_",
                           "callsuf_",  "calcoef_",    "cala0_",        "cala2_",
                           "overlap_",  "cala_",       "calga_",        "calg2_",
                           "dclisb0_",  "dcsbsub0_",   "calamp_",       "calbvecphi0_",
                           "calbveczero_", "calu_",    "interpolate_",  "locallycartesianderivatives_",
                           "udertotsgtsh_", "udertorsgtsh_", "utosynnsh_"};
    int n = (int)(sizeof(names) / sizeof(names[0]));
    if (n * sizeof(void *) != sizeof(REF_FNS)) return -2;
    for (int k = 0; k < n; k++) {
        dst[k] = dlsym(h, names[k]);
        if (!dst[k]) return -3 - k;
    }
    REF_LOADED = 1;
    return 0;
}

/* ----------------------------------------------------------------------- */
/* Driver: SGTsh-MPI.f90                                                    */
/* ----------------------------------------------------------------------- */
typedef struct {
    int nzone;
    double vrmin[ORC_MAXZONE], vrmax[ORC_MAXZONE], rrho[ORC_MAXZONE * 4], vsv[ORC_MAXZONE * 4],
        vsh[ORC_MAXZONE * 4], qmu[ORC_MAXZONE];
    double rmin, rmax, omegai, tlen, ratl;
    int maxlmax, synnswitch, rsgtswitch;
    int r_n, theta_n;
    double *r_, *rrsta, *theta;
    int *iista;
    double r0;
    int intir0;
    int nnlayer, nlayer[ORC_MAXZONE], isp[ORC_MAXZONE], jsp[ORC_MAXZONE];
    double *ra;
    double spo;
    int spn;
    double gra[3], gvra[3], grho[3], gecL[3], gecN[3], mu0;
    int vnp;
    double *vra, *rho, *ecL, *ecN;
    double *t, *h1, *h2, *h3, *h4;
    double gt[8], gh1[8], gh2[8], gh3[8], gh4[8];
    int nn, ns, ins;
} sh_setup_t;

static void sh_free(sh_setup_t *s)
{
    free(s->r_);
    free(s->rrsta);
    free(s->theta);
    free(s->iista);
    free(s->ra);
    free(s->vra);
    free(s->rho);
    free(s->ecL);
    free(s->ecN);
    free(s->t);
    free(s->h1);
    free(s->h2);
    free(s->h3);
    free(s->h4);
}

/* SGTsh-MPI.f90:66-348 */
static int sh_setup(const orc_input_t *in, const sh_fns_t *F, sh_setup_t *s)
{
    memset(s, 0, sizeof(*s));
    /* remove the fluid core and everything below (SGTsh-MPI.f90:83-137) */
    int ioc = 0;
    for (int i = 1; i <= in->nzone; i++)
        if (in->vsv[i - 1][0] == 0.0 && in->vsv[i - 1][1] == 0.0 && in->vsv[i - 1][2] == 0.0 &&
            in->vsv[i - 1][3] == 0.0)
            ioc = i;
    s->nzone = in->nzone - ioc;
    if (s->nzone < 1) return -1;
    for (int i = 1; i <= s->nzone; i++) {
        int src = ioc + i - 1;
        s->vrmin[i - 1] = in->vrmin[src];
        s->vrmax[i - 1] = in->vrmax[src];
        for (int j = 0; j < 4; j++) {
            s->rrho[j + 4 * (i - 1)] = in->rho[src][j];
            s->vsv[j + 4 * (i - 1)] = in->vsv[src][j];
            s->vsh[j + 4 * (i - 1)] = in->vsh[src][j];
        }
        s->qmu[i - 1] = in->qmu[src];
    }
    s->rmin = s->vrmin[0];
    s->rmax = s->vrmax[s->nzone - 1];
    s->tlen = in->tlen;
    s->omegai = -log(in->omegai_seed) / in->tlen;
    s->ratl = in->ratl;
    s->maxlmax = in->maxlmax;
    s->synnswitch = in->synnswitch;
    s->rsgtswitch = in->rsgtswitch;
    /* receivers above the CMB (SGTsh-MPI.f90:152-179) */
    double rmin_ = in->rmin_;
    int r_n = in->r_list_n > 0 ? in->r_list_n : (int)((in->rmax_ - in->rmin_) / in->rdelta_) + 1;
    int ir_ = 0;  /* (a radius list: its entries at or below the CMB, the list ascending) */
    for (int i = 1; i <= r_n; i++)
        if ((in->r_list_n > 0 ? in->r_list[i - 1] : rmin_ + (double)(i - 1) * in->rdelta_) <= s->rmin) ir_ = i;
    r_n = r_n - ir_;
    rmin_ = rmin_ + (double)ir_ * in->rdelta_;
    if (r_n < 1) return -2;
    s->r_n = r_n;
    s->r_ = calloc(r_n, sizeof(double));
    s->rrsta = calloc(3 * r_n, sizeof(double));
    s->iista = calloc(3 * r_n, sizeof(int));
    for (int i = 1; i <= r_n; i++)
        s->r_[i - 1] = in->r_list_n > 0 ? in->r_list[ir_ + i - 1] : rmin_ + (double)(i - 1) * in->rdelta_;
    s->theta_n = in->theta_list_n > 0 ? in->theta_list_n : (int)((in->thetamax - in->thetamin) / in->thetadelta) + 1;
    s->theta = calloc(s->theta_n, sizeof(double));
    for (int i = 1; i <= s->theta_n; i++)
        s->theta[i - 1] = in->theta_list_n > 0 ? in->theta_list[i - 1] : in->thetamin + (double)(i - 1) * in->thetadelta;
    s->r0 = in->r0min;
    s->intir0 = (int)(s->r0 * 1.0e3);
    if (s->r0 < s->rmin || s->r0 > s->rmax) return -3;

    /* grid (SGTsh-MPI.f90:276-287) */
    double vmin[ORC_MAXZONE], gridpar[ORC_MAXZONE], dzpar[ORC_MAXZONE];
    int one = 1, ndc = s->nzone - 1;
    F->calgrid(&s->nzone, s->vrmin, s->vrmax, s->vsv, &s->rmin, &s->rmax, &in->imax, &one, &s->tlen,
               vmin, gridpar, dzpar);
    F->calra(&s->nnlayer, gridpar, dzpar, &s->nzone, s->vrmin, s->vrmax, &s->rmin, &s->rmax,
             s->nlayer, &in->re);
    s->ra = calloc(s->nnlayer + s->nzone + 1, sizeof(double));
    for (int i = 0; i < 3 * r_n; i++) s->iista[i] = 1;
    F->calra2(&s->nnlayer, gridpar, dzpar, &s->nzone, s->vrmin, s->vrmax, &s->rmin, &s->rmax,
              s->nlayer, s->ra, &in->re, &s->r_n, s->r_, s->rrsta, s->iista);
    F->calsp(&ndc, s->nlayer, s->isp, s->jsp);
    F->calspo(&ndc, s->vrmax, &s->nnlayer, &s->r0, &s->rmin, &s->rmax, s->ra, s->isp, &s->spo,
              &s->spn);
    F->calgra(s->isp, s->ra, &s->r0, &s->spn, &s->spo, s->gra);

    /* matrix elements (SGTsh-MPI.f90:291-339) */
    int NT = s->nnlayer + 2 * s->nzone + 1, N4 = 4 * s->nnlayer;
    s->vra = calloc(NT, sizeof(double));
    s->rho = calloc(NT, sizeof(double));
    s->ecL = calloc(NT, sizeof(double));
    s->ecN = calloc(NT, sizeof(double));
    s->t = calloc(N4, sizeof(double));
    s->h1 = calloc(N4, sizeof(double));
    s->h2 = calloc(N4, sizeof(double));
    s->h3 = calloc(N4, sizeof(double));
    s->h4 = calloc(N4, sizeof(double));
    double *work = calloc(N4 + 8, sizeof(double));
    F->calstg(&s->nzone, s->rrho, s->vsv, s->vsh, &s->nnlayer, s->nlayer, s->ra, &s->rmax, &s->vnp,
              s->vra, s->rho, s->ecL, s->ecN);
    F->calgstg(&s->nzone, &s->nnlayer, &s->spn, s->rrho, s->vsv, s->vsh, s->gra, s->gvra, &s->rmax,
               s->grho, s->gecL, s->gecN, &s->r0, &s->mu0);
    int two = 2, zero = 0;
    for (int i = 1; i <= s->nzone; i++) {
        int *nl = &s->nlayer[i - 1];
        const double *rai = &s->ra[s->isp[i - 1] - 1];
        int j0 = s->jsp[i - 1] - 1;
        F->calmatc(nl, &s->vnp, s->vra, s->rho, &two, &zero, &zero, rai, &s->t[j0], &work[j0]);
        F->calmatc(nl, &s->vnp, s->vra, s->ecL, &two, &one, &one, rai, &s->h1[j0], &work[j0]);
        F->calmatc(nl, &s->vnp, s->vra, s->ecL, &one, &one, &zero, rai, &s->h2[j0], &work[j0]);
        F->calmatc(nl, &s->vnp, s->vra, s->ecL, &zero, &zero, &zero, rai, &s->h3[j0], &work[j0]);
        F->calmatc(nl, &s->vnp, s->vra, s->ecN, &zero, &zero, &zero, rai, &s->h4[j0], &work[j0]);
        F->caltl(nl, &s->vnp, s->vra, s->rho, rai, &work[j0]);
        F->calt(nl, &s->t[j0], &work[j0], &s->t[j0]);
        F->calhl(nl, &s->vnp, s->vra, s->ecL, rai, &work[j0]);
        F->calt(nl, &s->h3[j0], &work[j0], &s->h3[j0]);
        F->calhl(nl, &s->vnp, s->vra, s->ecN, rai, &work[j0]);
        F->calt(nl, &s->h4[j0], &work[j0], &s->h4[j0]);
    }
    int three = 3;
    F->calmatc(&two, &three, s->gvra, s->grho, &two, &zero, &zero, s->gra, s->gt, work);
    F->calmatc(&two, &three, s->gvra, s->gecL, &two, &one, &one, s->gra, s->gh1, work);
    F->calmatc(&two, &three, s->gvra, s->gecL, &one, &one, &zero, s->gra, s->gh2, work);
    F->calmatc(&two, &three, s->gvra, s->gecL, &zero, &zero, &zero, s->gra, s->gh3, work);
    F->calmatc(&two, &three, s->gvra, s->gecN, &zero, &zero, &zero, s->gra, s->gh4, work);
    F->caltl(&two, &three, s->gvra, s->grho, s->gra, work);
    F->calt(&two, s->gt, work, s->gt);
    F->calhl(&two, &three, s->gvra, s->gecL, s->gra, work);
    F->calt(&two, s->gh3, work, s->gh3);
    F->calhl(&two, &three, s->gvra, s->gecN, s->gra, work);
    F->calt(&two, s->gh4, work, s->gh4);
    free(work);
    s->nn = s->nnlayer + 1;
    s->ns = s->isp[s->spn - 1] + (int)trunc(s->spo);
    s->ins = 4 * s->ns - 3;
    return 0;
}

int orc_sh_info(const orc_input_t *in, orc_sh_info_t *info)
{
    sh_setup_t s;
    int rc = sh_setup(in, &C_FNS, &s);
    if (rc == 0) {
        info->r_n = s.r_n;
        info->theta_n = s.theta_n;
        info->nn = s.nn;
        info->nzone = s.nzone;
        info->r0 = s.r0;
        info->spo = s.spo;
        info->spn = s.spn;
        info->intir0 = s.intir0;
    }
    sh_free(&s);
    return rc;
}

/* Mixed table: bit k of mask selects the reference routine for entry k of
 * sh_fns_t (a test aid that localises a disagreement to one routine). */
static sh_fns_t MIX_FNS;
int orc_sh_mix(uint64_t mask)
{
    if (!REF_LOADED) return -10;
    void **dst = (void **)&MIX_FNS;
    void *const *c = (void *const *)&C_FNS, *const *r = (void *const *)&REF_FNS;
    for (size_t k = 0; k < sizeof(MIX_FNS) / sizeof(void *); k++) dst[k] = ((mask >> k) & 1) ? r[k] : c[k];
    return 0;
}

/* One frequency: SGTsh-MPI.f90:389-560.  use_ref: 0 restated routines,
 * 1 reference routines, 2 the mixed table of orc_sh_mix. */
int orc_sh_omega(const orc_input_t *in, int32_t i, int32_t use_ref, double complex *tsgt,
                 double complex *rsgt, double complex *synn, int32_t *llog_out)
{
    if (use_ref && !REF_LOADED) return -10;
    const sh_fns_t *F = use_ref == 1 ? &REF_FNS : use_ref == 2 ? &MIX_FNS : &C_FNS;
    sh_setup_t S, *s = &S;
    int rc = sh_setup(in, F, s);
    if (rc) {
        sh_free(s);
        return rc;
    }
    const int r_n = s->r_n, theta_n = s->theta_n, nn = s->nn, lda = 2;
    memset(tsgt, 0, sizeof(zc) * 10 * r_n * theta_n);
    memset(rsgt, 0, sizeof(zc) * 5 * r_n * theta_n);
    memset(synn, 0, sizeof(zc) * 5 * theta_n);
    double eps = -1.0;
    double omega = 2.0 * PI * (double)i / s->tlen;
    int lsuf;
    zc coef[ORC_MAXZONE];
    F->callsuf(&omega, &s->nzone, s->vrmax, s->vsv, &lsuf);
    F->calcoef(&s->nzone, &omega, s->qmu, coef);
    double *plm = calloc(12 * (size_t)theta_n, sizeof(double));
    zc *bvec = calloc(15 * (size_t)theta_n, sizeof(zc));
    zc *bvecdt = calloc(15 * (size_t)theta_n, sizeof(zc));
    zc *bvecdp = calloc(15 * (size_t)theta_n, sizeof(zc));
    zc *a0 = calloc(2 * (size_t)nn, sizeof(zc)), *a2 = calloc(2 * (size_t)nn, sizeof(zc));
    zc *a = calloc(2 * (size_t)nn, sizeof(zc));
    zc *cwork = calloc(4 * (size_t)s->nnlayer, sizeof(zc));
    zc *g0 = calloc(nn, sizeof(zc)), *dr = calloc(nn, sizeof(zc)), *z = calloc(nn, sizeof(zc));
    for (int j = 1; j <= s->nzone; j++) {
        int j0 = s->jsp[j - 1] - 1, i0 = s->isp[j - 1] - 1;
        F->cala0(&s->nlayer[j - 1], &omega, &s->omegai, &s->t[j0], &s->h1[j0], &s->h2[j0],
                 &s->h3[j0], &s->h4[j0], &coef[j - 1], &cwork[j0]);
        F->overlap(&s->nlayer[j - 1], &cwork[j0], &a0[2 * i0]);
        F->cala2(&s->nlayer[j - 1], &s->h4[j0], &coef[j - 1], &cwork[j0]);
        F->overlap(&s->nlayer[j - 1], &cwork[j0], &a2[2 * i0]);
    }
    int ismall = 0, llog = s->maxlmax, ier;
    double maxamp = -1.0;
    zc aa[4], ga[8], ga2[6], gdr[3], rvec[15];
    double mt[9];
    int one = 1, two = 2, zero = 0;
    const int sp0 = s->isp[s->spn - 1] - 1;
    for (int l = 0; l <= s->maxlmax; l++) {
        double lsq = sqrt((double)l * (double)(l + 1));
        for (int it = 0; it < theta_n; it++) {
            double th = s->theta[it] / 180.0 * PI;
            F->calbvecphi0(&l, &th, &plm[12 * it], &bvec[15 * it], &bvecdt[15 * it], &bvecdp[15 * it]);
        }
        memset(rvec, 0, sizeof(rvec));
        F->calbveczero(&l, rvec);
        if (ismall > 20) {
            if (llog > l) llog = l;
            continue;
        }
        memset(a, 0, sizeof(zc) * 2 * nn);
        memset(ga2, 0, sizeof(ga2));
        F->cala(&nn, &l, &lda, a0, a2, a);
        F->calga(&one, &omega, &s->omegai, &l, &s->t[s->ins - 1], &s->h1[s->ins - 1],
                 &s->h2[s->ins - 1], &s->h3[s->ins - 1], &s->h4[s->ins - 1], &coef[s->spn - 1], aa);
        F->calga(&two, &omega, &s->omegai, &l, s->gt, s->gh1, s->gh2, s->gh3, s->gh4,
                 &coef[s->spn - 1], ga);
        F->overlap(&two, ga, ga2);
        for (int m = -2; m <= 2; m++) {
            if (!(m != 0 && abs(m) <= abs(l))) continue;
            int ig2 = 0;
            for (int imt = 2; imt <= 6; imt++) {
                memset(g0, 0, sizeof(zc) * nn);
                o_setmt(imt, mt);
                F->calg2(&l, &m, &s->spo, &s->r0, mt, &s->mu0, &coef[s->spn - 1], ga, aa, ga2, gdr,
                         &g0[sp0], &ig2);
                if (m == -2 || m == -l) {
                    if (ig2 == 1) {
                        F->dclisb0(a, &nn, &one, &lda, g0, &eps, dr, z, &ier);
                        ig2++;
                    } else {
                        F->dcsbsub0(a, &nn, &one, &lda, g0, &eps, dr, z, &ier);
                    }
                } else {
                    F->dcsbsub0(a, &nn, &one, &lda, g0, &eps, dr, z, &ier);
                }
                if (imt == 3) {
                    orc_trace_calamp(l, &g0[nn - 1], 1);
                    F->calamp(&g0[nn - 1], &l, &lsuf, &maxamp, &ismall, &s->ratl);
                }
                if (s->synnswitch == 1) {
                    for (int it = 0; it < theta_n; it++) {
                        zc u[3] = {0, 0, 0};
                        F->calu(&g0[nn - 1], &lsq, &bvec[15 * it + 3 * (m + 2)], u);
                        F->utosynnSH(&imt, u, &synn[5 * it]);
                    }
                }
                for (int ir = 0; ir < r_n; ir++) {
                    zc gs[3], g0tmp = 0, g0dertmp = 0;
                    for (int k = 0; k < 3; k++) gs[k] = g0[s->iista[3 * ir + k] - 1];
                    F->interpolate(&one, &zero, &s->r_[ir], &s->rrsta[3 * ir], gs, &g0tmp);
                    F->interpolate(&one, &one, &s->r_[ir], &s->rrsta[3 * ir], gs, &g0dertmp);
                    for (int it = 0; it < theta_n; it++) {
                        zc u[3] = {0}, udr[3] = {0}, udt[3] = {0}, udp[3] = {0}, uder[9] = {0};
                        const int bo = 15 * it + 3 * (m + 2);
                        F->calu(&g0tmp, &lsq, &bvec[bo], u);
                        F->calu(&g0dertmp, &lsq, &bvec[bo], udr);
                        F->calu(&g0tmp, &lsq, &bvecdt[bo], udt);
                        F->calu(&g0tmp, &lsq, &bvecdp[bo], udp);
                        double th = s->theta[it] / 180.0 * PI;
                        F->lcd(u, udr, udt, udp, uder, &s->r_[ir], &th);
                        F->udertotsgtSH(&imt, uder, &tsgt[10 * (ir + r_n * it)]);
                    }
                }
            }
            if (abs(m) == 1) {
                for (int icomp = 2; icomp <= 3; icomp++) {
                    memset(g0, 0, sizeof(zc) * nn);
                    g0[nn - 1] = -conj(rvec[(icomp - 1) + 3 * (m + 2)]) / CMPLX(lsq, 0.0);
                    F->dcsbsub0(a, &nn, &one, &lda, g0, &eps, dr, z, &ier);
                    for (int ir = 0; ir < r_n; ir++) {
                        zc gs[3], g0tmp = 0, g0dertmp = 0;
                        for (int k = 0; k < 3; k++) gs[k] = g0[s->iista[3 * ir + k] - 1];
                        F->interpolate(&one, &zero, &s->r_[ir], &s->rrsta[3 * ir], gs, &g0tmp);
                        F->interpolate(&one, &one, &s->r_[ir], &s->rrsta[3 * ir], gs, &g0dertmp);
                        for (int it = 0; it < theta_n; it++) {
                            zc u[3] = {0}, udr[3] = {0}, udt[3] = {0}, udp[3] = {0}, uder[9] = {0};
                            const int bo = 15 * it + 3 * (m + 2);
                            F->calu(&g0tmp, &lsq, &bvec[bo], u);
                            F->calu(&g0dertmp, &lsq, &bvec[bo], udr);
                            F->calu(&g0tmp, &lsq, &bvecdt[bo], udt);
                            F->calu(&g0tmp, &lsq, &bvecdp[bo], udp);
                            double th = s->theta[it] / 180.0 * PI;
                            F->lcd(u, udr, udt, udp, uder, &s->r_[ir], &th);
                            F->udertorsgtSH(&icomp, uder, &rsgt[5 * (ir + r_n * it)]);
                        }
                    }
                }
            }
        }
    }
    *llog_out = llog - 1;
    free(plm);
    free(bvec);
    free(bvecdt);
    free(bvecdp);
    free(a0);
    free(a2);
    free(a);
    free(cwork);
    free(g0);
    free(dr);
    free(z);
    sh_free(s);
    return 0;
}
