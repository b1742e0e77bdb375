// fe_poly.h -- exact piecewise-polynomial element integrals of the DSM
// (src/SGTsh/calmat.f90:2-262, src/SGTpsv/calmat.f90:1-241), restated in
// the reference's operation order.  Only included by files compiled with
// -ffp-contract=off (see fe_setup.h).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>

namespace dsm {

using cd = std::complex<double>;

// (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re): the textbook product,
// which is what the reference's complex(8) arithmetic evaluates.
static inline cd cmul(cd a, cd b)
{
    return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

// Complex division by Smith's method, the reference's (gfortran's complex
// division and the GNU runtime's __divdc3 on finite operands), written out:
// the __divdc3 a shared library binds depends on its linker (clang's
// runtime scales by powers of two and rounds differently), and the device
// assembly (dev_common.h cdiv) uses these operations.
static inline cd cdiv(cd p, cd q)
{
    const double a = p.real(), b = p.imag(), c = q.real(), d = q.imag();
    if (std::fabs(c) < std::fabs(d)) {
        const double ratio = c / d;
        const double denom = (c * ratio) + d;
        return cd(((a * ratio) + b) / denom, ((b * ratio) - a) / denom);
    }
    const double ratio = d / c;
    const double denom = (d * ratio) + c;
    return cd(((b * ratio) + a) / denom, (b - (a * ratio)) / denom);
}

// pmulti: c = a*b for polynomials of n and m coefficients (l = n+m-1).
static inline void pmulti(int n, const double *a, int m, const double *b, int l, double *c)
{
    for (int i = 0; i < l; i++) c[i] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) c[i + j] = c[i + j] + a[i] * b[j];
}

// polint: integral of the polynomial p (n coefficients) over [x1,x2].
static inline double polint(int n, const double *p, double x1, double x2)
{
    double a[6], b[6];
    a[0] = 1.0;
    b[0] = 1.0;
    for (int i = 1; i < n; i++) {
        a[i] = a[i - 1] * x1;
        b[i] = b[i - 1] * x2;
    }
    double dx = x2 - x1, pint = 0.0;
    for (int i = 1; i <= n; i++) {
        double xx = 0.0;
        for (int j = 1; j <= i; j++) xx = xx + a[j - 1] * b[i - j] / (double)i;
        pint = pint + p[i - 1] * dx * xx;
    }
    return pint;
}

// pinteg: integral of p(r)*con(r) over [from,to], con linear between the
// structure nodes vra(1:vnp) (zero-length intervals at zone boundaries are
// skipped by the search).  snp carries the search position between calls.
static inline double pinteg(int &snp, int n, const double *p, double from, double to, int vnp,
                            const double *vra, const double *con)
{
    double q[2], pq[6], x1 = from, x2, pint = 0.0;
    if (snp >= vnp) snp = 1;
    for (;;) {
        if (vra[snp - 1] <= x1 && vra[snp] > x1) {
            x2 = std::min(to, vra[snp]);
        } else {
            snp = snp + 1;
            if (snp == vnp) snp = 1;
            continue;
        }
        q[1] = (con[snp] - con[snp - 1]) / (vra[snp] - vra[snp - 1]);
        q[0] = -q[1] * vra[snp - 1] + con[snp - 1];
        pmulti(n, p, 2, q, n + 1, pq);
        pint = pint + polint(n + 1, pq, x1, x2);
        if (x2 == to) break;
        x1 = x2;
    }
    return pint;
}

// calmatc: per-layer 2x2 blocks of \int r^rpow con W_p^(w1dn) W_q^(w2dn) dr,
// 4 values per layer; symmetrised when w1dn != w2dn.
static inline void calmatc(int nlayer, int vnp, const double *vra, const double *con, int rpow,
                           int w1dn, int w2dn, const double *ra, double *m, double *work)
{
    int snp = 1;
    for (int i = 1; i <= nlayer; i++) {
        double rh = ra[i] - ra[i - 1];
        double a[2][2], b[2][2];  // [column][row]
        if (w1dn == 0) {
            a[0][0] = ra[i] / rh;
            a[0][1] = -1.0 / rh;
            a[1][0] = -ra[i - 1] / rh;
            a[1][1] = 1.0 / rh;
        } else {
            a[0][0] = -1.0 / rh;
            a[0][1] = 0.0;
            a[1][0] = 1.0 / rh;
            a[1][1] = 0.0;
        }
        if (w2dn == 0) {
            b[0][0] = ra[i] / rh;
            b[0][1] = -1.0 / rh;
            b[1][0] = -ra[i - 1] / rh;
            b[1][1] = 1.0 / rh;
        } else {
            b[0][0] = -1.0 / rh;
            b[0][1] = 0.0;
            b[1][0] = 1.0 / rh;
            b[1][1] = 0.0;
        }
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
                double c[5] = {0, 0, 0, 0, 0};
                pmulti(2, a[j], 2, b[k], 3, c);
                for (int l = 3; l >= 1; l--) {
                    c[l + rpow - 1] = c[l - 1];
                    if (rpow > 0) c[l - 1] = 0.0;
                }
                work[4 * (i - 1) + 2 * j + k] = pinteg(snp, 5, c, ra[i - 1], ra[i], vnp, vra, con);
            }
    }
    if (w1dn != w2dn) {
        for (int i = 1; i <= 4 * nlayer; i++) {
            if (i % 4 == 0 || i % 4 == 1) m[i - 1] = 2.0 * work[i - 1];
            else if (i % 4 == 2) m[i - 1] = work[i - 1] + work[i];
            else m[i - 1] = work[i - 2] + work[i - 1];
        }
    } else {
        for (int i = 0; i < 4 * nlayer; i++) m[i] = work[i];
    }
}

// P-SV calmatc (src/SGTpsv/calmat.f90:1-62): the same integrals, no
// symmetrisation.
static inline void calmatc_psv(int nlayer, int vnp, const double *vra, const double *con, int rpow, int w1dn,
                               int w2dn, const double *ra, double *m)
{
    int snp = 1;
    for (int i = 1; i <= nlayer; i++) {
        double rh = ra[i] - ra[i - 1];
        double a[2][2], b[2][2];
        if (w1dn == 0) {
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
        if (w2dn == 0) {
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
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) {
                double c[5] = {0, 0, 0, 0, 0};
                pmulti(2, a[j], 2, b[k], 3, c);
                for (int l = 3; l >= 1; l--) {
                    c[l + rpow - 1] = c[l - 1];
                    if (rpow > 0) c[l - 1] = 0.0;
                }
                m[4 * (i - 1) + 2 * j + k] = pinteg(snp, 5, c, ra[i - 1], ra[i], vnp, vra, con);
            }
    }
}

// caltl: lumped mass (\int r^2 rho over each half layer).
static inline void caltl(int nlayer, int vnp, const double *vra, const double *rho, const double *ra,
                         double *tl)
{
    int snp = 1;
    const double c[3] = {0.0, 0.0, 1.0};
    for (int i = 1; i <= nlayer; i++) {
        double from = ra[i - 1], to = (ra[i - 1] + ra[i]) / 2.0;
        int nn = 4 * (i - 1);
        tl[nn] = pinteg(snp, 3, c, from, to, vnp, vra, rho);
        tl[nn + 1] = 0.0;
        tl[nn + 2] = 0.0;
        tl[nn + 3] = pinteg(snp, 3, c, to, ra[i], vnp, vra, rho);
    }
}

// calhl: lumped rigidity (\int mu over each half layer).
static inline void calhl(int nlayer, int vnp, const double *vra, const double *mu, const double *ra,
                         double *hl)
{
    int snp = 1;
    const double c[1] = {1.0};
    for (int i = 1; i <= nlayer; i++) {
        double from = ra[i - 1], to = (ra[i - 1] + ra[i]) / 2.0;
        int nn = 4 * (i - 1);
        hl[nn] = pinteg(snp, 1, c, from, to, vnp, vra, mu);
        hl[nn + 1] = 0.0;
        hl[nn + 2] = 0.0;
        hl[nn + 3] = pinteg(snp, 1, c, to, ra[i], vnp, vra, mu);
    }
}

// calt: average of the lumped and consistent matrices.
static inline void calt(int nlayer, const double *tl, const double *tc, double *t)
{
    for (int i = 0; i < 4 * nlayer; i++) t[i] = (tl[i] + tc[i]) / 2.0;
}

// SH overlap (calmat.f90:386-423): per-layer 2x2 blocks -> band a2(2,nlayer+1),
// a2(1,j) = A(j-1,j), a2(2,j) = A(j,j); adds into a2.
static inline void overlap2(int nlayer, const cd *a, cd *a2)
{
    const int nsize = nlayer + 1;
    for (int j = 1; j <= nsize; j++) {
        const int i1 = std::max(1, j - 1);
        for (int i = i1; i <= j; i++) {
            const int k = i - j + 2;
            cd &dst = a2[(k - 1) + 2 * (j - 1)];
            if (i == j) {
                if (i == 1) dst = dst + a[0];
                else if (i == nsize) dst = dst + a[4 * nlayer - 1];
                else dst = dst + a[4 * i - 5] + a[4 * i - 4];
            }
            if (i + 1 == j) dst = dst + a[4 * i - 3];
        }
    }
}

}  // namespace dsm
