/*
 * km_oracle.c -- CPU restatement of KernelMaker's catalog stage and of its
 * frequency-to-time transform (src/KernelMaker/fourier.f90).
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle of the GPU catalog stage.
 * The product library never links or calls this file.
 *
 * Arrays follow the reference's Fortran layouts (column-major, 1-based
 * index arithmetic written as [i-1]):
 *   sgtomega(1:num, fmin:fmax, 1:theta_n)   complex(8)
 *   sgtF(1:num, fmin:fmax)                  complex(8)
 *   h3(1:6, fmin:fmax), h4(1:6, fmin:fmax)  complex(8)
 *
 * The reference forms its reduction-time phase factors with
 * exp(cmplx(0.d0, x)), whose CMPLX has the default (single-precision) kind:
 * the argument is rounded to float and the factor is a complex(4) value,
 * exp(0) (cos, sin) of the float argument in float arithmetic.  km_phase()
 * takes cosf / sinf of the C library, which gives the reference's values
 * bit for bit where the reference is built against the same library.
 */
#include "dsm_oracle.h"

#include <math.h>
#include <string.h>

typedef double complex zc;
static const double KM_PI = 3.1415926535897932; /* module parameters, module.f90 */

/* exp(cmplx(0.d0, x)) of the reference: complex(4) */
static zc km_phase(double x)
{
    const float xf = (float)x;
    const float c = cosf(xf), s = sinf(xf);
    return CMPLX((double)c, (double)s);
}

/* rdsgtomega, fourier.f90:1-123, after the read: the records of one file
 * kind for frequencies fmin..fmax (rec: complex(4) a(num_sgt, theta_n) per
 * frequency, consecutive; i = 0 is not read and stays zero) added into the
 * P-SV-slot catalog `omega` (num_psv components): SYNN's off-diagonal
 * slots halved (:79-87), the reduction-time phase and 1e-21 (:93-104), the
 * SH -> P-SV slot map (:106-117).  ipsvorsh: 2 / 1 SYNN P-SV / SH, 20 / 10
 * RSGT, 200 / 100 TSGT. */
void orc_km_rdsgt(int ipsvorsh, int num_sgt, int num_psv, int fmin, int fmax, int theta_n, const double *thetaD,
                  double tlen, double c_red_reci, const float complex *rec, double complex *omega)
{
    const int nf = fmax - fmin + 1;
    zc sg[32];
    for (int i = fmin; i <= fmax; i++) {
        if (i == 0) continue;
        const float complex *r = rec + (size_t)(i - fmin) * num_sgt * theta_n;
        for (int it = 1; it <= theta_n; it++) {
            for (int c = 0; c < num_sgt; c++) sg[c] = (zc)r[(size_t)(it - 1) * num_sgt + c];
            if (ipsvorsh == 1)
                for (int c = 3; c <= 5; c++) sg[c - 1] = sg[c - 1] * 5.0e-1;
            if (ipsvorsh == 2) {
                sg[1] = sg[1] * 5.0e-1;
                for (int c = 8; c <= 10; c++) sg[c - 1] = sg[c - 1] * 5.0e-1;
            }
            const double redtime = -thetaD[it - 1] * c_red_reci;
            const zc ph = km_phase(-2.0 * KM_PI * (double)i / tlen * redtime);
            for (int c = 0; c < num_sgt; c++) sg[c] = sg[c] * ph * 1.0e-21;
            zc *o = omega + (size_t)num_psv * ((i - fmin) + (size_t)nf * (it - 1));
            if (ipsvorsh == 2 || ipsvorsh == 20 || ipsvorsh == 200) {
                for (int c = 0; c < num_psv; c++) o[c] = o[c] + sg[c];
            } else if (ipsvorsh == 1 || ipsvorsh == 10) {
                for (int c = 6; c <= 10; c++) o[c - 1] = o[c - 1] + sg[c - 6];
            } else if (ipsvorsh == 100) {
                o[6] = o[6] + sg[0];
                for (int c = 12; c <= 20; c++) o[c - 1] = o[c - 1] + sg[c - 11];
            }
        }
    }
}

/* findIdeltaRegion, fourier.f90:166-190 */
static void km_region(int ix, int nhalf, int nmax, int *istart, int *iend)
{
    *istart = ix - nhalf;
    *iend = ix + nhalf;
    if (*istart < 1) {
        *iend = *iend + (1 - *istart);
        *istart = 1;
    }
    if (*iend > nmax) {
        *istart = *istart - (*iend - nmax);
        *iend = nmax;
    }
    if (*istart < 1) *istart = 1;
    if (*iend > nmax) *iend = nmax;
}

/* clsgt, fourier.f90:128-162: the catalog at `distance` (degrees) by sinc
 * interpolation over the int(ipdistance / thetadelta / 2) distances on
 * either side, normalised by the sum of the weights, times the inverse
 * reduction-time phase; added into sgtF(1:num, fmin:fmax). */
void orc_km_clsgt(double distance, int num, int fmin, int fmax, int theta_n, const double *thetaD, double thetadelta,
                  double ipdistance, double tlen, double c_red_reci, const double complex *omega, double complex *sgtF)
{
    const int nf = fmax - fmin + 1;
    const int idelta = 1 + (int)((distance - thetaD[0]) / thetadelta);
    const int nhregion = (int)(ipdistance / thetadelta / 2.0);
    int is, ie;
    km_region(idelta, nhregion, theta_n, &is, &ie);
    double norm = 0.0;
    for (int kk = is; kk <= ie; kk++) {
        const double x = KM_PI * ((distance - thetaD[is - 1]) / thetadelta - (double)(kk - is));
        const double w = (x == 0.0) ? 1.0 : sin(x) / x;
        const zc *src = omega + (size_t)num * nf * (kk - 1);
        for (int j = 0; j < nf; j++)
            for (int c = 0; c < num; c++) {
                const size_t k = (size_t)num * j + c;
                sgtF[k] = (x == 0.0) ? sgtF[k] + src[k] : sgtF[k] + w * src[k];
            }
        norm = norm + w;
    }
    for (int j = fmin; j <= fmax; j++) {
        const zc ph = km_phase(-2.0 * KM_PI * (double)j / tlen * distance * c_red_reci);
        for (int c = 0; c < num; c++) {
            const size_t k = (size_t)num * (j - fmin) + c;
            sgtF[k] = sgtF[k] / norm * ph;
        }
    }
}

/* rsgt2h3freq, fourier.f90:278-318: RSGT -> 3rd-order tensor of the
 * receiver component compo (0 Z, 1 R, 2 T); trig = crq, srq, crq2, srq2.
 * (T: the reference's h3(3) = -h3(3) negates what h3(3) held before; h3 is
 * an in/out array here for that reason, zero for KernelMaker's points
 * since it zeroes h3 at every radius, KernelMaker-1.0.0.f90:722.) */
void orc_km_rsgt2h3(int compo, int fmin, int fmax, const double complex *rsgtF, const double *trig,
                    double complex *h3)
{
    const double crq = trig[0], srq = trig[1], crq2 = trig[2], srq2 = trig[3];
    for (int f = 0; f <= fmax - fmin; f++) {
        const zc *r = rsgtF + (size_t)10 * f;
        zc *h = h3 + (size_t)6 * f;
        if (compo == 0) {
            h[0] = r[0];
            h[1] = r[2] * crq2 - r[3];
            h[2] = -r[2] * crq2 - r[3];
            h[3] = 2.0 * r[1] * crq;
            h[4] = 2.0 * r[1] * srq;
            h[5] = 2.0 * r[2] * srq2;
        } else if (compo == 1) {
            h[0] = r[4];
            h[1] = -(r[6]) * crq2 - r[5];
            h[2] = (r[6]) * crq2 - r[5];
            h[3] = 4.0 * (r[8]) * crq;
            h[4] = 4.0 * (r[8]) * srq;
            h[5] = -2.0 * (r[6]) * srq2;
        } else {
            h[0] = 0.0;
            h[1] = (r[7]) * srq2;
            h[2] = -h[2];
            h[3] = 2.0 * (r[9]) * srq;
            h[4] = -2.0 * (r[9]) * crq;
            h[5] = -2.0 * (r[7]) * crq2;
        }
    }
}

/* tsgt2h4freq, fourier.f90:320-415: TSGT -> the 4th-order tensor double
 * dotted with the moment tensor mt (rr, tt, pp, rt, rp, tp).  trig = cqs,
 * sqs, csq, ssq, cqs2, sqs2, csq2, ssq2. */
void orc_km_tsgt2h4(int fmin, int fmax, const double complex *tsgtF, const double *trig, const double *mt,
                    double complex *h4)
{
    const double cqs = trig[0], sqs = trig[1], csq = trig[2], ssq = trig[3];
    const double cqs2 = trig[4], sqs2 = trig[5], csq2 = trig[6], ssq2 = trig[7];
    for (int f = 0; f <= fmax - fmin; f++) {
        const zc *t = tsgtF + (size_t)20 * f - 1; /* t[k]: tsgtF(k) */
        zc *h = h4 + (size_t)6 * f;
        const zc h01 = t[7] * cqs, h02 = t[7] * sqs, h03 = t[8] * csq, h04 = t[8] * ssq;
        const zc h05 = t[9] * cqs2, h06 = t[10] * csq2, h07 = t[11] * csq2, h08 = t[11] * ssq2;
        const zc h09 = t[12] * cqs2, h10 = t[12] * sqs2;
        const zc h11 = t[13] * (csq * cqs - ssq * sqs), h12 = t[13] * (ssq * cqs + csq * sqs);
        const zc h13 = t[14] * (csq * cqs + ssq * sqs), h14 = t[14] * (ssq * cqs - csq * sqs);
        const zc h15 = t[15] * (csq2 * cqs - ssq2 * sqs), h16 = t[15] * (ssq2 * cqs + csq2 * sqs);
        const zc h17 = t[17] * (csq * cqs2 - ssq * sqs2), h18 = t[17] * (ssq * cqs2 + csq * sqs2);
        const zc h19 = t[19] * (csq2 * cqs2 - ssq2 * sqs2), h20 = t[19] * (ssq2 * cqs2 + csq2 * sqs2);
        const zc h21 = t[20] * (csq2 * cqs2 + ssq2 * sqs2), h22 = t[20] * (ssq2 * cqs2 - csq2 * sqs2);
        const zc h23 = t[16] * (csq2 * cqs + ssq2 * sqs), h24 = t[16] * (ssq2 * cqs - csq2 * sqs);
        const zc h25 = t[18] * (csq * cqs2 + ssq * sqs2), h26 = t[18] * (ssq * cqs2 - csq * sqs2);
        h[0] = t[1] * mt[0] + (h05 - t[2]) * mt[1] + (-h05 - t[2]) * mt[2] - t[5] * (cqs * mt[3] + sqs * mt[4]) +
               t[9] * sqs2 * mt[5];
        h[1] = (h06 - t[3]) * mt[0] + (h21 + h19 - h07 - h09 + t[4]) * mt[1] + (-h21 - h19 - h07 + h09 + t[4]) * mt[2] +
               (h23 - h15 + h01) * mt[3] + (-h24 - h16 + h02) * mt[4] + (-h22 + h20 - h10) * mt[5];
        h[2] = (-h06 - t[3]) * mt[0] + (-h21 - h19 + h07 - h09 + t[4]) * mt[1] + (h21 + h19 + h07 + h09 + t[4]) * mt[2] +
               (-h23 + h15 + h01) * mt[3] + (h24 + h16 + h02) * mt[4] + (h22 - h20 - h10) * mt[5];
        h[3] = -t[6] * csq * mt[0] + (h25 - h17 + h03) * mt[1] + (-h25 + h17 + h03) * mt[2] + (-h13 + h11) * mt[3] +
               (h14 + h12) * mt[4] + (-h26 - h18) * mt[5];
        h[4] = (-t[6] * ssq) * mt[0] + (h26 - h18 + h04) * mt[1] + (-h26 + h18 + h04) * mt[2] + (-h14 + h12) * mt[3] +
               (-h13 - h11) * mt[4] + (h25 + h17) * mt[5];
        h[5] = t[10] * ssq2 * mt[0] + (h22 + h20 - h08) * mt[1] + (-h22 - h20 - h08) * mt[2] + (h24 - h16) * mt[3] +
               (h23 + h15) * mt[4] + (h21 - h19) * mt[5];
    }
}

/* bitrv2, fourier.f90:568-617 */
static void km_bitrv2(int n, double *a)
{
    const int m = n / 4, m2 = 2 * m, n2 = n - 2;
    int k = 0;
    for (int j = 0; j <= m2 - 4; j += 4) {
        double xr, xi;
        if (j < k) {
            xr = a[j];
            xi = a[j + 1];
            a[j] = a[k];
            a[j + 1] = a[k + 1];
            a[k] = xr;
            a[k + 1] = xi;
        } else if (j > k) {
            const int j1 = n2 - j, k1 = n2 - k;
            xr = a[j1];
            xi = a[j1 + 1];
            a[j1] = a[k1];
            a[j1 + 1] = a[k1 + 1];
            a[k1] = xr;
            a[k1 + 1] = xi;
        }
        const int k1 = m2 + k;
        xr = a[j + 2];
        xi = a[j + 3];
        a[j + 2] = a[k1];
        a[j + 3] = a[k1 + 1];
        a[k1] = xr;
        a[k1 + 1] = xi;
        int l = m;
        while (k >= l) {
            k = k - l;
            l = l / 2;
        }
        k = k + l;
    }
}

/* cdft, fourier.f90:481-565: in-place complex DFT of n/2 points c, twiddles
 * from (wr, wi) = (cos, sin)(pi / (n/2)) by the reference's recurrences */
void orc_km_cdft(int n, double wr, double wi, double complex *c, double *a)
{
    for (int i = 0; i < n / 2; i++) {
        a[2 * i] = creal(c[i]);
        a[2 * i + 1] = cimag(c[i]);
    }
    double wmr = wr, wmi = wi;
    int m = n;
    while (m > 4) {
        const int l = m / 2;
        double wkr = 1, wki = 0;
        double wdr = 1 - 2 * wmi * wmi;
        double wdi = 2 * wmi * wmr;
        const double ss = 2 * wdi;
        wmr = wdr;
        wmi = wdi;
        for (int j = 0; j <= n - m; j += m) {
            const int i = j + l;
            double xr = a[j] - a[i], xi = a[j + 1] - a[i + 1];
            a[j] = a[j] + a[i];
            a[j + 1] = a[j + 1] + a[i + 1];
            a[i] = xr;
            a[i + 1] = xi;
            xr = a[j + 2] - a[i + 2];
            xi = a[j + 3] - a[i + 3];
            a[j + 2] = a[j + 2] + a[i + 2];
            a[j + 3] = a[j + 3] + a[i + 3];
            a[i + 2] = wdr * xr - wdi * xi;
            a[i + 3] = wdr * xi + wdi * xr;
        }
        for (int k = 4; k <= l - 4; k += 4) {
            wkr = wkr - ss * wdi;
            wki = wki + ss * wdr;
            wdr = wdr - ss * wki;
            wdi = wdi + ss * wkr;
            for (int j = k; j <= n - m + k; j += m) {
                const int i = j + l;
                double xr = a[j] - a[i], xi = a[j + 1] - a[i + 1];
                a[j] = a[j] + a[i];
                a[j + 1] = a[j + 1] + a[i + 1];
                a[i] = wkr * xr - wki * xi;
                a[i + 1] = wkr * xi + wki * xr;
                xr = a[j + 2] - a[i + 2];
                xi = a[j + 3] - a[i + 3];
                a[j + 2] = a[j + 2] + a[i + 2];
                a[j + 3] = a[j + 3] + a[i + 3];
                a[i + 2] = wdr * xr - wdi * xi;
                a[i + 3] = wdr * xi + wdi * xr;
            }
        }
        m = l;
    }
    if (m > 2) {
        for (int j = 0; j <= n - 4; j += 4) {
            const double xr = a[j] - a[j + 2], xi = a[j + 1] - a[j + 3];
            a[j] = a[j] + a[j + 2];
            a[j + 1] = a[j + 1] + a[j + 3];
            a[j + 2] = xr;
            a[j + 3] = xi;
        }
    }
    if (n > 4) km_bitrv2(n, a);
    for (int i = 0; i < n / 2; i++) c[i] = CMPLX(a[2 * i], a[2 * i + 1]);
}

/* vectorFFT_double, fourier.f90:418-447: the spectrum ccvec(imin:imax)
 * extended to 2 np1 points with c(np1 + i) = conj(c(np1 - i)), i =
 * imin..np1-1, transformed by cdft, and rvec(i) = Re c(i) exp(omegai t_i) /
 * tlen * 1e3 for the window i = iws..iwe (t_i = i tlen / (2 np1)).
 * work: 2 np1 complex + 4 np1 doubles. */
void orc_km_vectorfft(int imin, int imax, int np1, const double complex *ccvec, double *rvec, double omegai,
                      double tlen, int iws, int iwe, double complex *cvec, double *a)
{
    const double pi = 3.141592653589793; /* fourier.f90:427 */
    memset(cvec, 0, sizeof(zc) * 2 * (size_t)np1);
    for (int i = imin; i <= imax; i++) cvec[i] = ccvec[i - imin];
    const double samplingHz = (double)(2 * np1) / tlen;
    for (int i = imin; i <= np1 - 1; i++) cvec[np1 + i] = conj(cvec[np1 - i]);
    orc_km_cdft(4 * np1, cos(pi / (2 * np1)), sin(pi / (2 * np1)), cvec, a);
    for (int i = iws; i <= iwe; i++)
        rvec[i - iws] = creal(cvec[i]) * exp(omegai * (double)i / samplingHz) / tlen * 1.0e3;
}
