// legendre_lane.h -- per-distance body of the Legendre / theta-basis table
// (legendre.hip), shared with the host test build (hostcheck.cpp).
#pragma once

#include "dev_common.h"

namespace dsm {

#ifndef DSM_PI_DEFINED
#define DSM_PI_DEFINED
static constexpr double DSM_PI = 3.1415926535897932;
#endif

// Upward recurrence of calplm (src/SGTsh/trial.f90:85-117).
__host__ __device__ inline void calplm(int l, int m, double x, double *p)
{
    if (l == m) {
        double pmm = 1.0;
        if (m > 0) {
            const double somx2 = sqrt((1.0 - x) * (1.0 + x));
            double fact = 1.0;
            for (int i = 1; i <= m; i++) {
                pmm = -pmm * fact * somx2;
                fact = fact + 2.0;
            }
        }
        p[2] = 0.0;
        p[1] = 0.0;
        p[0] = pmm;
    } else {
        p[2] = p[1];
        p[1] = p[0];
        if (l == m + 1) p[0] = x * (double)(2 * m + 1) * p[1];
        else p[0] = (x * (double)(2 * l - 1) * p[1] - (double)(l + m - 1) * p[2]) / (double)(l - m);
    }
}

// SH basis, |m| = mu in {1,2}, b = 0..5:
//   Y0 = c P/s, Y1 = c dP, Y2 = c (dP/s - x/(1-x^2) P),
//   Y3 = c (x/s dP - mu^2/(1-x^2) P + l(l+1) P),
//   Z23 = -mu^2 Y0/s + cot Y1,  Z33 = cot Y0 - Y1/s,
// with c = sqrt((2l+1)/(4 pi) (l-mu)!/(l+mu)!), dP = mu x/s P_l^mu + P_l^{mu+1}
// (trial.f90:1-81).  Zero at the poles (trial.f90:151-164) and for l < mu.
__host__ __device__ inline void legendre_sh_column(const double *theta_rad, int theta_n, int theta_pad, int lmax,
                                                   int lcap, double *tab, int it)
{
    if (it >= theta_pad) return;
    const size_t plane = (size_t)lcap * theta_pad;
    if (it >= theta_n) {  // padding columns
        for (int b = 0; b < 12; b++)
            for (int l = 0; l < lcap; l++) tab[(size_t)b * plane + (size_t)l * theta_pad + it] = 0.0;
        return;
    }
    const double th = theta_rad[it];
    const bool pole = (th == 0.0 || th == DSM_PI);
    const double x = cos(th), s = sin(th);
    const double cot = x / s;
    double plm[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int l = 0; l <= lmax; l++) {
        if (!pole)
            for (int m = 0; m <= (l < 3 ? l : 3); m++) calplm(l, m, x, plm[m]);
        const double xl2 = (double)l * (double)(l + 1);
        for (int mu = 1; mu <= 2; mu++) {
            double y[6] = {0, 0, 0, 0, 0, 0};
            if (!pole && l >= mu) {
                double fact = 1.0;
                for (int i = l - mu + 1; i <= l + mu; i++) fact = fact * (double)i;
                const double c = sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / fact);
                const double P = plm[mu][0], P1 = plm[mu + 1][0];
                const double dP = (double)mu * x / s * P + P1;
                const double mu2 = (double)(mu * mu);
                y[0] = c * P / s;
                y[1] = c * dP;
                y[2] = c * (dP / s - x / (1 - x * x) * P);
                y[3] = c * (x / s * dP - mu2 / (1 - x * x) * P + xl2 * P);
                y[4] = -mu2 * y[0] / s + cot * y[1];
                y[5] = cot * y[0] - y[1] / s;
            }
            for (int b = 0; b < 6; b++)
                tab[(size_t)((mu - 1) * 6 + b) * plane + (size_t)l * theta_pad + it] = y[b];
        }
    }
    for (int l = lmax + 1; l < lcap; l++)
        for (int b = 0; b < 12; b++) tab[(size_t)b * plane + (size_t)l * theta_pad + it] = 0.0;
}

// P-SV basis (caldvecphi0_withplm, src/SGTpsv/trial.f90:97-223, with the
// single-precision sin/cot of locallyCartesianDerivatives, others.f90:
// 310-344): for |m| = mu, c = sqrt((2l+1)/(4 pi) (l-mu)!/(l+mu)!),
// Pt = mu x/s P_l^mu + P_l^{mu+1}, sf = float(s), cf = float(x/s):
//   F0 = c P                F1 = c Pt
//   F2 = (-x/s Pt + mu^2/(1-x^2) P - l(l+1) P) c
//   F3 = (-x/(1-x^2) P + Pt/s) c          F4 = ((1/s) c) P
//   F5 = F0/sf   F6 = F1/sf   F7 = F4 cf   F8 = (((1/s) P) c)/sf   F9 = F1 cf
// Planes: mu=0 -> F0, F1, F2, F9 (planes 0..3); mu=1,2 -> F0..F9 (planes
// 4+10(mu-1)+k).  At the poles only the mu=0 F0 survives, as the
// default-kind cmplx(coeff)*(+-1) of the reference (trial.f90:112-131).
constexpr int PSV_NPLANES = 24;
__host__ __device__ inline int psv_plane(int mu, int f)
{
    if (mu == 0) return f == 9 ? 3 : f;
    return 4 + 10 * (mu - 1) + f;
}

// F0..F9 of degree l at one distance from pl[m] = P_l^m (m = 0..3; the
// recurrence's values, zero for m > l), s = sin(theta), x = cos(theta).
__host__ __device__ inline void legendre_psv_values(int l, double th, bool pole, double x, double s, double sf,
                                                    double cf, const double pl[4], double f[3][10])
{
    for (int mu = 0; mu < 3; mu++)
        for (int k = 0; k < 10; k++) f[mu][k] = 0.0;
    if (pole) {
        const double sign1 = (th == DSM_PI && (l & 1)) ? -1.0 : 1.0;
        f[0][0] = (double)(float)sqrt((double)(2 * l + 1) / (4.0 * DSM_PI)) * sign1;
        return;
    }
    const double xl2 = (double)l * (double)(l + 1);
    for (int mu = 0; mu <= (l < 2 ? l : 2); mu++) {
        double fact = 1.0;
        if (mu != 0)
            for (int i = l - mu + 1; i <= l + mu; i++) fact = fact * (double)i;
        const double c = sqrt((double)(2 * l + 1) / (4.0 * DSM_PI) / fact);
        const double P = pl[mu];
        const double Pt = (double)mu * x / s * P + pl[mu + 1];
        double *F = f[mu];
        F[0] = c * P;
        F[1] = c * Pt;
        F[2] = (-x / s * Pt + (double)mu * (double)mu / (1 - x * x) * P - xl2 * P) * c;
        F[3] = (-x / (1 - x * x) * P + 1.0 / s * Pt) * c;
        F[4] = 1.0 / s * c * P;
        F[5] = P * c / sf;
        F[6] = Pt * c / sf;
        F[7] = F[4] * cf;
        F[8] = 1.0 / s * P * c / sf;
        F[9] = c * Pt * cf;
    }
}

// The 24 planes of degree l at distance column it.
__host__ __device__ inline void legendre_psv_store(const double f[3][10], double *tab, size_t plane, int l,
                                                   int theta_pad, int it)
{
    for (int k = 0; k < 4; k++) {
        const int fk = (k == 3) ? 9 : k;
        tab[(size_t)psv_plane(0, fk) * plane + (size_t)l * theta_pad + it] = f[0][fk];
    }
    for (int mu = 1; mu <= 2; mu++)
        for (int k = 0; k < 10; k++) tab[(size_t)psv_plane(mu, k) * plane + (size_t)l * theta_pad + it] = f[mu][k];
}

__host__ __device__ inline void legendre_psv_column(const double *theta_rad, int theta_n, int theta_pad, int lmax,
                                                    int lcap, double *tab, int it)
{
    if (it >= theta_pad) return;
    const size_t plane = (size_t)lcap * theta_pad;
    if (it >= theta_n) {
        for (int b = 0; b < PSV_NPLANES; b++)
            for (int l = 0; l < lcap; l++) tab[(size_t)b * plane + (size_t)l * theta_pad + it] = 0.0;
        return;
    }
    const double th = theta_rad[it];
    const bool pole = (th == 0.0 || th == DSM_PI);
    const double x = cos(th), s = sin(th);
    const double sf = (double)(float)s, cf = (double)(float)(x / s);
    double plm[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int l = 0; l <= lmax; l++) {
        if (!pole)
            for (int m = 0; m <= (l < 3 ? l : 3); m++) calplm(l, m, x, plm[m]);
        const double pl[4] = {plm[0][0], plm[1][0], plm[2][0], plm[3][0]};
        double f[3][10];
        legendre_psv_values(l, th, pole, x, s, sf, cf, pl, f);
        legendre_psv_store(f, tab, plane, l, theta_pad, it);
    }
    for (int l = lmax + 1; l < lcap; l++)
        for (int b = 0; b < PSV_NPLANES; b++) tab[(size_t)b * plane + (size_t)l * theta_pad + it] = 0.0;
}

}  // namespace dsm
