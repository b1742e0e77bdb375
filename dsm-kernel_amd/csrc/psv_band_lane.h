// psv_band_lane.h -- one band column of the per-frequency P-SV band
// assembly (psv_band.hip): the operations of psv_omega_band
// (fe_setup_psv.cpp) for the column's (at most two) zones, host- and
// device-callable so the host build can check them against the host
// set-up.  See psv_band.hip for the scheme.
#pragma once

#include "dev_common.h"
#include "fe_setup.h"

namespace dsm {

// Band pieces of nslots frequencies (tables of psv_band_tables):
// band [slot0 + t][nn][4][P0, P1, Q0, Q1].
struct PsvBandArgs {
    const double *fe;
    const long long *zbase;
    const int *zsolid, *znl, *zkkdr;
    const int *colz;         // [nn][2]
    const int *ovr_ii;       // [nn]
    const double2 *ovr;      // [nn]
    int nz, nn;
    const double2 *scal;     // [t][1 + 3 nz]: (omega - i omegai)^2, then per zone coef1, coef2, 1/coef2
    int slot0;
    double2 *band;
};

namespace band_detail {
__host__ __device__ inline double2 sc(double2 c, double x) { return make_double2(c.x * x, c.y * x); }

struct BandZone {
    const double *f;  // the zone's block of element integrals
    int nl;
    double2 c1, c2, cf, om2;
    // array a (4 nl) at 1-based n; hm array b (4 (nl + 1)) at offset idx
    __host__ __device__ double F(int a, int n) const { return f[(size_t)a * 4 * nl + (n - 1)]; }
    __host__ __device__ double M(int b, int idx) const
    {
        return f[(size_t)PSV_FE_NSOLID4 * 4 * nl + (size_t)b * 4 * (nl + 1) + idx];
    }
};

// cala0's element entry k (1-based, 1..16 nl)
__host__ __device__ inline double2 cw0(const BandZone &z, int k)
{
    const double2 c1 = z.c1, c2 = z.c2;
    if ((k & 3) == 1) {
        const int n = (k + 3) / 4;
        if (n % 4 == 3) return cz();
        const double tt = z.F(FE_T, n);
        double2 hh0 = cadd(sc(c2, 4.0 * z.F(FE_H1X, n)), sc(c1, 16.0 / 3.0 * z.F(FE_H2N, n)));
        hh0 = cadd(hh0, sc(c1, -4.0 * z.F(FE_H2N, n)));
        hh0 = cadd(hh0, sc(c2, 2.0 * (z.F(FE_H3AY, n) + z.F(FE_H5AY, n))));
        hh0 = csub(hh0, sc(c1, 4.0 / 3.0 * (z.F(FE_H4AN, n) + z.F(FE_H6AN, n))));
        hh0 = cadd(hh0, sc(c2, 3.0 * z.F(FE_H7Z, n) - 2.0 * z.F(FE_H7Y, n)));
        hh0 = cadd(hh0, sc(c1, 4.0 / 3.0 * z.F(FE_H8N, n)));
        return csub(sc(z.om2, tt), hh0);
    }
    if ((k & 3) == 0) {
        const int n = k / 4;
        if (n % 4 == 3) return cz();
        const double tt = z.F(FE_T, n);
        double2 hh0 = cadd(sc(c1, z.F(FE_H2L, n)), sc(c1, -2.0 * z.F(FE_H2N, n)));
        hh0 = csub(hh0, sc(c1, z.F(FE_H4AL, n) + z.F(FE_H6AL, n)));
        hh0 = cadd(hh0, sc(c1, z.F(FE_H8L, n)));
        return csub(sc(z.om2, tt), hh0);
    }
    return cz();
}

// cala1's element entry k
__host__ __device__ inline double2 cw1(const BandZone &z, int k)
{
    const double2 c1 = z.c1, c2 = z.c2;
    if ((k & 3) == 2) {
        const int n = (k + 2) / 4;
        if (n % 4 == 3) return cz();
        double2 hh1 = csub(cneg(sc(c2, 2.0 * z.F(FE_H1X, n))), sc(c1, 8.0 / 3.0 * z.F(FE_H2N, n)));
        hh1 = csub(hh1, sc(c1, z.F(FE_H2L, n)));
        hh1 = csub(hh1, sc(c1, -2.0 * z.F(FE_H2N, n)));
        hh1 = csub(hh1, sc(c2, z.F(FE_H3Y, n)));
        hh1 = csub(hh1, sc(c1, -2.0 / 3.0 * z.F(FE_H4N, n)));
        hh1 = cadd(hh1, sc(c1, z.F(FE_H6L, n)));
        return cneg(hh1);
    }
    if ((k & 3) == 3) {
        const int n = (k + 1) / 4;
        if (n % 4 != 2) return cz();
        double2 hh1 = csub(cneg(sc(c2, 2.0 * z.F(FE_H1X, n))), sc(c1, 8.0 / 3.0 * z.F(FE_H2N, n)));
        hh1 = csub(hh1, sc(c1, -2.0 * z.F(FE_H2N, n)));
        hh1 = csub(hh1, sc(c1, z.F(FE_H2L, n)));
        hh1 = cadd(hh1, sc(c1, z.F(FE_H4L, n)));
        hh1 = csub(hh1, sc(c2, z.F(FE_H5Y, n)));
        hh1 = cadd(hh1, sc(c1, 2.0 / 3.0 * z.F(FE_H6N, n)));
        return cneg(hh1);
    }
    return cz();
}

// cala2's element entry k
__host__ __device__ inline double2 cw2(const BandZone &z, int k)
{
    if ((k & 3) == 1) {
        const int n = (k + 3) / 4;
        if (n % 4 == 3) return cz();
        return cneg(sc(z.c1, z.F(FE_H2L, n)));
    }
    if ((k & 3) == 0) {
        const int n = k / 4;
        if (n % 4 == 3) return cz();
        const double2 hh2 = cadd(sc(z.c2, z.F(FE_H1X, n)), sc(z.c1, 4.0 / 3.0 * z.F(FE_H2N, n)));
        return cneg(hh2);
    }
    return cz();
}

// overlapa's a2(ii, i) from the element entries F(k) = W(z, k): at most one
// assignment reaches each (ii, i), a plain entry or the sum of two
template <class W>
__host__ __device__ inline double2 overlapa_at(const BandZone &z, int ii, int i, W w)
{
    const int nsize = 2 * (z.nl + 1), msize = 16 * z.nl;
    double2 v = cz();
    if (ii == 4 && i == 1) v = w(z, 1);
    if (ii == 3 && i == 2) v = w(z, 2);
    if (ii == 4 && i == 2) v = w(z, 4);
    if (ii == 4 && i == nsize - 1) v = w(z, msize - 3);
    if (ii == 3 && i == nsize) v = w(z, msize - 2);
    if (ii == 4 && i == nsize) v = w(z, msize);
    if ((i & 1) && i >= 3 && i <= nsize - 1) {
        const int n1 = 16 * (i - 3) / 2 + 5;
        if (ii == 2) v = w(z, n1);
        if (ii == 3) v = w(z, n1 + 2);
        if (ii == 4 && i != nsize - 1) v = cadd(w(z, n1 + 8), w(z, n1 + 12));
    }
    if (!(i & 1) && i >= 4 && i <= nsize) {
        const int n1 = 16 * (i - 4) / 2 + 6;
        if (ii == 1) v = w(z, n1);
        if (ii == 2) v = w(z, n1 + 2);
        if (i != nsize) {
            if (ii == 3) v = cadd(w(z, n1 + 8), w(z, n1 + 12));
            if (ii == 4) v = cadd(w(z, n1 + 10), w(z, n1 + 14));
        }
    }
    return v;
}

// calhml's term subtracted from a1(ii, j), ii in {1, 3}; none: zero flag
__host__ __device__ inline bool hml_at(const BandZone &z, int ii, int j, double2 &t)
{
    const double2 c1 = z.c1, c2 = z.c2;
    if (!(j & 1)) {  // j = 2 .. 2 (nl + 1)
        if (j < 2 || j > 2 * (z.nl + 1) || (ii == 1 && j == 2)) return false;
        const int m = (-ii + 3) / 2, n = (ii + j - 3) / 2;
        const int idx = (m + 2) + 4 * (n - 1);
        t = cadd(cneg(sc(c2, z.M(FE_H3MY, idx))), sc(c1, 2.0 / 3.0 * z.M(FE_H4M2N, idx)));
        t = cadd(t, sc(c1, z.M(FE_H6M2L, idx)));
        return true;
    }
    if (j < 3 || j > 2 * z.nl + 1 || (ii == 1 && j == 3)) return false;
    const int m = (-ii + 5) / 2, n = (ii + j - 4) / 2;
    const int idx = (m + 1) + 4 * (n - 1);
    t = csub(sc(c1, z.M(FE_H4M1L, idx)), sc(c2, z.M(FE_H5MY, idx)));
    t = cadd(t, sc(c1, 2.0 / 3.0 * z.M(FE_H6M1N, idx)));
    return true;
}

// calb0 / calb2's element entry k of a fluid zone (entries k % 4 == 3 zero)
__host__ __device__ inline double2 cwb0(const BandZone &z, int k)
{
    if (k % 4 == 3) return cz();
    const double p1 = z.f[(size_t)0 * 4 * z.nl + (k - 1)], p3 = z.f[(size_t)2 * 4 * z.nl + (k - 1)];
    const double2 a = cdiv(make_double2(-p1, 0.0), z.om2);
    const double2 b = make_double2(z.cf.x * p3 - z.cf.y * 0.0, z.cf.x * 0.0 + z.cf.y * p3);  // cmul(cf, (p3, 0))
    return cadd(a, b);
}
__host__ __device__ inline double2 cwb2(const BandZone &z, int k)
{
    if (k % 4 == 3) return cz();
    const double p2 = z.f[(size_t)1 * 4 * z.nl + (k - 1)];
    return cdiv(make_double2(-p2, 0.0), z.om2);
}

// overlapb's b2(ii, i), ii in {3, 4}
template <class W>
__host__ __device__ inline double2 overlapb_at(const BandZone &z, int ii, int i, W w)
{
    const int nl = z.nl;
    double2 v = cz();
    if (ii == 4) {
        if (i == 1) v = w(z, 1);
        if (i >= 2 && i <= nl) v = cadd(w(z, 4 * (i - 1)), w(z, 4 * (i - 1) + 1));
        if (i == nl + 1) v = w(z, 4 * nl);
    } else if (ii == 3) {
        if (i >= 2 && i <= nl + 1) v = w(z, 4 * (i - 2) + 2);
    }
    return v;
}

}  // namespace band_detail

// Column j (0-based) of frequency t of the batch.
__host__ __device__ inline void psv_band_column(const PsvBandArgs &a, int t, int j)
{
    using namespace band_detail;
    const double2 *sc_t = a.scal + (size_t)t * (1 + 3 * a.nz);
    double2 e[4][4];  // [ii][P0, P1, Q0, Q1]
    bool seen[4] = {false, false, false, false};
#pragma unroll
    for (int ii = 0; ii < 4; ii++)
#pragma unroll
        for (int q = 0; q < 4; q++) e[ii][q] = cz();
    auto put = [&](int ii, double2 v0, double2 v1) {  // ii 1-based
        const int r = ii - 1;
        if (!seen[r]) {
            e[r][0] = v0;
            e[r][1] = v1;
            seen[r] = true;
        } else {
            e[r][2] = v0;
            e[r][3] = v1;
        }
    };
    for (int u = 0; u < 2; u++) {
        const int iz = a.colz[2 * j + u];
        if (iz < 0) continue;
        BandZone z;
        z.f = a.fe + a.zbase[iz];
        z.nl = a.znl[iz];
        z.om2 = sc_t[0];
        z.c1 = sc_t[1 + 3 * iz];
        z.c2 = sc_t[2 + 3 * iz];
        z.cf = sc_t[3 + 3 * iz];
        const int i = (j + 1) - a.zkkdr[iz] + 1;  // 1-based column in the zone
        if (a.zsolid[iz]) {
            auto w0 = [](const BandZone &zz, int k) -> double2 { return cw0(zz, k); };
            auto w1 = [](const BandZone &zz, int k) -> double2 { return cw1(zz, k); };
            auto w2 = [](const BandZone &zz, int k) -> double2 { return cw2(zz, k); };
            // cala's order per column: ii 2, 4 (a0, a2), then ii 1, 3 (0, a1)
            for (int ii = 2; ii <= 4; ii += 2) put(ii, overlapa_at(z, ii, i, w0), overlapa_at(z, ii, i, w2));
            for (int ii = 1; ii <= 3; ii += 2) {
                double2 v = overlapa_at(z, ii, i, w1), tm;
                if (hml_at(z, ii, i, tm)) v = csub(v, tm);
                put(ii, cz(), v);
            }
        } else {
            auto b0 = [](const BandZone &zz, int k) -> double2 { return cwb0(zz, k); };
            auto b2 = [](const BandZone &zz, int k) -> double2 { return cwb2(zz, k); };
            for (int ii = 3; ii <= 4; ii++) put(ii, overlapb_at(z, ii, i, b0), overlapb_at(z, ii, i, b2));
        }
    }
    const int oi = a.ovr_ii[j];
    if (oi >= 0) {  // calbc
        e[oi][0] = a.ovr[j];
        e[oi][1] = e[oi][2] = e[oi][3] = cz();
    }
    double2 *out = a.band + ((size_t)a.slot0 + t) * a.nn * 16 + (size_t)j * 16;
#pragma unroll
    for (int ii = 0; ii < 4; ii++)
#pragma unroll
        for (int q = 0; q < 4; q++) out[4 * ii + q] = e[ii][q];
}

}  // namespace dsm
