// fe_setup_sh.cpp -- omega-independent set-up of the SH engine and the
// per-frequency band pieces.  See fe_setup.h for why this file is built
// with -ffp-contract=off and keeps the reference's operation order.
#include <cmath>
#include <cstring>

#include "fe_setup.h"
#include "fe_poly.h"

namespace dsm {

static const double PI = 3.1415926535897932;

// Reference's 3x3 complex Gaussian elimination with partial pivoting, used
// only to get the interpolation weights (src/SGTsh/glu2.f90:2-143); the
// matrix is real, so it is done in real arithmetic with the same steps.
static void glu3(double a[3][3], double b[3])
{
    const int n = 3;
    const double eps = 3.52e-15;
    int ip[3];
    for (int k = 0; k < n; k++) {
        double amax = std::fabs(a[k][k]);
        int ipk = k;
        for (int i = k + 1; i < n; i++) {
            double aik = std::fabs(a[i][k]);
            if (aik > amax) {
                ipk = i;
                amax = aik;
            }
        }
        ip[k] = ipk;
        if (amax <= eps) return;  // singular: the reference leaves b unchanged
        if (ipk != k) std::swap(a[ipk][k], a[k][k]);
        double wk[3] = {0, 0, 0};
        for (int i = k + 1; i < n; i++) {
            a[i][k] = -a[i][k] / a[k][k];
            wk[i] = a[i][k];
        }
        for (int j = k + 1; j < n; j++) {
            if (ipk != k) std::swap(a[ipk][j], a[k][j]);
            double t = a[k][j];
            for (int i = k + 1; i < n; i++) a[i][j] = a[i][j] + wk[i] * t;
        }
    }
    for (int k = 0; k < n; k++) {
        if (ip[k] != k) std::swap(b[ip[k]], b[k]);
        double t = b[k];
        for (int i = k + 1; i < n; i++) b[i] = b[i] + a[i][k] * t;
    }
    b[n - 1] = b[n - 1] / a[n - 1][n - 1];
    for (int k = n - 2; k >= 0; k--) {
        double t = b[k + 1];
        for (int i = 0; i <= k; i++) b[i] = b[i] - a[i][k + 1] * t;
        b[k] = b[k] / a[k][k];
    }
}

void stencil_weights(double rsta, const double rr[3], int nderiv, double w[3])
{
    double dh[3];
    for (int i = 0; i < 3; i++) dh[i] = rr[i] - rsta;
    if (dh[1] == 0.0 && nderiv == 0) {  // others.f90:864-871 shortcut
        w[0] = 0.0;
        w[1] = 1.0;
        w[2] = 0.0;
        return;
    }
    double a[3][3];
    for (int i = 0; i < 3; i++) {
        a[0][i] = 1.0;
        a[1][i] = dh[i];
        a[2][i] = dh[i] * dh[i] / 2.0;
    }
    double b[3] = {0, 0, 0};
    b[nderiv] = 1.0;
    glu3(a, b);
    for (int i = 0; i < 3; i++) w[i] = b[i];
}

// Evaluate a zone polynomial at r with the reference's running power
// (coef = coef * (r/rmax)), others.f90:676-685.
static inline double poly_at(const double *c4, double r, double rmax)
{
    double v = 0.0, coef = 1.0;
    for (int j = 0; j < 4; j++) {
        if (j > 0) coef = coef * (r / rmax);
        v = v + c4[j] * coef;
    }
    return v;
}

int sh_setup(const dsm_sgt_input_t &in, ShSetup &s, std::string &err)
{
    s = ShSetup();
    if (in.nzone < 1 || in.nzone > DSM_SGT_MAXZONE) {
        err = "nzone out of range";
        return DSM_SGT_EINVAL;
    }
    // Keep the zones above the last fluid zone (SGTsh-MPI.f90:83-137).
    int ioc = 0;
    for (int i = 0; i < in.nzone; i++)
        if (in.vsv[i][0] == 0.0 && in.vsv[i][1] == 0.0 && in.vsv[i][2] == 0.0 && in.vsv[i][3] == 0.0)
            ioc = i + 1;
    s.nzone = in.nzone - ioc;
    if (s.nzone < 1) {
        err = "no solid zone above the fluid core";
        return DSM_SGT_EINVAL;
    }
    const int nz = s.nzone;
    s.vrmin.resize(nz);
    s.vrmax.resize(nz);
    s.rrho.resize(4 * nz);
    s.vsv.resize(4 * nz);
    s.vsh.resize(4 * nz);
    s.qmu.resize(nz);
    for (int i = 0; i < nz; i++) {
        int z = ioc + i;
        s.vrmin[i] = in.vrmin[z];
        s.vrmax[i] = in.vrmax[z];
        for (int j = 0; j < 4; j++) {
            s.rrho[4 * i + j] = in.rho[z][j];
            s.vsv[4 * i + j] = in.vsv[z][j];
            s.vsh[4 * i + j] = in.vsh[z][j];
        }
        s.qmu[i] = in.qmu[z];
    }
    s.rmin = s.vrmin[0];
    s.rmax = s.vrmax[nz - 1];
    s.tlen = in.tlen;
    if (!(in.tlen > 0) || !(in.omegai_seed > 0) || (in.r_list_n <= 0 && !(in.rdelta_ > 0)) ||
        (in.theta_list_n <= 0 && !(in.thetadelta > 0)) || in.maxlmax < 0 || in.imax < 0) {
        err = "invalid tlen / omegai / rdelta / thetadelta / maxlmax / imax";
        return DSM_SGT_EINVAL;
    }
    s.omegai = -std::log(in.omegai_seed) / in.tlen;
    s.ratl = in.ratl;
    s.maxlmax = in.maxlmax;
    s.imax = in.imax;

    // Receivers strictly above the CMB (SGTsh-MPI.f90:152-179); a radius
    // list (ascending) loses its entries at or below it the same way.
    const bool rl = in.r_list_n > 0, tl = in.theta_list_n > 0;
    double rmin_ = in.rmin_;
    int r_n = rl ? in.r_list_n : (int)((in.rmax_ - in.rmin_) / in.rdelta_) + 1;
    int ir_ = 0;
    for (int i = 1; i <= r_n; i++)
        if ((rl ? in.r_list[i - 1] : rmin_ + (double)(i - 1) * in.rdelta_) <= s.rmin) ir_ = i;
    r_n -= ir_;
    rmin_ = rmin_ + (double)ir_ * in.rdelta_;
    if (r_n < 1) {
        err = "no receiver above the fluid core";
        return DSM_SGT_EINVAL;
    }
    s.r_n = r_n;
    s.r_.resize(r_n);
    for (int i = 1; i <= r_n; i++) s.r_[i - 1] = rl ? in.r_list[ir_ + i - 1] : rmin_ + (double)(i - 1) * in.rdelta_;
    for (int i = 0; i < r_n; i++)
        if (s.r_[i] > s.rmax) {
            err = "receiver above the surface";
            return DSM_SGT_EINVAL;
        }
    s.theta_n = theta_grid(in, s.theta_deg);
    if (s.theta_n < 1) {
        err = "empty distance grid";
        return DSM_SGT_EINVAL;
    }
    s.r0 = in.r0min;
    s.intir0 = (int)(s.r0 * 1.0e3);
    if (s.r0 < s.rmin || s.r0 > s.rmax) {
        err = "source radius outside the solid part of the model";
        return DSM_SGT_EINVAL;
    }

    // calgrid (SH: S velocity), others.f90:383-459
    std::vector<double> vmin(nz), gridpar(nz), dzpar(nz);
    {
        const int lmin = 1;
        for (int iz = 0; iz < nz; iz++) {
            double coef1 = 1.0, coef2 = 1.0, vs1 = 0.0, vs2 = 0.0;
            for (int j = 0; j < 4; j++) {
                if (j > 0) {
                    coef1 = coef1 * (s.vrmin[iz] / s.rmax);
                    coef2 = coef2 * (s.vrmax[iz] / s.rmax);
                }
                vs1 = vs1 + s.vsv[4 * iz + j] * coef1;
                vs2 = vs2 + s.vsv[4 * iz + j] * coef2;
            }
            double rh = s.vrmax[iz] - s.vrmin[iz];
            double omega = 2.0 * PI * (double)in.imax / s.tlen;
            vmin[iz] = (vs1 >= vs2) ? vs2 : vs1;
            double amax = s.vrmax[iz];
            double gtmp = (omega * omega) / (vmin[iz] * vmin[iz]) -
                          (((double)lmin + 0.5) * ((double)lmin + 0.5)) / (amax * amax);
            if (gtmp > 0.0) {
                dzpar[iz] = std::sqrt(1.0 / gtmp);
                gridpar[iz] = rh / dzpar[iz];
            } else {
                dzpar[iz] = 0.0;
                gridpar[iz] = 0.0;
            }
        }
        double g = 0.0;
        for (int iz = 0; iz < nz; iz++) g = g + gridpar[iz];
        for (int iz = 0; iz < nz; iz++) {
            if (gridpar[iz] > 0.0) gridpar[iz] = gridpar[iz] / g;
            else gridpar[iz] = (s.vrmax[iz] - s.vrmin[iz]) / (s.rmax - s.rmin) * 0.1;
        }
        g = 0.0;
        for (int iz = 0; iz < nz; iz++) g = g + gridpar[iz];
        for (int iz = 0; iz < nz; iz++) gridpar[iz] = gridpar[iz] / g;
    }
    // calra, others.f90:463-495
    s.nlayer.resize(nz);
    s.nnlayer = 0;
    for (int iz = 0; iz < nz; iz++) {
        double rh = s.vrmax[iz] - s.vrmin[iz];
        int ntmp = (dzpar[iz] == 0.0)
                       ? 1
                       : (int)(std::sqrt(3.3 / in.re) * rh / dzpar[iz] / 2.0 / PI / 7.0e-1 + 1);
        s.nlayer[iz] = ntmp < 5 ? 5 : ntmp;
        s.nnlayer += s.nlayer[iz];
    }
    // calra2, others.f90:500-563
    s.ra.assign(s.nnlayer + nz + 1, 0.0);
    s.ra[0] = s.rmin;
    {
        int itmp = 1;
        for (int iz = 0; iz < nz; iz++) {
            double rh = s.vrmax[iz] - s.vrmin[iz];
            for (int i = 1; i <= s.nlayer[iz]; i++) {
                itmp++;
                s.ra[itmp - 1] = s.vrmin[iz] + rh * (double)i / (double)s.nlayer[iz];
            }
        }
    }
    s.rrsta.assign(3 * r_n, 0.0);
    s.iista.assign(3 * r_n, 0);
    {
        int itmp = 1;
        for (int iz = 0; iz < nz; iz++)
            for (int i = 1; i <= s.nlayer[iz]; i++) {
                for (int ista = 0; ista < r_n; ista++)
                    if (s.ra[itmp - 1] < s.r_[ista] && s.r_[ista] <= s.ra[itmp]) {
                        int base = (i != s.nlayer[iz]) ? itmp : itmp - 1;
                        for (int k = 0; k < 3; k++) {
                            s.rrsta[3 * ista + k] = s.ra[base + k - 1];
                            s.iista[3 * ista + k] = base + k;
                        }
                    }
                itmp++;
            }
    }
    for (int ista = 0; ista < r_n; ista++)
        if (s.iista[3 * ista] == 0) {
            err = "receiver radius not inside the model grid";
            return DSM_SGT_EINVAL;
        }
    s.w0.resize(3 * r_n);
    s.w1.resize(3 * r_n);
    for (int ista = 0; ista < r_n; ista++) {
        stencil_weights(s.r_[ista], &s.rrsta[3 * ista], 0, &s.w0[3 * ista]);
        stencil_weights(s.r_[ista], &s.rrsta[3 * ista], 1, &s.w1[3 * ista]);
    }
    // calsp, others.f90:569-584
    s.isp.resize(nz);
    s.jsp.resize(nz);
    s.isp[0] = 1;
    s.jsp[0] = 1;
    for (int i = 1; i < nz; i++) {
        s.isp[i] = s.isp[i - 1] + s.nlayer[i - 1];
        s.jsp[i] = s.jsp[i - 1] + 4 * s.nlayer[i - 1];
    }
    // calspo, others.f90:588-631
    {
        const int ndc = nz - 1, NL = s.nnlayer;
        double r0 = s.r0, spo;
        if (r0 == s.rmax) {
            spo = (double)NL - 0.01;
            r0 = s.ra[NL - 1] + (spo - (double)(NL - 1)) * (s.ra[NL] - s.ra[NL - 1]);
        } else {
            int itmp;
            for (itmp = 2; itmp <= ndc + NL + 2; itmp++)
                if (r0 < s.ra[itmp - 1]) break;
            spo = (double)(itmp - 2) + (r0 - s.ra[itmp - 2]) / (s.ra[itmp - 1] - s.ra[itmp - 2]);
            if (spo - (double)(itmp - 2) < 0.01) {
                spo = (double)(itmp - 2) + 0.01;
                r0 = s.ra[itmp - 2] + (spo - (double)(itmp - 2)) * (s.ra[itmp - 1] - s.ra[itmp - 2]);
            }
            if (spo - (double)(itmp - 2) > 0.99) {
                spo = (double)(itmp - 2) + 0.99;
                r0 = s.ra[itmp - 2] + (spo - (double)(itmp - 2)) * (s.ra[itmp - 1] - s.ra[itmp - 2]);
            }
        }
        int spn = 0;
        for (int itmp = 1; itmp <= ndc + 1; itmp++) {
            spn++;
            if (r0 <= s.vrmax[itmp - 1]) break;
        }
        spo = spo - (double)(s.isp[spn - 1] - 1);
        s.r0 = r0;
        s.spo = spo;
        s.spn = spn;
    }
    // calgra, others.f90:637-647
    {
        int itmp = s.isp[s.spn - 1] + (int)std::trunc(s.spo);
        s.gra[0] = s.ra[itmp - 1];
        s.gra[1] = s.r0;
        s.gra[2] = s.ra[itmp];
    }
    // calstg, others.f90:651-694: structure nodes (zone boundaries doubled)
    const int NT = s.nnlayer + 2 * nz + 1;
    std::vector<double> vra(NT, 0.0), rho(NT, 0.0), ecL(NT, 0.0), ecN(NT, 0.0);
    int vnp = 0;
    {
        int itmp = 0, jtmp = 0;
        for (int iz = 0; iz < nz; iz++) {
            for (int i = 1; i <= s.nlayer[iz] + 1; i++) {
                itmp++;
                jtmp++;
                vra[itmp - 1] = s.ra[jtmp - 1];
                double trho = poly_at(&s.rrho[4 * iz], vra[itmp - 1], s.rmax);
                double tvsv = poly_at(&s.vsv[4 * iz], vra[itmp - 1], s.rmax);
                double tvsh = poly_at(&s.vsh[4 * iz], vra[itmp - 1], s.rmax);
                rho[itmp - 1] = trho;
                ecL[itmp - 1] = rho[itmp - 1] * tvsv * tvsv;
                ecN[itmp - 1] = rho[itmp - 1] * tvsh * tvsh;
            }
            jtmp--;
        }
        vnp = itmp;
    }
    // calgstg, others.f90:699-742
    double gvra[3], grho[3], gecL[3], gecN[3];
    for (int i = 0; i < 3; i++) {
        gvra[i] = s.gra[i];
        const int z = s.spn - 1;
        double trho = poly_at(&s.rrho[4 * z], gvra[i], s.rmax);
        double tvsv = poly_at(&s.vsv[4 * z], gvra[i], s.rmax);
        double tvsh = poly_at(&s.vsh[4 * z], gvra[i], s.rmax);
        grho[i] = trho;
        gecL[i] = grho[i] * tvsv * tvsv;
        gecN[i] = grho[i] * tvsh * tvsh;
    }
    s.mu0 = gecL[1];

    // element integrals, SGTsh-MPI.f90:314-339
    const int N4 = 4 * s.nnlayer;
    s.t.assign(N4, 0.0);
    s.h1.assign(N4, 0.0);
    s.h2.assign(N4, 0.0);
    s.h3.assign(N4, 0.0);
    s.h4.assign(N4, 0.0);
    std::vector<double> work(N4 + 8, 0.0);
    for (int i = 0; i < nz; i++) {
        const int nl = s.nlayer[i];
        const double *rai = &s.ra[s.isp[i] - 1];
        const int j0 = s.jsp[i] - 1;
        calmatc(nl, vnp, vra.data(), rho.data(), 2, 0, 0, rai, &s.t[j0], &work[j0]);
        calmatc(nl, vnp, vra.data(), ecL.data(), 2, 1, 1, rai, &s.h1[j0], &work[j0]);
        calmatc(nl, vnp, vra.data(), ecL.data(), 1, 1, 0, rai, &s.h2[j0], &work[j0]);
        calmatc(nl, vnp, vra.data(), ecL.data(), 0, 0, 0, rai, &s.h3[j0], &work[j0]);
        calmatc(nl, vnp, vra.data(), ecN.data(), 0, 0, 0, rai, &s.h4[j0], &work[j0]);
        caltl(nl, vnp, vra.data(), rho.data(), rai, &work[j0]);
        calt(nl, &s.t[j0], &work[j0], &s.t[j0]);
        calhl(nl, vnp, vra.data(), ecL.data(), rai, &work[j0]);
        calt(nl, &s.h3[j0], &work[j0], &s.h3[j0]);
        calhl(nl, vnp, vra.data(), ecN.data(), rai, &work[j0]);
        calt(nl, &s.h4[j0], &work[j0], &s.h4[j0]);
    }
    calmatc(2, 3, gvra, grho, 2, 0, 0, s.gra, s.gt, work.data());
    calmatc(2, 3, gvra, gecL, 2, 1, 1, s.gra, s.gh1, work.data());
    calmatc(2, 3, gvra, gecL, 1, 1, 0, s.gra, s.gh2, work.data());
    calmatc(2, 3, gvra, gecL, 0, 0, 0, s.gra, s.gh3, work.data());
    calmatc(2, 3, gvra, gecN, 0, 0, 0, s.gra, s.gh4, work.data());
    caltl(2, 3, gvra, grho, s.gra, work.data());
    calt(2, s.gt, work.data(), s.gt);
    calhl(2, 3, gvra, gecL, s.gra, work.data());
    calt(2, s.gh3, work.data(), s.gh3);
    calhl(2, 3, gvra, gecN, s.gra, work.data());
    calt(2, s.gh4, work.data(), s.gh4);

    s.nn = s.nnlayer + 1;
    s.ns = s.isp[s.spn - 1] + (int)std::trunc(s.spo);
    s.ins = 4 * s.ns - 3;
    return DSM_SGT_OK;
}

void sh_omega_setup(const ShSetup &s, int i, ShOmega &o)
{
    o.omega = 2.0 * PI * (double)i / s.tlen;
    // callsuf, others.f90:748-766: sums the coefficients (surface value).
    {
        double tvs = 0.0;
        for (int k = 0; k < 4; k++) tvs = tvs + s.vsv[4 * (s.nzone - 1) + k] * 1.0;
        o.lsuf = (int)(o.omega * s.vrmax[s.nzone - 1] / tvs - 0.5) + 1;
    }
    // calcoef, others.f90:772-800
    std::vector<cd> coef(s.nzone);
    for (int iz = 0; iz < s.nzone; iz++) {
        if (s.qmu[iz] <= 0.0) coef[iz] = 1.0;
        else {
            double aa = (o.omega == 0.0) ? 1.0 : 1.0 + std::log(o.omega / (2.0 * PI)) / (PI * s.qmu[iz]);
            double bb = 1.0 / (2.0 * s.qmu[iz]);
            cd c(aa, bb);
            coef[iz] = cmul(c, c);
        }
    }
    cd w(o.omega, -s.omegai);
    o.comega2 = cmul(w, w);
    o.coef_spn = coef[s.spn - 1];
    o.a0.assign(2 * s.nn, 0.0);
    o.a2.assign(2 * s.nn, 0.0);
    std::vector<cd> cw(4 * (size_t)s.nnlayer);
    for (int j = 0; j < s.nzone; j++) {
        const int nl = s.nlayer[j], j0 = s.jsp[j] - 1, i0 = s.isp[j] - 1;
        // cala0 calmat.f90:267-285
        for (int k = 0; k < 4 * nl; k++) {
            double h = s.h1[j0 + k] - s.h2[j0 + k] + s.h3[j0 + k] - 2.0 * s.h4[j0 + k];
            cw[j0 + k] = cmul(o.comega2, cd(s.t[j0 + k], 0.0)) - cmul(coef[j], cd(h, 0.0));
        }
        overlap2(nl, &cw[j0], &o.a0[2 * i0]);
        // cala2 calmat.f90:290-304
        for (int k = 0; k < 4 * nl; k++) cw[j0 + k] = -cmul(coef[j], cd(s.h4[j0 + k], 0.0));
        overlap2(nl, &cw[j0], &o.a2[2 * i0]);
    }
}

}  // namespace dsm
