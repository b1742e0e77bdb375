// hostcheck.cpp -- TEST BUILD ONLY (libdsmsgt_hostcheck.so, never linked
// into libdsmsgt.so): runs the product's SH and P-SV formulations on the
// CPU -- the same set-up code, the same per-lane bodies of the HIP kernels
// (sh_lane.h, psv_lane.h, legendre_lane.h) and the same contraction tables
// (sh_jobs.h, psv_jobs.h),
// with the MFMA contraction replaced by plain loops -- so that the
// reformulation of the reference's per-theta loops can be checked against
// the oracle without a GPU.
#include <algorithm>
#include <cstring>
#include <vector>

#include "dsm_sgt.h"
#include "fe_setup.h"
#include "legendre_lane.h"
#include "psv_jobs.h"
#include "psv_band_lane.h"
#include "psv_lane.h"
#include "sh_jobs.h"
#include "sh_lane.h"

using namespace dsm;

extern "C" int dsm_hostcheck_sh(const dsm_sgt_input_t *in, int32_t i, double *tsgt, double *rsgt, double *synn,
                                int32_t *llog_out)
{
    ShSetup s;
    std::string err;
    int rc = sh_setup(*in, s, err);
    if (rc) return rc;
    ShOmega o;
    sh_omega_setup(s, i, o);
    const int nn = s.nn, R = s.r_n, T = s.theta_n, L = s.maxlmax;
    const int lcap = ((L + 1 + 15) / 16) * 16;
    std::vector<int> need(s.iista.begin(), s.iista.end());
    need.push_back(nn);
    std::sort(need.begin(), need.end(), std::greater<int>());
    need.erase(std::unique(need.begin(), need.end()), need.end());
    std::vector<int> rec_slot(3 * R);
    for (int k = 0; k < 3 * R; k++)
        rec_slot[k] = (int)(std::find(need.begin(), need.end(), s.iista[k]) - need.begin());
    std::vector<double2> band(4 * (size_t)nn);
    for (int j = 0; j < nn; j++) {
        band[0 * nn + j] = make_double2(o.a0[2 * j].real(), o.a0[2 * j].imag());
        band[1 * nn + j] = make_double2(o.a0[2 * j + 1].real(), o.a0[2 * j + 1].imag());
        band[2 * nn + j] = make_double2(o.a2[2 * j].real(), o.a2[2 * j].imag());
        band[3 * nn + j] = make_double2(o.a2[2 * j + 1].real(), o.a2[2 * j + 1].imag());
    }
    const int nneed = (int)need.size();
    const int nlb = (L + 64) / 64, nthr = nlb * 64;
    std::vector<double2> scratch((size_t)(2 * ((nn + SH_SEG - 1) / SH_SEG) + 20) * nthr);
    std::vector<double2> xn((size_t)nneed * SH_NSOL * lcap);
    std::vector<double2> E((size_t)SH_NQ * R * lcap), Es((size_t)SH_NQS * lcap);
    std::vector<double> amp(lcap), rinv(R);
    for (int r = 0; r < R; r++) rinv[r] = 1.0 / s.r_[r];
    int slot0 = 0;
    ShSrcConst src{};
    for (int k = 0; k < 4; k++) {
        src.t4[k] = s.t[s.ins - 1 + k];
        src.h1[k] = s.h1[s.ins - 1 + k];
        src.h2[k] = s.h2[s.ins - 1 + k];
        src.h3[k] = s.h3[s.ins - 1 + k];
        src.h4[k] = s.h4[s.ins - 1 + k];
    }
    for (int k = 0; k < 8; k++) {
        src.gt[k] = s.gt[k];
        src.gh1[k] = s.gh1[k];
        src.gh2[k] = s.gh2[k];
        src.gh3[k] = s.gh3[k];
        src.gh4[k] = s.gh4[k];
    }
    src.r0 = s.r0;
    src.mu0 = s.mu0;
    double2 om[2] = {make_double2(o.comega2.real(), o.comega2.imag()),
                     make_double2(o.coef_spn.real(), o.coef_spn.imag())};
    ShSolveArgs sa{};
    sa.band = band.data();
    sa.slot_list = &slot0;
    sa.nlb = nlb;
    sa.l0 = 0;
    sa.lend = L;
    sa.nn = nn;
    sa.k = s.ns;
    sa.nneed = nneed;
    sa.need = need.data();
    sa.scratch = scratch.data();
    sa.sstride = nthr;
    sa.xn = xn.data();
    sa.lcap = lcap;
    sa.om = om;
    sa.src = &src;
    ShCoefArgs ca{};
    ca.xn = xn.data();
    ca.slot_list = &slot0;
    ca.nlb = nlb;
    ca.l0 = 0;
    ca.lend = L;
    ca.lcap = lcap;
    ca.nneed = nneed;
    ca.r_n = R;
    ca.rec_slot = rec_slot.data();
    ca.w0 = s.w0.data();
    ca.w1 = s.w1.data();
    ca.rinv = rinv.data();
    ca.E = E.data();
    ca.Es = Es.data();
    ca.amp = amp.data();
    for (int b = 0; b < nlb; b++)
        for (int t = 0; t < 64; t++) {
            sh_solve_lane(sa, b, t);
            sh_coef_lane(ca, b, t);
        }
    // l cut-off (calamp replay, SH form)
    int ismall = 0, llog = L, nl = L + 1;
    double maxamp = -1.0;
    for (int l = 0; l <= L; l++) {
        if (ismall > 20) {
            llog = l;
            nl = l;
            break;
        }
        for (int m = -2; m <= 2; m++) {
            if (m == 0 || std::abs(m) > l) continue;
            const double a = std::abs(m) == 2 ? amp[l] : 0.0;
            double ratio = 0.0;
            if (a > maxamp) maxamp = a;
            if (a != 0.0 && maxamp != 0.0) ratio = a / maxamp;
            if (ratio < s.ratl && l > o.lsuf) ismall++;
            else ismall = 0;
        }
    }
    *llog_out = llog - 1;
    // theta functions
    const int tpad = T;
    std::vector<double> th(T), tab((size_t)12 * lcap * tpad);
    for (int k = 0; k < T; k++) th[k] = s.theta_deg[k] / 180.0 * DSM_PI;
    for (int k = 0; k < T; k++) legendre_sh_column(th.data(), T, tpad, L, lcap, tab.data(), k);
    // contraction
    std::memset(tsgt, 0, sizeof(double) * 2 * 10 * R * T);
    std::memset(rsgt, 0, sizeof(double) * 2 * 5 * R * T);
    std::memset(synn, 0, sizeof(double) * 2 * 5 * T);
    const size_t plane = (size_t)lcap * tpad;
    for (int jb = 0; jb < SH_NJOBS; jb++) {
        const ShJobDef &d = SH_JOBS[jb];
        const int Re = d.kind == 2 ? 1 : R, nc = d.kind == 0 ? 10 : 5;
        double *out = d.kind == 0 ? tsgt : (d.kind == 1 ? rsgt : synn);
        for (int r = 0; r < Re; r++)
            for (int t = 0; t < T; t++) {
                double re = 0.0, im = 0.0;
                for (int j = 0; j < d.nb; j++) {
                    const double *Tb = tab.data() + (size_t)((d.mu - 1) * 6 + d.b[j]) * plane + t;
                    const double2 *Eq = d.kind == 2 ? &Es[(size_t)(d.q0 + j) * lcap]
                                                    : &E[((size_t)(d.q0 + j) * R + r) * lcap];
                    for (int l = 0; l < nl; l++) {
                        re += Eq[l].x * Tb[(size_t)l * tpad];
                        im += Eq[l].y * Tb[(size_t)l * tpad];
                    }
                }
                const size_t o2 = d.kind == 2 ? ((size_t)t * nc + d.comp) : (((size_t)t * R + r) * nc + d.comp);
                out[2 * o2] = re;
                out[2 * o2 + 1] = im;
            }
    }
    return 0;
}

// P-SV band pieces of frequency i as the host set-up forms them
// (psv_omega_setup + psv_pack_band): band[nn][4][4] complex.
extern "C" int dsm_hostcheck_psv_band(const dsm_sgt_input_t *in, int32_t i, double *band)
{
    PsvSetup s;
    std::string err;
    int rc = psv_setup(*in, s, err);
    if (rc) return rc;
    PsvOmega o;
    psv_omega_setup(s, i, o);
    psv_pack_band(s, o, band);
    return 0;
}

// The same band through psv_band_column (the device assembly's code, run
// on the host), from psv_band_tables and psv_omega_scalars.
extern "C" int dsm_hostcheck_psv_band_lane(const dsm_sgt_input_t *in, int32_t i, double *band)
{
    PsvSetup s;
    std::string err;
    int rc = psv_setup(*in, s, err);
    if (rc) return rc;
    PsvBandTables bt;
    psv_band_tables(s, bt);
    PsvOmega o;
    psv_omega_scalars(s, i, o);
    const int nz = s.nzone;
    std::vector<double2> scal(1 + 3 * nz);
    auto d2 = [](const cd &c) { return make_double2(c.real(), c.imag()); };
    scal[0] = d2(o.comega2);
    for (int iz = 0; iz < nz; iz++) {
        scal[1 + 3 * iz] = d2(o.zc1[iz]);
        scal[2 + 3 * iz] = d2(o.zc2[iz]);
        scal[3 + 3 * iz] = d2(o.zcf[iz]);
    }
    PsvBandArgs a{};
    a.fe = bt.fe.data();
    a.zbase = bt.zbase.data();
    a.zsolid = bt.zsolid.data();
    a.znl = bt.znl.data();
    a.zkkdr = bt.zkkdr.data();
    a.colz = bt.colz.data();
    a.ovr_ii = bt.ovr_ii.data();
    a.ovr = (const double2 *)bt.ovr.data();
    a.nz = nz;
    a.nn = s.nn;
    a.scal = scal.data();
    a.slot0 = 0;
    a.band = (double2 *)band;
    for (int j = 0; j < s.nn; j++) psv_band_column(a, 0, j);
    return 0;
}

// P-SV: one frequency through psv_l0_lane / psv_solve_lane / psv_coef_lane,
// the calamp replay and loop contractions of the job table.  Output layout
// of the oracle: tsgt[theta][r][20], rsgt[theta][r][10], synn[theta][10].
extern "C" int dsm_hostcheck_psv(const dsm_sgt_input_t *in, int32_t i, double *tsgt, double *rsgt, double *synn,
                                 int32_t *llog_out)
{
    PsvSetup s;
    std::string err;
    int rc = psv_setup(*in, s, err);
    if (rc) return rc;
    PsvOmega o;
    psv_omega_setup(s, i, o);
    const PsvGeom g = psv_geometry(s);
    const int nn = s.nn, R = s.r_n, T = s.theta_n, L = s.maxlmax;
    const int lcap = ((L + 1 + 15) / 16) * 16;
    std::vector<double2> band(16 * (size_t)nn);
    psv_pack_band(s, o, (double *)band.data());
    std::vector<int> kind(nn);
    psv_pack_kind(s, kind.data());
    std::vector<PsvSrc> src(1);
    for (int k = 0; k < 160; k++) {
        src[0].anum[k] = make_double2(o.anum[k].real(), o.anum[k].imag());
        src[0].bnum[k] = make_double2(o.bnum[k].real(), o.bnum[k].imag());
    }
    src[0].coef1 = make_double2(o.coef1_spn.real(), o.coef1_spn.imag());
    src[0].coef2 = make_double2(o.coef2_spn.real(), o.coef2_spn.imag());
    PsvSrcConst sc{s.ecC0, s.ecF0, s.ecL0, s.r0, s.ra[s.mtmp - 1], s.ra[s.mtmp]};
    std::vector<double2> fk(R);
    for (int r = 0; r < R; r++) fk[r] = make_double2(o.fluidK[r].real(), o.fluidK[r].imag());
    const int nneed = (int)g.need.size();
    const int nlb = (L + 64) / 64, nthr = nlb * 64;
    const int n = nn - s.row0 + 1, klo = g.rowlo - s.row0 + 1, fs = s.jtmp - s.row0 + 1;
    std::vector<double2> scratch((size_t)(3 * (n - klo + 1) + 4 * (n - fs + 1)) * nthr);
    std::vector<double2> xn((size_t)nneed * PSV_NSOL * lcap);
    std::vector<double2> E((size_t)PSV_NQ * R * lcap), Es((size_t)PSV_NQS * lcap), ampuv(2 * (size_t)lcap);
    int slot0 = 0;
    PsvSolveArgs sa{};
    sa.band = band.data();
    sa.kind = kind.data();
    sa.slot_list = &slot0;
    sa.nlb = nlb;
    sa.l0 = 0;
    sa.lend = L;
    sa.nn = nn;
    sa.row0 = s.row0;
    sa.jtmp = s.jtmp;
    sa.rowlo = g.rowlo;
    sa.nneed = nneed;
    sa.need = g.need.data();
    sa.scratch = scratch.data();
    sa.sstride = nthr;
    sa.xn = xn.data();
    sa.lcap = lcap;
    sa.src = src.data();
    sa.sc = sc;
    PsvCoefArgs ca{};
    ca.xn = xn.data();
    ca.slot_list = &slot0;
    ca.nlb = nlb;
    ca.l0 = 0;
    ca.lend = L;
    ca.lcap = lcap;
    ca.nneed = nneed;
    ca.r_n = R;
    ca.rec_row = g.rec_row.data();
    ca.fluid = s.fluid_rec.data();
    ca.w0 = s.w0.data();
    ca.w1 = s.w1.data();
    ca.irf = g.irf.data();
    ca.fluidK = fk.data();
    ca.amp_i0 = g.amp_i0;
    ca.amp_i1 = g.amp_i1;
    ca.syn_i0 = g.syn_i0;
    ca.syn_i1 = g.syn_i1;
    ca.E = E.data();
    ca.Es = Es.data();
    ca.ampuv = ampuv.data();
    PsvL0Args za{};
    std::vector<double2> work(5 * (size_t)s.nn0);
    za.band = band.data();
    za.kind = kind.data();
    za.slot_list = &slot0;
    za.nslots = 1;
    za.nn = nn;
    za.nn0 = s.nn0;
    za.row0_0 = s.row0_0;
    za.jtmp = s.jtmp;
    za.js = s.js0;
    za.node_row = s.node_row.data();
    za.node_c1 = s.node_c1.data();
    za.work = work.data();
    za.r_n = R;
    za.jsta0 = s.jsta0.data();
    za.fluid = s.fluid_rec.data();
    za.w0 = s.w0.data();
    za.w1 = s.w1.data();
    za.irf = g.irf.data();
    za.fluidK = fk.data();
    za.src = src.data();
    za.sc = sc;
    za.E = E.data();
    za.Es = Es.data();
    za.lcap = lcap;
    psv_l0_serial(za, 0);
    for (int b = 0; b < nlb; b++)
        for (int t = 0; t < 64; t++) {
            psv_solve_lane(sa, b, t);
            psv_coef_lane(ca, b, t);
        }
    // l cut-off (calamp replay, P-SV form: imt 1 at every m, l >= lsuf)
    int ismall = 0, llog = L, nl = L + 1;
    double maxamp = -1.0;
    for (int l = 0; l <= L; l++) {
        if (ismall > 20) {
            llog = l;
            nl = l;
            break;
        }
        if (l == 0) continue;
        const double a1 = std::hypot(ampuv[2 * l].x, ampuv[2 * l].y), a2 = std::hypot(ampuv[2 * l + 1].x, ampuv[2 * l + 1].y);
        const double a0 = std::sqrt(a1 * a1 + a2 * a2);
        for (int m = -2; m <= 2; m++) {
            if (std::abs(m) > l) continue;
            const double a = (m == 0) ? a0 : 0.0;
            double ratio = 0.0;
            if (a > maxamp) maxamp = a;
            if (a != 0.0 && maxamp != 0.0) ratio = a / maxamp;
            if (ratio < s.ratl && l >= o.lsuf) ismall++;
            else ismall = 0;
        }
    }
    *llog_out = llog - 1;
    std::vector<double> th(T), tab((size_t)PSV_NPLANES * lcap * T);
    for (int k = 0; k < T; k++) th[k] = s.theta_deg[k] / 180.0 * DSM_PI;
    for (int k = 0; k < T; k++) legendre_psv_column(th.data(), T, T, L, lcap, tab.data(), k);
    std::memset(tsgt, 0, sizeof(double) * 2 * 20 * R * T);
    std::memset(rsgt, 0, sizeof(double) * 2 * 10 * R * T);
    std::memset(synn, 0, sizeof(double) * 2 * 10 * T);
    const size_t plane = (size_t)lcap * T;
    for (int pass = 0; pass < 2; pass++)
        for (int jb = 0; jb < PSV_NJOBS; jb++) {
            const PsvJobDef &d = PSV_JOBS[jb];
            if (d.pass != pass) continue;
            const int Re = d.kind == 2 ? 1 : R, nc = d.kind == 0 ? 20 : 10;
            double *out = d.kind == 0 ? tsgt : (d.kind == 1 ? rsgt : synn);
            for (int r = 0; r < Re; r++)
                for (int t = 0; t < T; t++) {
                    double re = 0.0, im = 0.0;
                    for (int j = 0; j < d.nb; j++) {
                        const double *Tb = tab.data() + (size_t)psv_plane(d.mu, d.f[j]) * plane + t;
                        const double2 *Eq = d.kind == 2 ? &Es[(size_t)(d.q0 + j) * lcap]
                                                        : &E[((size_t)(d.q0 + j) * R + r) * lcap];
                        for (int l = 0; l < nl; l++) {
                            re += Eq[l].x * Tb[(size_t)l * T];
                            im += Eq[l].y * Tb[(size_t)l * T];
                        }
                    }
                    const int comps[2] = {d.comp0, d.comp1};
                    const double signs[2] = {(double)d.sign0, (double)d.sign1};
                    for (int tg = 0; tg < 2; tg++) {
                        if (comps[tg] < 0) continue;
                        const size_t o2 = d.kind == 2 ? ((size_t)t * nc + comps[tg]) : (((size_t)t * R + r) * nc + comps[tg]);
                        if (pass == 0) {
                            out[2 * o2] = signs[tg] * re;
                            out[2 * o2 + 1] = signs[tg] * im;
                        } else {
                            out[2 * o2] += signs[tg] * re;
                            out[2 * o2 + 1] += signs[tg] * im;
                        }
                    }
                }
        }
    return 0;
}

// The depth cut of the batched P-SV forms on the host: the first solved band
// row (1-based, full system) of frequency i and degree l, by the rule of
// psv_depth_kernel (psv_kernels.hip) taken row by row -- the decay exponent
// sum min(sqrt((l + 1/2)^2 / r^2 - omega^2 / v^2) dr, dcap) over the
// intervals below the source row, the cut where it passes dcut, at least
// three rows below the source -- with the per-row radius and wave speed of
// psv_row_profile.  Also the system's first row (row0), the source row
// (jtmp), the row count nn and, if urow is not null, urow[j - 1] = 1 for the
// rows holding a radial displacement U (every fluid row; the solid rows at
// even offsets from their zone's first row kkdr, as the receiver stencils
// ksta - 1 = kkdr + 2 (iista - 1) read them).
extern "C" int dsm_hostcheck_psv_depth(const dsm_sgt_input_t *in, int32_t i, int32_t l, double dcut, double dcap,
                                       int32_t *first_row, int32_t *row0, int32_t *jtmp, int32_t *nn,
                                       int32_t *urow)
{
    PsvSetup s;
    std::string err;
    int rc = psv_setup(*in, s, err);
    if (rc) return rc;
    PsvOmega o;
    psv_omega_scalars(s, i, o);
    std::vector<double> r, v;
    psv_row_profile(s, r, v);
    const double al = (double)std::max(1, l) + 0.5, om = o.omega;
    int kst = s.row0;
    double sum = 0.0;
    for (int j = s.jtmp - 1; j >= s.row0; j--) {  // interval between rows j and j + 1 (1-based)
        const double rm = 0.5 * (r[j - 1] + r[j]), vm = 0.5 * (v[j - 1] + v[j]), dr = r[j] - r[j - 1];
        const double q = al / rm, p = om / vm, k2 = q * q - p * p;
        if (k2 > 0.0 && dr > 0.0) sum += std::min(std::sqrt(k2) * dr, dcap);
        if (sum > dcut) {
            kst = j;
            break;
        }
    }
    *first_row = std::min(kst, s.jtmp - 3);
    if (urow)
        for (int iz = 0; iz < s.nzone; iz++) {
            const int j0 = s.kkdr[iz], j1 = (iz + 1 < s.nzone) ? s.kkdr[iz + 1] - 1 : s.nn;
            for (int j = j0; j <= j1; j++) urow[j - 1] = (s.iphase[iz] != 1 || (j - j0) % 2 == 0) ? 1 : 0;
        }
    *row0 = s.row0;
    *jtmp = s.jtmp;
    *nn = s.nn;
    return 0;
}
