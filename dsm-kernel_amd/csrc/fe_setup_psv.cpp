// fe_setup_psv.cpp -- omega-independent set-up of the P-SV engine
// (SGTpsv-MPI.f90:120-480) and its per-frequency band pieces
// (SGTpsv-MPI.f90:566-625).  Built with -ffp-contract=off in the
// reference's operation order; see fe_setup.h.
#include <algorithm>
#include <atomic>
#include <functional>
#include <thread>
#include <cmath>
#include <cstring>
#include <limits>

#include "fe_setup.h"
#include "fe_poly.h"

namespace dsm {

static const double PI = 3.1415926535897932;

namespace {

// Runs independent host tasks on up to 16 threads and waits for them.
void run_tasks(std::vector<std::function<void()>> &tasks)
{
    std::atomic<size_t> next{0};
    auto worker = [&] {
        for (size_t i; (i = next.fetch_add(1)) < tasks.size();) tasks[i]();
    };
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({tasks.size(), (size_t)16, (size_t)hc});
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
}


// value of the zone polynomial coef[0..3] at r (x = r/rmax powers built
// incrementally, others.f90:590-682)
struct Moduli {
    double rho, vpv, vph, vsv, vsh, eta;
};
Moduli eval_zone(const PsvSetup &s, int iz, double r, double rmax)
{
    Moduli m{0, 0, 0, 0, 0, 0};
    double coef = 1.0;
    for (int j = 0; j < 4; j++) {
        if (j == 0) coef = 1.0;
        else coef = coef * (r / rmax);
        const int o = j + 4 * iz;
        m.rho = m.rho + s.rrho[o] * coef;
        m.vpv = m.vpv + s.vpv[o] * coef;
        m.vph = m.vph + s.vph[o] * coef;
        m.vsv = m.vsv + s.vsv[o] * coef;
        m.vsh = m.vsh + s.vsh[o] * coef;
        m.eta = m.eta + s.eta[o] * coef;
    }
    return m;
}

// calh5 (calmat.f90:246-277): vra/con start at the zone's first structure node
void calh5(int nlayer, const double *vra, const double *con, const double *ra, double *h5)
{
    for (int i = 0; i < 4 * nlayer; i++) h5[i] = 0.0;
    int it = 0;
    for (;;) {
        it++;
        if (vra[it - 1] == ra[0]) {
            if (vra[it] == vra[it - 1]) it++;
            break;
        }
    }
    it--;
    for (int i = 1; i <= nlayer; i++) {
        it++;
        h5[4 * i - 4] = -3.0 / 8.0 * con[it - 1] * ra[i - 1] - 1.0 / 8.0 * con[it] * ra[i];
        h5[4 * i - 3] = -h5[4 * i - 4];
        h5[4 * i - 2] = -1.0 / 8.0 * con[it - 1] * ra[i - 1] - 3.0 / 8.0 * con[it] * ra[i];
        h5[4 * i - 1] = -h5[4 * i - 2];
    }
}

// calhm1 (calmat.f90:282-318): hm1(-1:2, nlayer+1), stored [4*(j-1) + i+1]
void calhm1(int N, const double *vra, const double *con, const double *ra, double *hm1)
{
    auto H = [&](int i, int j) -> double & { return hm1[(i + 1) + 4 * (j - 1)]; };
    int it = 0;
    for (;;) {
        it++;
        if (vra[it - 1] == ra[0]) {
            if (vra[it] == vra[it - 1]) it++;
            break;
        }
    }
    auto C = [&](int k) { return con[k - 1]; };
    auto R = [&](int k) { return ra[k - 1]; };
    H(0, 1) = H(0, 1) - 7.0 / 12.0 * C(it) * R(1);
    H(1, 1) = H(1, 1) + 8.0 / 12.0 * C(it) * R(1);
    H(2, 1) = H(2, 1) - 1.0 / 12.0 * C(it) * R(1);
    for (int i = 2; i <= N - 1; i++) {
        it++;
        H(-1, i) = H(-1, i) - 5.0 / 12.0 * C(it) * R(i);
        H(0, i) = H(0, i) - 3.0 / 12.0 * C(it) * R(i);
        H(1, i) = H(1, i) + 9.0 / 12.0 * C(it) * R(i);
        H(2, i) = H(2, i) - 1.0 / 12.0 * C(it) * R(i);
    }
    it++;
    H(-1, N) = H(-1, N) - 5.0 / 12.0 * C(it) * R(N);
    H(0, N) = H(0, N) - 3.0 / 12.0 * C(it) * R(N);
    H(1, N) = H(1, N) + 8.0 / 12.0 * C(it) * R(N);
    it++;
    H(-1, N + 1) = H(-1, N + 1) - 5.0 / 12.0 * C(it) * R(N + 1);
    H(0, N + 1) = H(0, N + 1) + 5.0 / 12.0 * C(it) * R(N + 1);
}

// calhm2 (calmat.f90:322-358): hm2(-2:1, nlayer+1), stored [4*(j-1) + i+2]
void calhm2(int N, const double *vra, const double *con, const double *ra, double *hm2)
{
    auto H = [&](int i, int j) -> double & { return hm2[(i + 2) + 4 * (j - 1)]; };
    int it = 0;
    for (;;) {
        it++;
        if (vra[it - 1] == ra[0]) {
            if (vra[it] == vra[it - 1]) it++;
            break;
        }
    }
    auto C = [&](int k) { return con[k - 1]; };
    auto R = [&](int k) { return ra[k - 1]; };
    H(0, 1) = H(0, 1) - 5.0 / 12.0 * C(it) * R(1);
    H(1, 1) = H(1, 1) + 5.0 / 12.0 * C(it) * R(1);
    it++;
    H(-1, 2) = H(-1, 2) - 8.0 / 12.0 * C(it) * R(2);
    H(0, 2) = H(0, 2) + 3.0 / 12.0 * C(it) * R(2);
    H(1, 2) = H(1, 2) + 5.0 / 12.0 * C(it) * R(2);
    for (int i = 3; i <= N; i++) {
        it++;
        H(-2, i) = H(-2, i) + 1.0 / 12.0 * C(it) * R(i);
        H(-1, i) = H(-1, i) - 9.0 / 12.0 * C(it) * R(i);
        H(0, i) = H(0, i) + 3.0 / 12.0 * C(it) * R(i);
        H(1, i) = H(1, i) + 5.0 / 12.0 * C(it) * R(i);
    }
    it++;
    H(-2, N + 1) = H(-2, N + 1) + 1.0 / 12.0 * C(it) * R(N + 1);
    H(-1, N + 1) = H(-1, N + 1) - 8.0 / 12.0 * C(it) * R(N + 1);
    H(0, N + 1) = H(0, N + 1) + 7.0 / 12.0 * C(it) * R(N + 1);
}

// mtrnp (calmat.f90:363-386): transpose of the per-layer 2x2 blocks
void mtrnp(int nlayer, const double *a1, double *a2)
{
    for (int i = 0; i < nlayer; i++) {
        a2[4 * i] = a1[4 * i];
        a2[4 * i + 1] = a1[4 * i + 2];
        a2[4 * i + 2] = a1[4 * i + 1];
        a2[4 * i + 3] = a1[4 * i + 3];
    }
}

// mtrnp2 (calmat.f90:390-404): hm1(-p:q, nlayer+1) -> hm2(-q:p, nlayer+1)
void mtrnp2(int nlayer, int p, int q, const double *hm1, double *hm2)
{
    const int w = p + q + 1;
    for (int i = -q; i <= p; i++)
        for (int j = 1; j <= nlayer + 1; j++) {
            const int m = -i, n = i + j;
            if (n >= 1 && n <= nlayer + 1) hm2[(i + q) + w * (j - 1)] = hm1[(m + p) + w * (n - 1)];
        }
}

}  // namespace

int theta_grid(const dsm_sgt_input_t &in, std::vector<double> &deg)
{
    const bool tl = in.theta_list_n > 0;
    const int n = tl ? in.theta_list_n : (int)((in.thetamax - in.thetamin) / in.thetadelta) + 1;
    deg.resize(std::max(0, n));
    for (int i = 1; i <= n; i++) deg[i - 1] = tl ? in.theta_list[i - 1] : in.thetamin + (double)(i - 1) * in.thetadelta;
    return n;
}

int psv_setup(const dsm_sgt_input_t &in, PsvSetup &s, std::string &err)
{
    s = PsvSetup{};
    const int nz = in.nzone;
    if (nz < 1 || nz > DSM_SGT_MAXZONE) {
        err = "nzone out of range";
        return DSM_SGT_EINVAL;
    }
    s.nzone = nz;
    s.ndc = nz - 1;
    s.vrmin.assign(in.vrmin, in.vrmin + nz);
    s.vrmax.assign(in.vrmax, in.vrmax + nz);
    s.rrho.resize(4 * nz);
    s.vpv.resize(4 * nz);
    s.vph.resize(4 * nz);
    s.vsv.resize(4 * nz);
    s.vsh.resize(4 * nz);
    s.eta.resize(4 * nz);
    for (int i = 0; i < nz; i++)
        for (int j = 0; j < 4; j++) {
            s.rrho[j + 4 * i] = in.rho[i][j];
            s.vpv[j + 4 * i] = in.vpv[i][j];
            s.vph[j + 4 * i] = in.vph[i][j];
            s.vsv[j + 4 * i] = in.vsv[i][j];
            s.vsh[j + 4 * i] = in.vsh[i][j];
            s.eta[j + 4 * i] = in.eta[i][j];
        }
    s.qmu.assign(in.qmu, in.qmu + nz);
    s.qkappa.assign(in.qkappa, in.qkappa + nz);
    s.rmin = s.vrmin[0];
    s.rmax = s.vrmax[nz - 1];
    s.tlen = in.tlen;
    s.omegai = -std::log(in.omegai_seed) / in.tlen;
    s.ratl = in.ratl;
    s.maxlmax = in.maxlmax;
    s.imax = in.imax;
    const bool rl = in.r_list_n > 0, tl = in.theta_list_n > 0;
    if (in.tlen <= 0 || (!rl && in.rdelta_ <= 0) || (!tl && in.thetadelta <= 0) || in.maxlmax < 1) {
        err = "tlen, rdelta, thetadelta and maxlmax must be positive";
        return DSM_SGT_EINVAL;
    }
    // receivers (SGTpsv-MPI.f90:127-157): no trim in P-SV; or a radius list
    const int r_n = rl ? in.r_list_n : (int)((in.rmax_ - in.rmin_) / in.rdelta_) + 1;
    if (r_n < 1) {
        err = "no receiver radius";
        return DSM_SGT_EINVAL;
    }
    s.r_n = r_n;
    s.r_.resize(r_n);
    for (int i = 1; i <= r_n; i++) s.r_[i - 1] = rl ? in.r_list[i - 1] : in.rmin_ + (double)(i - 1) * in.rdelta_;
    // calstg_for_stackpoints (others.f90:687-762): lambda = A, qkappa
    s.lambda.assign(r_n, 0.0);
    s.qkp.assign(r_n, 0.0);
    for (int ir = 0; ir < r_n; ir++) {
        const double r = s.r_[ir];
        for (int iz = 0; iz < nz; iz++)
            if (r > s.vrmin[iz] && r <= s.vrmax[iz]) {
                Moduli m = eval_zone(s, iz, r, s.vrmax[nz - 1]);
                s.lambda[ir] = m.rho * m.vph * m.vph;
                s.qkp[ir] = s.qkappa[iz];
            }
    }
    s.theta_n = theta_grid(in, s.theta_deg);
    s.r0 = in.r0min;
    s.intir0 = (int)(s.r0 * 1.0e3);

    // calnl (others.f90:348-369)
    s.iphase.assign(nz, 1);
    int nsl = 0, nll = 0;
    for (int i = 0; i < nz; i++) {
        const double *v = &s.vsv[4 * i];
        if (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0 && v[3] == 0.0) {
            nll++;
            s.iphase[i] = 2;
        } else {
            nsl++;
        }
    }
    if (s.r0 < s.rmin || s.r0 > s.rmax) {
        err = "Location of the source is improper.";
        return DSM_SGT_EINVAL;
    }
    if (s.iphase[nz - 1] != 1) {
        err = "the top zone must be solid";
        return DSM_SGT_EUNSUPPORTED;
    }
    // calgrid, P-SV form (others.f90:374-450): Vp in fluid zones
    std::vector<double> vmin(nz), gridpar(nz), dzpar(nz);
    for (int iz = 0; iz < nz; iz++) {
        const double *v = (s.vsv[4 * iz] == 0.0) ? &s.vpv[4 * iz] : &s.vsv[4 * iz];
        double c1 = 1, c2 = 1, vs1 = 0.0, vs2 = 0.0;
        for (int j = 0; j < 4; j++) {
            if (j == 0) c1 = 1.0;
            else c1 = c1 * (s.vrmin[iz] / s.rmax);
            if (j == 0) c2 = 1.0;
            else c2 = c2 * (s.vrmax[iz] / s.rmax);
            vs1 = vs1 + v[j] * c1;
            vs2 = vs2 + v[j] * c2;
        }
        const double rh = s.vrmax[iz] - s.vrmin[iz];
        const double omega = 2.0 * PI * (double)in.imax / s.tlen;
        vmin[iz] = (vs1 >= vs2) ? vs2 : vs1;
        const double amax = s.vrmax[iz];
        const double gtmp = (omega * omega) / (vmin[iz] * vmin[iz]) - ((1.0 + 0.5) * (1.0 + 0.5)) / (amax * amax);
        if (gtmp > 0.0) {
            dzpar[iz] = std::sqrt(1.0 / gtmp);
            gridpar[iz] = rh / dzpar[iz];
        } else {
            dzpar[iz] = 0.0;
            gridpar[iz] = 0.0;
        }
    }
    // (gridpar is normalised by the reference but not used by the P-SV grid)
    // calra_psv (others.f90:1292-1340)
    s.nlayer.assign(nz, 0);
    int nslay = 0, nllay = 0, inlayer = 0;
    for (int iz = 0; iz < nz; iz++) {
        const double rh = s.vrmax[iz] - s.vrmin[iz];
        int ntmp;
        if (dzpar[iz] == 0.0) ntmp = 1;
        else ntmp = (int)(std::sqrt(3.3 / in.re) * rh / dzpar[iz] / 2.0 / PI / 7.0e-1 + 1);
        s.nlayer[iz] = ntmp < 5 ? 5 : ntmp;
        if (s.iphase[iz] == 1) nslay += s.nlayer[iz];
        else nllay += s.nlayer[iz];
        inlayer += s.nlayer[iz];
    }
    s.nnlayer = inlayer;
    s.nslay = nslay;
    s.nllay = nllay;
    // calra2_psv (others.f90:1344-1419)
    s.ra.assign(inlayer + nz + 1, 0.0);
    s.ra[0] = s.rmin;
    {
        int it = 1;
        for (int iz = 0; iz < nz; iz++) {
            const double rh = s.vrmax[iz] - s.vrmin[iz];
            for (int i = 1; i <= s.nlayer[iz]; i++) {
                it++;
                s.ra[it - 1] = s.vrmin[iz] + rh * (double)i / (double)s.nlayer[iz];
            }
        }
    }
    s.rrsta.assign(3 * r_n, 0.0);
    s.iista.assign(3 * r_n, 0);
    s.istazone.assign(r_n, 0);
    {
        int it = 1;
        for (int iz = 0; iz < nz; iz++)
            for (int i = 1; i <= s.nlayer[iz]; i++) {
                for (int ista = 0; ista < r_n; ista++) {
                    const double rs = s.r_[ista];
                    if (s.ra[it - 1] < rs && rs <= s.ra[it]) {
                        s.istazone[ista] = iz + 1;
                        const int b0 = (i != s.nlayer[iz]) ? 0 : -1;
                        for (int k = 0; k < 3; k++) {
                            s.rrsta[3 * ista + k] = s.ra[it - 1 + b0 + k];
                            s.iista[3 * ista + k] = i + b0 + k;
                        }
                    }
                }
                it++;
            }
    }
    for (int ir = 0; ir < r_n; ir++)
        if (s.istazone[ir] < 1) {
            err = "receiver radius " + std::to_string(s.r_[ir]) + " outside the layered model";
            return DSM_SGT_EINVAL;
        }
    // calsp (others.f90:455-522): first node of every zone
    s.isp.assign(nz, 1);
    for (int i = 1; i < nz; i++) s.isp[i] = s.isp[i - 1] + s.nlayer[i - 1];
    // calspo (others.f90:527-584)
    {
        const double *ra = s.ra.data();
        if (s.r0 == s.rmax) {
            s.spo = (double)inlayer - 0.01;
            s.r0 = ra[inlayer - 1] + (s.spo - (double)(inlayer - 1)) * (ra[inlayer] - ra[inlayer - 1]);
        } else {
            int it = 2;
            while (!(s.r0 < ra[it - 1])) it++;
            s.spo = (double)(it - 2) + (s.r0 - ra[it - 2]) / (ra[it - 1] - ra[it - 2]);
            if (s.spo - (double)(it - 2) < 0.01) {
                s.spo = (double)(it - 2) + 0.01;
                s.r0 = ra[it - 2] + (s.spo - (double)(it - 2)) * (ra[it - 1] - ra[it - 2]);
            }
            if (s.spo - (double)(it - 2) > 0.99) {
                s.spo = (double)(it - 2) + 0.99;
                s.r0 = ra[it - 2] + (s.spo - (double)(it - 2)) * (ra[it - 1] - ra[it - 2]);
            }
        }
        int spn = 0, it = 1;
        for (;;) {
            spn++;
            if (s.r0 <= s.vrmax[it - 1]) break;
            it++;
        }
        if (s.iphase[spn - 1] != 1) {
            err = "the source lies in a fluid zone";
            return DSM_SGT_EUNSUPPORTED;
        }
        s.spn = spn;
        s.spo = s.spo - (double)(s.isp[spn - 1] - 1);
    }

    // calstg (others.f90:590-682): structure nodes, moduli, source moduli
    const int NT = inlayer + 2 * nz + 1;
    std::vector<double> vra(NT, 0.0), rho(NT, 0.0), kappa(NT, 0.0), ecKx(NT, 0.0), ecKy(NT, 0.0), ecKz(NT, 0.0),
        ecL(NT, 0.0), ecN(NT, 0.0);
    int vnp = 0;
    {
        int it = 0, jt = 0;
        for (int iz = 0; iz < nz; iz++) {
            for (int i = 1; i <= s.nlayer[iz] + 1; i++) {
                it++;
                jt++;
                vra[it - 1] = s.ra[jt - 1];
                Moduli m = eval_zone(s, iz, vra[it - 1], s.rmax);
                rho[it - 1] = m.rho;
                ecL[it - 1] = rho[it - 1] * m.vsv * m.vsv;
                ecN[it - 1] = rho[it - 1] * m.vsh * m.vsh;
                const double ecA = m.rho * m.vph * m.vph, ecC = m.rho * m.vpv * m.vpv;
                const double ecF = m.eta * (ecA - 2.0 * ecL[it - 1]);
                kappa[it - 1] = (4.0 * ecA + ecC + 4.0 * ecF - 4.0 * ecN[it - 1]) / 9.0;
                ecKx[it - 1] = ecA - 4.0 / 3.0 * ecN[it - 1];
                ecKy[it - 1] = ecF + 2.0 / 3.0 * ecN[it - 1];
                ecKz[it - 1] = (ecC + 2.0 * ecF) / 3.0;
            }
            jt--;
        }
        vnp = it;
        Moduli m = eval_zone(s, s.spn - 1, s.r0, s.rmax);
        s.ecL0 = m.rho * m.vsv * m.vsv;
        const double ecA0 = m.rho * m.vph * m.vph;
        s.ecC0 = m.rho * m.vpv * m.vpv;
        s.ecF0 = m.eta * (ecA0 - 2.0 * s.ecL0);
    }
    std::vector<double> rhoinv(NT, 0.0), kappainv(NT, 0.0);
    for (int i = 0; i < vnp; i++) {
        rhoinv[i] = 1.0 / rho[i];
        kappainv[i] = 1.0 / kappa[i];
    }
    // element integrals per zone (SGTpsv-MPI.f90:395-435): the matrices of
    // every zone are independent, so they run as host tasks on a few
    // threads (each matrix computed by the same operations as in one
    // sequence), then the zones' derived operators in turn
    s.fe.resize(nz);
    std::vector<std::function<void()>> tasks;
    const double *V = vra.data();
    for (int iz = 0; iz < nz; iz++) {
        PsvZoneFE &z = s.fe[iz];
        const int nl = s.nlayer[iz];
        z.nlayer = nl;
        z.solid = s.iphase[iz] == 1;
        const double *rai = &s.ra[s.isp[iz] - 1];
        auto vec = [&](std::vector<double> &v) { v.assign(4 * nl, 0.0); };
        // a matrix with its lumped part (calhl/caltl, then calt)
        auto lumped = [&, nl, rai](const std::vector<double> *con, std::vector<double> *dst, int rpow, bool hl) {
            tasks.push_back([=, &vnp] {
                std::vector<double> w(4 * nl);
                calmatc_psv(nl, vnp, V, con->data(), rpow, 0, 0, rai, dst->data());
                if (hl) calhl(nl, vnp, V, con->data(), rai, w.data());
                else caltl(nl, vnp, V, con->data(), rai, w.data());
                calt(nl, dst->data(), w.data(), dst->data());
            });
        };
        auto plain = [&, nl, rai](const std::vector<double> *con, std::vector<double> *dst, int rpow, int w1dn, int w2dn) {
            tasks.push_back([=, &vnp] { calmatc_psv(nl, vnp, V, con->data(), rpow, w1dn, w2dn, rai, dst->data()); });
        };
        if (z.solid) {
            for (auto *v : {&z.t, &z.h1x, &z.h1y, &z.h1z, &z.h2L, &z.h2N, &z.h3ax, &z.h3ay, &z.h3az, &z.h4aL, &z.h4aN,
                            &z.h5ax, &z.h5ay, &z.h5az, &z.h6aL, &z.h6aN, &z.h3x, &z.h3y, &z.h3z, &z.h4L, &z.h4N,
                            &z.h5x, &z.h5y, &z.h5z, &z.h6L, &z.h6N, &z.h7x, &z.h7y, &z.h7z, &z.h8L, &z.h8N})
                vec(*v);
            lumped(&rho, &z.t, 2, false);
            lumped(&ecKx, &z.h1x, 0, true);
            lumped(&ecKy, &z.h1y, 0, true);
            lumped(&ecKz, &z.h1z, 0, true);
            lumped(&ecL, &z.h2L, 0, true);
            lumped(&ecN, &z.h2N, 0, true);
            plain(&ecKx, &z.h5ax, 1, 0, 1);
            plain(&ecKy, &z.h5ay, 1, 0, 1);
            plain(&ecKz, &z.h5az, 1, 0, 1);
            plain(&ecL, &z.h6aL, 1, 0, 1);
            plain(&ecN, &z.h6aN, 1, 0, 1);
            plain(&ecKx, &z.h7x, 2, 1, 1);
            plain(&ecKy, &z.h7y, 2, 1, 1);
            plain(&ecKz, &z.h7z, 2, 1, 1);
            plain(&ecL, &z.h8L, 2, 1, 1);
            plain(&ecN, &z.h8N, 2, 1, 1);
        } else {
            vec(z.p1);
            vec(z.p2);
            vec(z.p3);
            plain(&rhoinv, &z.p1, 2, 1, 1);
            lumped(&rhoinv, &z.p2, 0, true);
            lumped(&kappainv, &z.p3, 2, false);
        }
    }
    run_tasks(tasks);
    for (int iz = 0; iz < nz; iz++) {
        PsvZoneFE &z = s.fe[iz];
        if (!z.solid) continue;
        const int nl = z.nlayer;
        const double *rai = &s.ra[s.isp[iz] - 1];
        std::vector<double> work(4 * nl);
        mtrnp(nl, z.h5ax.data(), z.h3ax.data());
        mtrnp(nl, z.h5ay.data(), z.h3ay.data());
        mtrnp(nl, z.h5az.data(), z.h3az.data());
        mtrnp(nl, z.h6aL.data(), z.h4aL.data());
        mtrnp(nl, z.h6aN.data(), z.h4aN.data());
        // modified first-derivative operator (SGTpsv-MPI.f90:438-478);
        // caltstg recomputes the same structure values
        const int jt = s.isp[iz] + iz - 1;  // zone's first structure node, 0-based
        const double *vz = &vra[jt];
        calh5(nl, vz, &ecKx[jt], rai, work.data());
        for (int k = 0; k < 4 * nl; k++) z.h5x[k] = z.h5ax[k] - work[k];
        calh5(nl, vz, &ecKy[jt], rai, work.data());
        for (int k = 0; k < 4 * nl; k++) z.h5y[k] = z.h5ay[k] - work[k];
        calh5(nl, vz, &ecKz[jt], rai, work.data());
        for (int k = 0; k < 4 * nl; k++) z.h5z[k] = z.h5az[k] - work[k];
        calh5(nl, vz, &ecL[jt], rai, work.data());
        for (int k = 0; k < 4 * nl; k++) z.h6L[k] = z.h6aL[k] - work[k];
        calh5(nl, vz, &ecN[jt], rai, work.data());
        for (int k = 0; k < 4 * nl; k++) z.h6N[k] = z.h6aN[k] - work[k];
        mtrnp(nl, z.h5x.data(), z.h3x.data());
        mtrnp(nl, z.h5y.data(), z.h3y.data());
        mtrnp(nl, z.h5z.data(), z.h3z.data());
        mtrnp(nl, z.h6L.data(), z.h4L.data());
        mtrnp(nl, z.h6N.data(), z.h4N.data());
        const int nh = 4 * (nl + 1);
        for (auto *v : {&z.h3mx, &z.h3my, &z.h3mz, &z.h5mx, &z.h5my, &z.h5mz, &z.h4m1L, &z.h4m1N, &z.h4m2L,
                        &z.h4m2N, &z.h6m1L, &z.h6m1N, &z.h6m2L, &z.h6m2N})
            v->assign(nh, 0.0);
        calhm1(nl, vz, &ecKx[jt], rai, z.h5mx.data());
        calhm1(nl, vz, &ecKy[jt], rai, z.h5my.data());
        calhm1(nl, vz, &ecKz[jt], rai, z.h5mz.data());
        calhm1(nl, vz, &ecL[jt], rai, z.h6m1L.data());
        calhm1(nl, vz, &ecN[jt], rai, z.h6m1N.data());
        calhm2(nl, vz, &ecL[jt], rai, z.h6m2L.data());
        calhm2(nl, vz, &ecN[jt], rai, z.h6m2N.data());
        mtrnp2(nl, 1, 2, z.h5mx.data(), z.h3mx.data());
        mtrnp2(nl, 1, 2, z.h5my.data(), z.h3my.data());
        mtrnp2(nl, 1, 2, z.h5mz.data(), z.h3mz.data());
        mtrnp2(nl, 1, 2, z.h6m1L.data(), z.h4m2L.data());
        mtrnp2(nl, 1, 2, z.h6m1N.data(), z.h4m2N.data());
        mtrnp2(nl, 2, 1, z.h6m2L.data(), z.h4m1L.data());
        mtrnp2(nl, 2, 1, z.h6m2N.data(), z.h4m1N.data());
    }
    // calspdr (others.f90:870-898), ksta and nn (SGTpsv-MPI.f90:627-635)
    s.kkdr.assign(nz, 1);
    for (int iz = 0; iz + 1 < nz; iz++) {
        if (s.iphase[iz] == 1) {
            if (s.iphase[iz + 1] == 1) s.kkdr[iz + 1] = s.kkdr[iz] + 2 * s.nlayer[iz];
            else s.kkdr[iz + 1] = s.kkdr[iz] + 2 * (s.nlayer[iz] + 1);
        } else {
            if (s.iphase[iz + 1] == 1) s.kkdr[iz + 1] = s.kkdr[iz] + (s.nlayer[iz] + 1);
            else s.kkdr[iz + 1] = s.kkdr[iz] + s.nlayer[iz];
        }
    }
    s.ksta.resize(r_n);
    for (int ir = 0; ir < r_n; ir++) s.ksta[ir] = s.kkdr[s.istazone[ir] - 1] + 2 * s.iista[3 * ir] - 1;
    s.nn = s.kkdr[nz - 1] + 2 * s.nlayer[nz - 1] + 1;
    s.row0 = (s.rmin == 0.0) ? 3 : 1;
    s.row0_0 = (s.rmin == 0.0) ? 2 : 1;
    // source rows and layer (SGTpsv-MPI.f90:643-648)
    s.jtmp = s.kkdr[s.spn - 1] + 2 * (int)s.spo;
    s.mtmp = s.isp[s.spn - 1] + (int)s.spo;
    if (s.spo == (double)(int)s.spo) {
        s.jtmp -= 2;
        s.mtmp -= 1;
    }
    if (s.jtmp < s.row0 + 2 || s.jtmp + 3 > s.nn) {
        err = "source rows outside the solved band";
        return DSM_SGT_EUNSUPPORTED;
    }
    s.fluid_rec.resize(r_n);
    for (int ir = 0; ir < r_n; ir++) {
        s.fluid_rec[ir] = s.iphase[s.istazone[ir] - 1] == 2;
        if (s.ksta[ir] - 1 < s.row0 || s.ksta[ir] + 4 > s.nn) {
            err = "receiver stencil outside the solved band";
            return DSM_SGT_EUNSUPPORTED;
        }
    }
    // band entry kinds (cala/calbc)
    s.term_kind.assign(4 * (size_t)s.nn, 0);
    s.term_two.assign(4 * (size_t)s.nn, 0);
    auto add_kind = [&](int col, int ii, int kind) {
        const size_t k = 4 * (size_t)(col - 1) + (ii - 1);
        if (s.term_kind[k] == 0) s.term_kind[k] = kind;
        else s.term_two[k] = 1;
    };
    for (int iz = 0; iz < nz; iz++) {
        if (s.iphase[iz] == 1) {
            for (int jj = s.kkdr[iz]; jj <= s.kkdr[iz] + 2 * s.nlayer[iz] + 1; jj++) {
                add_kind(jj, 2, 1);
                add_kind(jj, 4, 1);
                add_kind(jj, 1, 2);
                add_kind(jj, 3, 2);
            }
        } else {
            for (int jj = s.kkdr[iz]; jj <= s.kkdr[iz] + s.nlayer[iz]; jj++) {
                add_kind(jj, 3, 1);
                add_kind(jj, 4, 1);
            }
        }
    }
    for (int i = 0; i + 1 < nz; i++) {
        if (s.iphase[i] == 1 && s.iphase[i + 1] == 2) {
            const size_t k = 4 * (size_t)(s.kkdr[i + 1] - 1) + 1;
            s.term_kind[k] = 3;
            s.term_two[k] = 0;
        }
        if (s.iphase[i] == 2 && s.iphase[i + 1] == 1) {
            const size_t k = 4 * (size_t)(s.kkdr[i + 1] - 1) + 2;
            s.term_kind[k] = 3;
            s.term_two[k] = 0;
        }
    }
    // l = 0: node vector of rea2 (calmat.f90:1247-1315)
    s.node_row.clear();
    s.node_c1.clear();
    s.jsta0.assign(r_n, 0);
    {
        int it = 0;
        for (int iz = 0; iz < nz; iz++) {
            for (int ir = 0; ir < r_n; ir++)
                if (iz + 1 == s.istazone[ir]) s.jsta0[ir] = it + s.iista[3 * ir];
            const int kk = s.kkdr[iz];
            if (s.iphase[iz] == 1) {
                const int i1 = it + 1;
                const int i2 =
                    (iz + 1 != nz) ? i1 + (s.kkdr[iz + 1] - kk) / 2 - 1 : i1 + (s.nn + 1 - kk) / 2 - 1;
                for (int i = i1; i <= i2; i++) {
                    s.node_row.push_back(kk + 2 * (i - i1));
                    int c1;
                    if (i == i1) c1 = (iz == 0) ? 0 : (s.iphase[iz - 1] == 2 ? 3 : 2);
                    else c1 = 2;
                    s.node_c1.push_back(c1);
                }
                it = i2;
            } else {
                const int i1 = it + 1;
                const int i2 = i1 + (s.kkdr[iz + 1] - kk) - 1;
                for (int i = i1; i <= i2; i++) {
                    s.node_row.push_back(kk + (i - i1));
                    int c1;
                    if (i == i1) c1 = (iz == 0) ? 0 : (s.iphase[iz - 1] == 1 ? 2 : 3);
                    else c1 = 3;
                    s.node_c1.push_back(c1);
                }
                it = i2;
            }
        }
        s.nn0 = it;
    }
    {
        const auto it = std::find(s.node_row.begin(), s.node_row.end(), s.jtmp);
        if (it == s.node_row.end() || it + 1 == s.node_row.end() || *(it + 1) != s.jtmp + 2) {
            err = "source rows are not consecutive nodes of the l=0 system";
            return DSM_SGT_EUNSUPPORTED;
        }
        s.js0 = (int)(it - s.node_row.begin());
    }
    for (int ir = 0; ir < r_n; ir++)
        if (s.jsta0[ir] < s.row0_0 || s.jsta0[ir] + 2 > s.nn0) {
            err = "receiver stencil outside the l=0 system";
            return DSM_SGT_EUNSUPPORTED;
        }
    // interpolation weights (interpolate, others.f90:1170-1239)
    s.w0.resize(3 * r_n);
    s.w1.resize(3 * r_n);
    for (int ir = 0; ir < r_n; ir++) {
        stencil_weights(s.r_[ir], &s.rrsta[3 * ir], 0, &s.w0[3 * ir]);
        stencil_weights(s.r_[ir], &s.rrsta[3 * ir], 1, &s.w1[3 * ir]);
    }
    return DSM_SGT_OK;
}

// ---------------------------------------------------------------------------
// per frequency (calmat.f90:409-742, 815-903)
// ---------------------------------------------------------------------------
namespace {

// (re(a) + s*re(b), im(a) + s*im(b)): cala's assembly of one entry
inline cd cpair(cd a, double s, cd b) { return cd(a.real() + s * b.real(), a.imag() + s * b.imag()); }

// overlapa (calmat.f90:670-714): 16*nlayer -> a2(4, 2(nlayer+1))
void overlapa(int nlayer, const cd *a, cd *a2)
{
    const int nsize = 2 * (nlayer + 1), msize = 16 * nlayer;
    auto A = [&](int i, int j) -> cd & { return a2[(i - 1) + 4 * (j - 1)]; };
    auto F = [&](int k) { return a[k - 1]; };
    for (int i = 0; i < 4 * nsize; i++) a2[i] = 0.0;
    A(4, 1) = F(1);
    A(3, 2) = F(2);
    A(4, 2) = F(4);
    A(4, nsize - 1) = F(msize - 3);
    A(3, nsize) = F(msize - 2);
    A(4, nsize) = F(msize);
    for (int i = 3; i <= nsize - 1; i += 2) {
        int n1 = 16 * (i - 3) / 2 + 5;
        A(2, i) = F(n1);
        n1 = n1 + 2;
        A(3, i) = F(n1);
        if (i != nsize - 1) {
            n1 = n1 + 6;
            const int n2 = n1 + 4;
            A(4, i) = F(n1) + F(n2);
        }
    }
    for (int i = 4; i <= nsize; i += 2) {
        int n1 = 16 * (i - 4) / 2 + 6;
        A(1, i) = F(n1);
        n1 = n1 + 2;
        A(2, i) = F(n1);
        if (i != nsize) {
            n1 = n1 + 6;
            int n2 = n1 + 4;
            A(3, i) = F(n1) + F(n2);
            n1 = n1 + 2;
            n2 = n1 + 4;
            A(4, i) = F(n1) + F(n2);
        }
    }
}

// overlapb (calmat.f90:720-742): 4*nlayer -> b2(4, nlayer+1)
void overlapb(int nlayer, const cd *b, cd *b2)
{
    auto B = [&](int i, int j) -> cd & { return b2[(i - 1) + 4 * (j - 1)]; };
    auto F = [&](int k) { return b[k - 1]; };
    for (int i = 0; i < 4 * (nlayer + 1); i++) b2[i] = 0.0;
    B(4, 1) = F(1);
    for (int i = 2; i <= nlayer; i++) {
        const int n1 = 4 * (i - 1), n2 = n1 + 1;
        B(4, i) = F(n1) + F(n2);
    }
    B(4, nlayer + 1) = F(4 * nlayer);
    for (int i = 2; i <= nlayer + 1; i++) B(3, i) = F(4 * (i - 2) + 2);
}

}  // namespace

void psv_omega_scalars(const PsvSetup &s, int i, PsvOmega &o)
{
    const int nz = s.nzone;
    o.omega = 2.0 * PI * (double)i / s.tlen;
    const double omega = o.omega, omegai = s.omegai;
    // callsuf (others.f90:1269-1289): surface Vsv (the coefficient stays 1)
    {
        double tvs = 0.0;
        for (int k = 0; k < 4; k++) tvs = tvs + s.vsv[k + 4 * (nz - 1)] * 1.0;
        o.lsuf = (int)(omega * s.vrmax[nz - 1] / tvs - 0.5) + 1;
    }
    // calcoef (calmat.f90:409-447)
    o.zc1.assign(nz, 0.0);
    o.zc2.assign(nz, 0.0);
    o.zcf.assign(nz, 0.0);
    std::vector<cd> &coef1 = o.zc1, &coef2 = o.zc2, &coef = o.zcf;
    for (int k = 0; k < nz; k++) {
        if (s.qmu[k] <= 0.0) coef1[k] = 1.0;
        else {
            const double aa = (omega == 0.0) ? 1.0 : 1.0 + std::log(omega / (2.0 * PI)) / (PI * s.qmu[k]);
            const double bb = 1.0 / (2.0 * s.qmu[k]);
            coef1[k] = cmul(cd(aa, bb), cd(aa, bb));
        }
        if (s.qkappa[k] <= 0.0) {
            coef2[k] = 1.0;
        } else {
            const double aa = (omega == 0.0) ? 1.0 : 1.0 + std::log(omega / (2.0 * PI)) / (PI * s.qkappa[k]);
            const double bb = 1.0 / (2.0 * s.qkappa[k]);
            coef2[k] = cmul(cd(aa, bb), cd(aa, bb));
        }
        coef[k] = cdiv(cd(1.0, 0.0), coef2[k]);
    }
    const int spn = s.spn;
    o.coef1_spn = coef1[spn - 1];
    o.coef2_spn = coef2[spn - 1];
    const cd w(omega, -omegai);
    o.womega = w;
    o.comega2 = cmul(w, w);
    const cd comega2 = o.comega2;
    // calabnum (calmat.f90:815-903)
    o.anum.assign(160, 0.0);
    o.bnum.assign(160, 0.0);
    {
        const double *ra = &s.ra[s.mtmp - 1];
        const double r0 = s.r0;
        auto AN = [&](std::vector<cd> &x, int i, int j, int k) -> cd & { return x[(i - 1) + 4 * (j - 1) + 16 * (k - 1)]; };
        for (int k = 1; k <= 10; k++) {
            double r;
            if (k <= 5) r = ra[0] + (double)(k - 1) / 4.0 * (r0 - ra[0]);
            else r = ra[1] - (double)(10 - k) / 4.0 * (ra[1] - r0);
            Moduli m{0, 0, 0, 0, 0, 0};
            double c = 1.0;
            for (int j = 0; j < 4; j++) {
                if (j == 0) c = 1.0;
                else c = c * (r / s.rmax);
                const int q = j + 4 * (spn - 1);
                m.rho = m.rho + s.rrho[q] * c;
                m.vpv = m.vpv + s.vpv[q] * c;
                m.vph = m.vph + s.vph[q] * c;
                m.vsv = m.vsv + s.vsv[q] * c;
                m.vsh = m.vsh + s.vsh[q] * c;
                m.eta = m.eta + s.eta[q] * c;
            }
            const double tecA = m.rho * m.vph * m.vph, tecC = m.rho * m.vpv * m.vpv, tecL = m.rho * m.vsv * m.vsv;
            const double tecN = m.rho * m.vsh * m.vsh, tmu = m.rho * m.vsv * m.vsv;
            const double tAk = tecA - tmu * 4.0 / 3.0, tCk = tecC - tmu * 4.0 / 3.0;
            const cd c1 = o.coef1_spn, c2 = o.coef2_spn;
            const cd xAk = cd(tAk * c2.real(), tAk * c2.imag()), xCk = cd(tCk * c2.real(), tCk * c2.imag());
            const cd xrho(m.rho, 0.0);
            const cd xmu = cd(tmu * c1.real(), tmu * c1.imag());
            const cd xecL = cd(tecL * c1.real(), tecL * c1.imag()), xecN = cd(tecN * c1.real(), tecN * c1.imag());
            const cd xecA = xAk + cd(xmu.real() * (4.0 / 3.0), xmu.imag() * (4.0 / 3.0));
            const cd xecC = xCk + cd(xmu.real() * (4.0 / 3.0), xmu.imag() * (4.0 / 3.0));
            const cd xLL = cd(2.0 * xecL.real(), 2.0 * xecL.imag());
            const cd xAmL = xecA - xLL;
            const cd xecF = cd(m.eta * xAmL.real(), m.eta * xAmL.imag());
            const cd xFdC = cdiv(xecF, xecC);
            const cd xAmF2dC = xecA - cmul(xecF, xFdC);
            const double r2 = r * r;
            auto rs = [](double a, cd b) { return cd(a * b.real(), a * b.imag()); };
            AN(o.anum, 1, 1, k) = rs(-(2.0 / r), xFdC);
            AN(o.anum, 1, 2, k) = cdiv(cd(1.0, 0.0), xecC);
            AN(o.anum, 2, 1, k) = cmul(-xrho, comega2) + rs(4.0 / r2, xAmF2dC - xecN);
            AN(o.anum, 2, 2, k) = rs(2.0 / r, xFdC - 1.0);
            AN(o.anum, 3, 1, k) = cd(-(1.0 / r), 0.0);
            AN(o.anum, 3, 3, k) = cd(1.0 / r, 0.0);
            AN(o.anum, 3, 4, k) = cdiv(cd(1.0, 0.0), xecL);
            AN(o.anum, 4, 1, k) = rs(-(2.0 / r2), xAmF2dC - xecN);
            AN(o.anum, 4, 2, k) = rs(-(1.0 / r), xFdC);
            AN(o.anum, 4, 3, k) = cmul(-xrho, comega2) - rs(2.0 / r2, xecN);
            AN(o.anum, 4, 4, k) = cd(-(3.0 / r), 0.0);
            AN(o.bnum, 1, 3, k) = rs(1.0 / r, xFdC);
            AN(o.bnum, 2, 3, k) = rs(-(2.0 / r2), xAmF2dC - xecN);
            AN(o.bnum, 2, 4, k) = cd(1.0 / r, 0.0);
            AN(o.bnum, 4, 3, k) = rs(1.0 / r2, xAmF2dC);
        }
    }
    // calupfluid (others.f90:1009-1031): omega/3 / lambda~ per receiver
    o.fluidK.assign(s.r_n, 0.0);
    {
        const double deuxsurpi = 0.63661977236, unsurdeuxpi = 0.15915494309,
                     unsurtrois = 0.33333333333333333333;
        for (int ir = 0; ir < s.r_n; ir++) {
            if (!s.fluid_rec[ir]) continue;
            const double qk = s.qkp[ir];
            cd fl = cmul(cd(s.lambda[ir], 0.0), cd(1.0, 1.0 / qk));
            fl = cmul(fl, cd(1 + deuxsurpi / qk * std::log(unsurdeuxpi * w.real()), 0.0));
            o.fluidK[ir] = cdiv(cd(w.real() * unsurtrois, w.imag() * unsurtrois), fl);
        }
    }
}

void psv_omega_band(const PsvSetup &s, PsvOmega &o)
{
    const int nz = s.nzone;
    const std::vector<cd> &coef1 = o.zc1, &coef2 = o.zc2, &coef = o.zcf;
    const cd comega2 = o.comega2;
    // per-zone band blocks a0/a1/a2 (SGTpsv-MPI.f90:572-604)
    const size_t NN = (size_t)s.nn;
    o.P0.assign(4 * NN, 0.0);
    o.P1.assign(4 * NN, 0.0);
    o.Q0.assign(4 * NN, 0.0);
    o.Q1.assign(4 * NN, 0.0);
    std::vector<int> seen(4 * NN, 0);
    auto put = [&](int col, int ii, cd v0, cd v1) {
        const size_t k = 4 * (size_t)(col - 1) + (ii - 1);
        if (!seen[k]) {
            o.P0[k] = v0;
            o.P1[k] = v1;
            seen[k] = 1;
        } else {
            o.Q0[k] = v0;
            o.Q1[k] = v1;
        }
    };
    for (int iz = 0; iz < nz; iz++) {
        const PsvZoneFE &z = s.fe[iz];
        const int nl = z.nlayer;
        const cd c1 = coef1[iz], c2 = coef2[iz];
        auto sc = [](const cd &c, double x) { return cd(c.real() * x, c.imag() * x); };
        if (z.solid) {
            std::vector<cd> cw(16 * nl), a0z(8 * (nl + 1)), a1z(8 * (nl + 1)), a2z(8 * (nl + 1));
            auto Fz = [](const std::vector<double> &v, int k) { return v[k - 1]; };
            // cala0 (calmat.f90:452-494)
            for (auto &x : cw) x = 0.0;
            for (int ii = 1; ii <= 16 * nl - 3; ii += 4) {
                const int n = (ii + 3) / 4;
                if (n % 4 == 3) continue;
                const double tt = Fz(z.t, n);
                cd hh0 = sc(c2, 4.0 * Fz(z.h1x, n)) + sc(c1, 16.0 / 3.0 * Fz(z.h2N, n));
                hh0 = hh0 + sc(c1, -4.0 * Fz(z.h2N, n));
                hh0 = hh0 + sc(c2, 2.0 * (Fz(z.h3ay, n) + Fz(z.h5ay, n)));
                hh0 = hh0 - sc(c1, 4.0 / 3.0 * (Fz(z.h4aN, n) + Fz(z.h6aN, n)));
                hh0 = hh0 + sc(c2, 3.0 * Fz(z.h7z, n) - 2.0 * Fz(z.h7y, n));
                hh0 = hh0 + sc(c1, 4.0 / 3.0 * Fz(z.h8N, n));
                cw[ii - 1] = sc(comega2, tt) - hh0;
            }
            for (int ii = 4; ii <= 16 * nl; ii += 4) {
                const int n = ii / 4;
                if (n % 4 == 3) continue;
                const double tt = Fz(z.t, n);
                cd hh0 = sc(c1, Fz(z.h2L, n)) + sc(c1, -2.0 * Fz(z.h2N, n));
                hh0 = hh0 - sc(c1, Fz(z.h4aL, n) + Fz(z.h6aL, n));
                hh0 = hh0 + sc(c1, Fz(z.h8L, n));
                cw[ii - 1] = sc(comega2, tt) - hh0;
            }
            overlapa(nl, cw.data(), a0z.data());
            // cala1 (calmat.f90:499-537)
            for (auto &x : cw) x = 0.0;
            for (int ii = 2; ii <= 16 * nl - 2; ii += 4) {
                const int n = (ii + 2) / 4;
                if (n % 4 == 3) continue;
                cd hh1 = -sc(c2, 2.0 * Fz(z.h1x, n)) - sc(c1, 8.0 / 3.0 * Fz(z.h2N, n));
                hh1 = hh1 - sc(c1, Fz(z.h2L, n));
                hh1 = hh1 - sc(c1, -2.0 * Fz(z.h2N, n));
                hh1 = hh1 - sc(c2, Fz(z.h3y, n));
                hh1 = hh1 - sc(c1, -2.0 / 3.0 * Fz(z.h4N, n));
                hh1 = hh1 + sc(c1, Fz(z.h6L, n));
                cw[ii - 1] = -hh1;
            }
            for (int ii = 3; ii <= 16 * nl - 1; ii += 4) {
                const int n = (ii + 1) / 4;
                if (n % 4 != 2) continue;
                cd hh1 = -sc(c2, 2.0 * Fz(z.h1x, n)) - sc(c1, 8.0 / 3.0 * Fz(z.h2N, n));
                hh1 = hh1 - sc(c1, -2.0 * Fz(z.h2N, n));
                hh1 = hh1 - sc(c1, Fz(z.h2L, n));
                hh1 = hh1 + sc(c1, Fz(z.h4L, n));
                hh1 = hh1 - sc(c2, Fz(z.h5y, n));
                hh1 = hh1 + sc(c1, 2.0 / 3.0 * Fz(z.h6N, n));
                cw[ii - 1] = -hh1;
            }
            overlapa(nl, cw.data(), a1z.data());
            // cala2 (calmat.f90:543-567)
            for (auto &x : cw) x = 0.0;
            for (int ii = 1; ii <= 16 * nl - 3; ii += 4) {
                const int n = (ii + 3) / 4;
                if (n % 4 == 3) continue;
                cw[ii - 1] = -sc(c1, Fz(z.h2L, n));
            }
            for (int ii = 4; ii <= 16 * nl; ii += 4) {
                const int n = ii / 4;
                if (n % 4 == 3) continue;
                const cd hh2 = sc(c2, Fz(z.h1x, n)) + sc(c1, 4.0 / 3.0 * Fz(z.h2N, n));
                cw[ii - 1] = -hh2;
            }
            overlapa(nl, cw.data(), a2z.data());
            // calhml (calmat.f90:629-664) adds into a1
            auto HM2 = [&](const std::vector<double> &v, int ii, int jj) { return v[(ii + 2) + 4 * (jj - 1)]; };
            auto HM1 = [&](const std::vector<double> &v, int ii, int jj) { return v[(ii + 1) + 4 * (jj - 1)]; };
            for (int j = 2; j <= 2 * (nl + 1); j += 2)
                for (int ii = 1; ii <= 3; ii += 2) {
                    if (ii == 1 && j == 2) continue;
                    const int m = (-ii + 3) / 2, n = (ii + j - 3) / 2;
                    cd t = -sc(c2, HM2(z.h3my, m, n)) + sc(c1, 2.0 / 3.0 * HM2(z.h4m2N, m, n));
                    t = t + sc(c1, HM2(z.h6m2L, m, n));
                    cd &dst = a1z[(ii - 1) + 4 * (j - 1)];
                    dst = dst - t;
                }
            for (int j = 3; j <= 2 * nl + 1; j += 2)
                for (int ii = 1; ii <= 3; ii += 2) {
                    if (ii == 1 && j == 3) continue;
                    const int m = (-ii + 5) / 2, n = (ii + j - 4) / 2;
                    cd t = sc(c1, HM1(z.h4m1L, m, n)) - sc(c2, HM1(z.h5my, m, n));
                    t = t + sc(c1, 2.0 / 3.0 * HM1(z.h6m1N, m, n));
                    cd &dst = a1z[(ii - 1) + 4 * (j - 1)];
                    dst = dst - t;
                }
            // cala: contributions of this zone
            for (int c = 0; c <= 2 * nl + 1; c++) {
                const int col = s.kkdr[iz] + c;
                for (int ii = 2; ii <= 4; ii += 2) put(col, ii, a0z[(ii - 1) + 4 * c], a2z[(ii - 1) + 4 * c]);
                for (int ii = 1; ii <= 3; ii += 2) put(col, ii, cd(0.0, 0.0), a1z[(ii - 1) + 4 * c]);
            }
        } else {
            std::vector<cd> cw(4 * nl), a0z(4 * (nl + 1)), a2z(4 * (nl + 1));
            const cd cf = coef[iz];
            // calb0 / calb2 (calmat.f90:572-622)
            for (auto &x : cw) x = 0.0;
            for (int k = 1; k <= 4 * nl; k++) {
                if (k % 4 == 3) continue;
                cw[k - 1] = cdiv(cd(-z.p1[k - 1], 0.0), comega2) + cmul(cf, cd(z.p3[k - 1], 0.0));
            }
            overlapb(nl, cw.data(), a0z.data());
            for (auto &x : cw) x = 0.0;
            for (int k = 1; k <= 4 * nl; k++) {
                if (k % 4 == 3) continue;
                cw[k - 1] = cdiv(cd(-z.p2[k - 1], 0.0), comega2);
            }
            overlapb(nl, cw.data(), a2z.data());
            for (int c = 0; c <= nl; c++) {
                const int col = s.kkdr[iz] + c;
                for (int ii = 3; ii <= 4; ii++) put(col, ii, a0z[(ii - 1) + 4 * c], a2z[(ii - 1) + 4 * c]);
            }
        }
    }
    // calbc (calmat.f90:795-809)
    for (int k = 0; k + 1 < nz; k++) {
        const double rr = s.vrmax[k] * s.vrmax[k];
        if (s.iphase[k] == 1 && s.iphase[k + 1] == 2) {
            const size_t q = 4 * (size_t)(s.kkdr[k + 1] - 1) + 1;
            o.P0[q] = cd(rr, 0.0);
            o.P1[q] = o.Q0[q] = o.Q1[q] = 0.0;
        }
        if (s.iphase[k] == 2 && s.iphase[k + 1] == 1) {
            const size_t q = 4 * (size_t)(s.kkdr[k + 1] - 1) + 2;
            o.P0[q] = -cd(rr, 0.0);
            o.P1[q] = o.Q0[q] = o.Q1[q] = 0.0;
        }
    }
}

void psv_omega_setup(const PsvSetup &s, int i, PsvOmega &o)
{
    psv_omega_scalars(s, i, o);
    psv_omega_band(s, o);
}

void psv_band_tables(const PsvSetup &s, PsvBandTables &t)
{
    const int nz = s.nzone, nn = s.nn;
    t = PsvBandTables{};
    for (int iz = 0; iz < nz; iz++) {
        const PsvZoneFE &z = s.fe[iz];
        const int nl = z.nlayer;
        t.zbase.push_back((long long)t.fe.size());
        t.zsolid.push_back(z.solid ? 1 : 0);
        t.znl.push_back(nl);
        t.zkkdr.push_back(s.kkdr[iz]);
        auto app = [&](const std::vector<double> &v, size_t n) {
            for (size_t k = 0; k < n; k++) t.fe.push_back(k < v.size() ? v[k] : 0.0);
        };
        if (z.solid) {
            const std::vector<double> *a4[PSV_FE_NSOLID4] = {&z.t,   &z.h1x,  &z.h2N,  &z.h3ay, &z.h5ay, &z.h4aN, &z.h6aN,
                                                             &z.h7z, &z.h7y,  &z.h8N,  &z.h2L,  &z.h4aL, &z.h6aL, &z.h8L,
                                                             &z.h3y, &z.h4N,  &z.h6L,  &z.h4L,  &z.h5y,  &z.h6N};
            const std::vector<double> *am[PSV_FE_NSOLIDM] = {&z.h3my, &z.h4m2N, &z.h6m2L, &z.h4m1L, &z.h5my, &z.h6m1N};
            for (auto *v : a4) app(*v, 4 * (size_t)nl);
            for (auto *v : am) app(*v, 4 * (size_t)(nl + 1));
        } else {
            app(z.p1, 4 * (size_t)nl);
            app(z.p2, 4 * (size_t)nl);
            app(z.p3, 4 * (size_t)nl);
        }
    }
    t.colz.assign(2 * (size_t)nn, -1);
    for (int iz = 0; iz < nz; iz++) {
        const int ncol = s.fe[iz].solid ? 2 * s.fe[iz].nlayer + 2 : s.fe[iz].nlayer + 1;
        for (int c = 0; c < ncol; c++) {
            const int col = s.kkdr[iz] + c;  // 1-based
            int *cz = &t.colz[2 * (size_t)(col - 1)];
            if (cz[0] < 0) cz[0] = iz;
            else cz[1] = iz;
        }
    }
    t.ovr_ii.assign(nn, -1);
    t.ovr.assign(2 * (size_t)nn, 0.0);
    for (int k = 0; k + 1 < nz; k++) {  // calbc, as in psv_omega_band
        const double rr = s.vrmax[k] * s.vrmax[k];
        const int j = s.kkdr[k + 1] - 1;
        if (s.iphase[k] == 1 && s.iphase[k + 1] == 2) {
            t.ovr_ii[j] = 1;
            t.ovr[2 * (size_t)j] = rr;
            t.ovr[2 * (size_t)j + 1] = 0.0;
        }
        if (s.iphase[k] == 2 && s.iphase[k + 1] == 1) {
            const cd v = -cd(rr, 0.0);
            t.ovr_ii[j] = 2;
            t.ovr[2 * (size_t)j] = v.real();
            t.ovr[2 * (size_t)j + 1] = v.imag();
        }
    }
}

void psv_pack_band(const PsvSetup &s, const PsvOmega &o, double *band)
{
    for (int j = 0; j < s.nn; j++)
        for (int ii = 0; ii < 4; ii++) {
            const size_t k = 4 * (size_t)j + ii;
            const cd v[4] = {o.P0[k], o.P1[k], o.Q0[k], o.Q1[k]};
            for (int t = 0; t < 4; t++) {
                band[2 * (16 * (size_t)j + 4 * ii + t)] = v[t].real();
                band[2 * (16 * (size_t)j + 4 * ii + t) + 1] = v[t].imag();
            }
        }
}

void psv_pack_kind(const PsvSetup &s, int *kind)
{
    for (int j = 0; j < s.nn; j++) {
        int bits = 0;
        for (int ii = 0; ii < 4; ii++) {
            const size_t k = 4 * (size_t)j + ii;
            bits |= ((s.term_kind[k] & 3) | (s.term_two[k] << 2)) << (4 * ii);
        }
        kind[j] = bits;
    }
}

void psv_row_profile(const PsvSetup &s, std::vector<double> &r, std::vector<double> &v)
{
    r.assign(s.nn, 0.0);
    v.assign(s.nn, 0.0);
    const int nz = s.nzone;
    for (int iz = 0; iz < nz; iz++) {
        const int j0 = s.kkdr[iz], j1 = (iz + 1 < nz) ? s.kkdr[iz + 1] - 1 : s.nn;
        const bool solid = s.iphase[iz] == 1;
        for (int j = j0; j <= j1; j++) {
            int i = solid ? (j - j0) / 2 : (j - j0);
            i = std::min(i, s.nlayer[iz]);
            const double rr = s.ra[s.isp[iz] - 1 + i];
            const Moduli m = eval_zone(s, iz, rr, s.rmax);
            const double vs = solid ? std::min(m.vsv, m.vsh) : std::min(m.vpv, m.vph);
            r[j - 1] = rr;
            v[j - 1] = vs > 0.0 ? vs : std::min(m.vpv, m.vph);
        }
    }
}

PsvGeom psv_geometry(const PsvSetup &s)
{
    PsvGeom g;
    std::vector<int> rows;
    for (int r = 0; r < s.r_n; r++)
        for (int k = 0; k < 6; k++) rows.push_back(s.ksta[r] - 1 + k);
    rows.push_back(s.nn - 1);
    rows.push_back(s.nn);
    std::sort(rows.begin(), rows.end(), std::greater<int>());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    g.need = rows;
    auto idx = [&](int row) { return (int)(std::find(rows.begin(), rows.end(), row) - rows.begin()); };
    g.rec_row.resize(6 * (size_t)s.r_n);
    for (int r = 0; r < s.r_n; r++)
        for (int k = 0; k < 6; k++) g.rec_row[6 * r + k] = idx(s.ksta[r] - 1 + k);
    g.amp_i0 = idx(s.ksta[s.r_n - 1] - 1);
    g.amp_i1 = idx(s.ksta[s.r_n - 1]);
    g.syn_i0 = idx(s.nn - 1);
    g.syn_i1 = idx(s.nn);
    g.rowlo = rows.back();
    g.irf.resize(s.r_n);
    for (int r = 0; r < s.r_n; r++) g.irf[r] = 1.0 / (double)(float)s.r_[r];
    return g;
}

}  // namespace dsm
