// engine.cpp -- the SGT engine behind the C-ABI of include/dsm_sgt.h.
//
// Pipeline for a batch of frequencies (one HIP stream):
//   host   : per-frequency band pieces a0/a2 (fe_setup_sh.cpp)
//   loop over l chunks until every frequency of the batch is cut off:
//     sh_solve  : (omega,l) band LDL^T + basis solves        -> stencil values
//     sh_coef   : excitation, interpolation, strain maps     -> contraction rows, |g0(nn)|
//     host      : replay of the reference's l cut-off (calamp) on |g0(nn)|
//   zero_tail  : drop the rows past each frequency's cut-off
//   synth      : FP64 MFMA contraction over l into the FP64 catalog in HBM
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "dsm_sgt.h"
#include "fe_setup.h"
#include "kernels.h"
#include "sh_jobs.h"

using namespace dsm;

namespace {

constexpr double PI = 3.1415926535897932;

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    template <class T> T *as() const { return (T *)p; }
};

struct TimedLaunch {
    hipEvent_t a, b;
    int kind;  // 0 solve, 1 legendre, 2 synth, 3 other
};

}  // namespace

struct dsm_sgt {
    int mode = 0, precision = 0, device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    dsm_sgt_input_t in{};
    ShSetup sh;
    // geometry of the device arrays
    int theta_pad = 0, lcap = 0, nneed = 0, nslot = 0, lchunk = 0;
    int ipos0 = 0, npos = 0;  // catalog covers frequencies [ipos0, ipos0+npos)
    std::vector<int> need, rec_slot;
    std::vector<int> llog;    // per position: the reference's logged llog-1
    std::vector<char> done;   // per position: computed
    // device buffers
    DevBuf d_theta, d_tab, d_need, d_recslot, d_w0, d_w1, d_rinv, d_src;
    DevBuf d_band, d_om, d_slots, d_nl, d_slotout, d_scratch, d_xn, d_E, d_Es, d_amp, d_jobs;
    DevBuf d_cat_t, d_cat_r, d_cat_s;
    std::vector<SynthJob> jobs;
    int njobs_main = 0, njobs_syn = 0;
    // timing
    bool timing = true;
    std::vector<TimedLaunch> pending;
    dsm_sgt_stats_t stats{};
};

static int fail(dsm_sgt_t *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    return code;
}

#define HIPCHK(h, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(h, e_ == hipErrorOutOfMemory ? DSM_SGT_ENOMEM : DSM_SGT_EHIP,            \
                        std::string(#call) + ": " + hipGetErrorString(e_));                      \
    } while (0)

static void ev_begin(dsm_sgt_t *h, TimedLaunch &t, int kind)
{
    t.kind = kind;
    if (!h->timing) return;
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
    (void)hipEventRecord(t.a, h->stream);
}
static void ev_end(dsm_sgt_t *h, TimedLaunch &t)
{
    if (!h->timing) return;
    (void)hipEventRecord(t.b, h->stream);
    h->pending.push_back(t);
}
static void ev_collect(dsm_sgt_t *h)
{
    for (auto &t : h->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        if (t.kind == 0) h->stats.solve_ms += ms;
        else if (t.kind == 1) h->stats.legendre_ms += ms;
        else if (t.kind == 2) h->stats.synth_ms += ms;
        else h->stats.other_ms += ms;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    h->pending.clear();
}

// Synthesis jobs of the SH contraction (table in sh_jobs.h).
static void sh_synth_jobs(dsm_sgt_t *h)
{
    h->jobs.clear();
    void *outs[3] = {h->d_cat_t.p, h->d_cat_r.p, h->d_cat_s.p};
    const int ncomp[3] = {10, 5, 5};
    h->njobs_main = 0;
    for (int k = 0; k < SH_NJOBS; k++) {
        const ShJobDef &d = SH_JOBS[k];
        SynthJob j{};
        j.q0 = d.q0;
        j.nb = d.nb;
        j.mu = d.mu;
        for (int t = 0; t < 4; t++) j.b[t] = d.b[t];
        j.comp = d.comp;
        j.ncomp = ncomp[d.kind];
        j.out = (double2 *)outs[d.kind];
        h->jobs.push_back(j);
        if (d.kind < 2) h->njobs_main++;
    }
    h->njobs_syn = (int)h->jobs.size() - h->njobs_main;
}

static int sh_create(dsm_sgt_t *h)
{
    int rc = sh_setup(h->in, h->sh, h->err);
    if (rc) return rc;
    const ShSetup &s = h->sh;
    const int R = s.r_n;
    h->theta_pad = ((s.theta_n + 127) / 128) * 128;  // synthesis N tile
    h->lcap = ((s.maxlmax + 1 + 15) / 16) * 16;
    // stencil nodes, descending, always including the surface node nn
    std::vector<int> nodes(s.iista.begin(), s.iista.end());
    nodes.push_back(s.nn);
    std::sort(nodes.begin(), nodes.end(), std::greater<int>());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
    h->need = nodes;
    h->nneed = (int)nodes.size();
    h->rec_slot.resize(3 * R);
    for (int i = 0; i < 3 * R; i++)
        h->rec_slot[i] = (int)(std::find(nodes.begin(), nodes.end(), s.iista[i]) - nodes.begin());
    const int k = s.ns;  // first source node (1-based)
    if (k < 1 || k + 1 > s.nn) return fail(h, DSM_SGT_EINVAL, "source node outside the band");

    // geometry of the batches
    h->nslot = 32;
    h->lchunk = 512;
    h->ipos0 = std::max(0, h->in.imin);
    h->npos = std::max(0, h->in.imax - h->ipos0 + 1);
    h->llog.assign(h->npos, 0);
    h->done.assign(h->npos, 0);

    std::vector<double> th(s.theta_n);
    for (int i = 0; i < s.theta_n; i++) th[i] = s.theta_deg[i] / 180.0 * PI;
    std::vector<double> rinv(R);
    for (int i = 0; i < R; i++) rinv[i] = 1.0 / s.r_[i];
    ShSrcConst src{};
    for (int i = 0; i < 4; i++) {
        src.t4[i] = s.t[s.ins - 1 + i];
        src.h1[i] = s.h1[s.ins - 1 + i];
        src.h2[i] = s.h2[s.ins - 1 + i];
        src.h3[i] = s.h3[s.ins - 1 + i];
        src.h4[i] = s.h4[s.ins - 1 + i];
    }
    for (int i = 0; i < 8; i++) {
        src.gt[i] = s.gt[i];
        src.gh1[i] = s.gh1[i];
        src.gh2[i] = s.gh2[i];
        src.gh3[i] = s.gh3[i];
        src.gh4[i] = s.gh4[i];
    }
    src.r0 = s.r0;
    src.mu0 = s.mu0;

    HIPCHK(h, h->d_theta.ensure(sizeof(double) * s.theta_n));
    HIPCHK(h, h->d_tab.ensure(sizeof(double) * 12 * (size_t)h->lcap * h->theta_pad));
    HIPCHK(h, h->d_need.ensure(sizeof(int) * h->nneed));
    HIPCHK(h, h->d_recslot.ensure(sizeof(int) * 3 * R));
    HIPCHK(h, h->d_w0.ensure(sizeof(double) * 3 * R));
    HIPCHK(h, h->d_w1.ensure(sizeof(double) * 3 * R));
    HIPCHK(h, h->d_rinv.ensure(sizeof(double) * R));
    HIPCHK(h, h->d_src.ensure(sizeof(ShSrcConst)));
    HIPCHK(h, hipMemcpy(h->d_theta.p, th.data(), sizeof(double) * s.theta_n, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_need.p, h->need.data(), sizeof(int) * h->nneed, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_recslot.p, h->rec_slot.data(), sizeof(int) * 3 * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_w0.p, s.w0.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_w1.p, s.w1.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_rinv.p, rinv.data(), sizeof(double) * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_src.p, &src, sizeof(src), hipMemcpyHostToDevice));

    const int S = h->nslot, nn = s.nn;
    HIPCHK(h, h->d_band.ensure(sizeof(double2) * 4 * (size_t)nn * S));
    HIPCHK(h, h->d_om.ensure(sizeof(double2) * 2 * S));
    HIPCHK(h, h->d_slots.ensure(sizeof(int) * S));
    HIPCHK(h, h->d_nl.ensure(sizeof(int) * S));
    HIPCHK(h, h->d_slotout.ensure(sizeof(int) * S));
    const size_t nthreads = (size_t)S * h->lchunk;
    HIPCHK(h, h->d_scratch.ensure(sizeof(double2) * (size_t)(nn + 10 * (nn - k + 1)) * nthreads));
    HIPCHK(h, h->d_xn.ensure(sizeof(double2) * (size_t)S * h->nneed * SH_NSOL * h->lcap));
    HIPCHK(h, h->d_E.ensure(sizeof(double2) * (size_t)S * 36 * R * h->lcap));
    HIPCHK(h, h->d_Es.ensure(sizeof(double2) * (size_t)S * 5 * h->lcap));
    HIPCHK(h, h->d_amp.ensure(sizeof(double) * (size_t)S * h->lcap));
    const size_t np = (size_t)std::max(1, h->npos);
    HIPCHK(h, h->d_cat_t.ensure(sizeof(double2) * np * R * s.theta_n * 10));
    HIPCHK(h, h->d_cat_r.ensure(sizeof(double2) * np * R * s.theta_n * 5));
    HIPCHK(h, h->d_cat_s.ensure(sizeof(double2) * np * s.theta_n * 5));
    HIPCHK(h, hipMemsetAsync(h->d_cat_t.p, 0, h->d_cat_t.n, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_cat_r.p, 0, h->d_cat_r.n, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_cat_s.p, 0, h->d_cat_s.n, h->stream));
    sh_synth_jobs(h);
    HIPCHK(h, h->d_jobs.ensure(sizeof(SynthJob) * h->jobs.size()));
    HIPCHK(h, hipMemcpy(h->d_jobs.p, h->jobs.data(), sizeof(SynthJob) * h->jobs.size(), hipMemcpyHostToDevice));

    TimedLaunch t;
    ev_begin(h, t, 1);
    launch_legendre_sh(h->d_theta.as<double>(), s.theta_n, h->theta_pad, s.maxlmax, h->lcap, h->d_tab.as<double>(),
                       h->stream);
    HIPCHK(h, hipGetLastError());
    ev_end(h, t);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ev_collect(h);
    return DSM_SGT_OK;
}

// The reference's l cut-off state (calamp, src/SGTsh/others.f90:805-824).
struct CutState {
    int ismall = 0, lsuf = 0, nl = -1, llog = 0;
    double maxamp = -1.0;
    bool done = false;
};
static inline void calamp(CutState &c, double amp, int l, double ratl, bool sh)
{
    double ampratio = 0.0;
    if (amp > c.maxamp) c.maxamp = amp;
    if (amp != 0.0 && c.maxamp != 0.0) ampratio = amp / c.maxamp;
    const bool past = sh ? (l > c.lsuf) : (l >= c.lsuf);
    if (ampratio < ratl && past) c.ismall++;
    else c.ismall = 0;
}

static int sh_run_batch(dsm_sgt_t *h, const std::vector<int> &freqs)
{
    const ShSetup &s = h->sh;
    const int S = (int)freqs.size(), nn = s.nn, R = s.r_n, LC = h->lcap;
    std::vector<double2> band(4 * (size_t)nn * S), om(2 * S);
    std::vector<CutState> cut(S);
    std::vector<int> slotout(S);
    for (int k = 0; k < S; k++) {
        ShOmega o;
        sh_omega_setup(s, freqs[k], o);
        for (int j = 0; j < nn; j++) {
            band[(size_t)(k * 4 + 0) * nn + j] = make_double2(o.a0[2 * j].real(), o.a0[2 * j].imag());
            band[(size_t)(k * 4 + 1) * nn + j] = make_double2(o.a0[2 * j + 1].real(), o.a0[2 * j + 1].imag());
            band[(size_t)(k * 4 + 2) * nn + j] = make_double2(o.a2[2 * j].real(), o.a2[2 * j].imag());
            band[(size_t)(k * 4 + 3) * nn + j] = make_double2(o.a2[2 * j + 1].real(), o.a2[2 * j + 1].imag());
        }
        om[2 * k] = make_double2(o.comega2.real(), o.comega2.imag());
        om[2 * k + 1] = make_double2(o.coef_spn.real(), o.coef_spn.imag());
        cut[k].lsuf = o.lsuf;
        cut[k].llog = s.maxlmax;
        slotout[k] = freqs[k] - h->ipos0;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_band.p, band.data(), sizeof(double2) * band.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_om.p, om.data(), sizeof(double2) * om.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_slotout.p, slotout.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));

    std::vector<int> active(S);
    for (int k = 0; k < S; k++) active[k] = k;
    std::vector<double> amp((size_t)S * LC);
    int l0 = 0;
    while (!active.empty()) {
        const int lend = std::min(l0 + h->lchunk - 1, s.maxlmax);
        const int nlb = (lend - l0 + 64) / 64;
        const int nblocks = (int)active.size() * nlb;
        HIPCHK(h, hipMemcpyAsync(h->d_slots.p, active.data(), sizeof(int) * active.size(), hipMemcpyHostToDevice,
                                 h->stream));
        ShSolveArgs sa{};
        sa.band = h->d_band.as<double2>();
        sa.slot_list = h->d_slots.as<int>();
        sa.nlb = nlb;
        sa.l0 = l0;
        sa.lend = lend;
        sa.nn = nn;
        sa.k = s.ns;
        sa.nneed = h->nneed;
        sa.need = h->d_need.as<int>();
        sa.scratch = h->d_scratch.as<double2>();
        sa.sstride = (long)nblocks * 64;
        sa.xn = h->d_xn.as<double2>();
        sa.lcap = LC;
        sa.om = h->d_om.as<double2>();
        sa.src = h->d_src.as<ShSrcConst>();
        TimedLaunch t0;
        ev_begin(h, t0, 0);
        launch_sh_solve(sa, nblocks, h->stream);
        HIPCHK(h, hipGetLastError());
        ShCoefArgs ca{};
        ca.xn = h->d_xn.as<double2>();
        ca.slot_list = h->d_slots.as<int>();
        ca.nlb = nlb;
        ca.l0 = l0;
        ca.lend = lend;
        ca.lcap = LC;
        ca.nneed = h->nneed;
        ca.r_n = R;
        ca.rec_slot = h->d_recslot.as<int>();
        ca.w0 = h->d_w0.as<double>();
        ca.w1 = h->d_w1.as<double>();
        ca.rinv = h->d_rinv.as<double>();
        ca.E = h->d_E.as<double2>();
        ca.Es = h->d_Es.as<double2>();
        ca.amp = h->d_amp.as<double>();
        launch_sh_coef(ca, nblocks, h->stream);
        HIPCHK(h, hipGetLastError());
        ev_end(h, t0);
        h->stats.solve_launches++;
        h->stats.systems += (long)active.size() * (lend - l0 + 1);
        // algorithmic bytes: band read once, 14 nonzero right-hand sides'
        // (value, derivative) at every receiver written once (w = 2)
        h->stats.solve_bytes += (double)active.size() * (lend - l0 + 1) * 16.0 * (2.0 * nn + 4.0 * 14 * R);
        // |g0(nn)| of the chunk for every slot: one strided copy
        HIPCHK(h, hipMemcpy2DAsync(&amp[l0], sizeof(double) * LC, h->d_amp.as<double>() + l0, sizeof(double) * LC,
                                   sizeof(double) * (lend - l0 + 1), S, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        // replay the cut-off in the reference's order (SGTsh-MPI.f90:423-478)
        std::vector<int> still;
        for (int k : active) {
            CutState &c = cut[k];
            for (int l = l0; l <= lend && !c.done; l++) {
                if (c.ismall > 20) {
                    if (c.llog > l) c.llog = l;
                    c.nl = l;
                    c.done = true;
                    break;
                }
                for (int m = -2; m <= 2; m++) {
                    if (m == 0 || std::abs(m) > l) continue;
                    const double a = (std::abs(m) == 2) ? amp[(size_t)k * LC + l] : 0.0;
                    calamp(c, a, l, s.ratl, true);
                }
            }
            if (!c.done && lend >= s.maxlmax) {
                c.nl = s.maxlmax + 1;
                c.done = true;
            }
            if (!c.done) still.push_back(k);
        }
        active.swap(still);
        l0 = lend + 1;
    }
    std::vector<int> nl(S), all(S);
    for (int k = 0; k < S; k++) {
        nl[k] = cut[k].nl;
        all[k] = k;
        h->llog[slotout[k]] = cut[k].llog - 1;
        h->done[slotout[k]] = 1;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_nl.p, nl.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_slots.p, all.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    TimedLaunch tz;
    ev_begin(h, tz, 3);
    launch_zero_tail(h->d_E.as<double2>(), S, 36 * R, h->d_slots.as<int>(), h->d_nl.as<int>(), LC, h->stream);
    launch_zero_tail(h->d_Es.as<double2>(), S, 5, h->d_slots.as<int>(), h->d_nl.as<int>(), LC, h->stream);
    HIPCHK(h, hipGetLastError());
    ev_end(h, tz);

    SynthArgs ga{};
    ga.E = h->d_E.as<double2>();
    ga.nq = 36;
    ga.r_eff = R;
    ga.tab = h->d_tab.as<double>();
    ga.theta_n = s.theta_n;
    ga.theta_pad = h->theta_pad;
    ga.lcap = LC;
    ga.nslot = S;
    ga.nl = h->d_nl.as<int>();
    ga.slot_out = h->d_slotout.as<int>();
    ga.jobs = h->d_jobs.as<SynthJob>();
    ga.njobs = h->njobs_main;
    TimedLaunch tg;
    ev_begin(h, tg, 2);
    launch_synth_f64(ga, h->stream);
    HIPCHK(h, hipGetLastError());
    ev_end(h, tg);
    // SYNN rows (surface node, no receiver dimension)
    ga.E = h->d_Es.as<double2>();
    ga.nq = SH_NQS;
    ga.r_eff = 1;
    ga.jobs = h->d_jobs.as<SynthJob>() + h->njobs_main;
    ga.njobs = h->njobs_syn;
    TimedLaunch tg2;
    ev_begin(h, tg2, 3);
    launch_synth_f64(ga, h->stream);
    HIPCHK(h, hipGetLastError());
    ev_end(h, tg2);
    h->stats.synth_launches += 1;
    double sumL = 0.0;
    for (int k = 0; k < S; k++) sumL += nl[k];
    // 4 flops per (complex coefficient x real basis) per (r,theta); 36 rows
    // per receiver
    h->stats.synth_flops += 4.0 * sumL * s.theta_n * 36.0 * R;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ev_collect(h);
    return DSM_SGT_OK;
}

// ------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------
extern "C" {

int dsm_sgt_create(const dsm_sgt_input_t *in, int32_t mode, int32_t precision, int32_t device,
                   dsm_sgt_t **handle)
{
    if (!in || !handle) return DSM_SGT_EINVAL;
    *handle = nullptr;
    dsm_sgt_t *h = new (std::nothrow) dsm_sgt_t();
    if (!h) return DSM_SGT_ENOMEM;
    h->in = *in;
    h->mode = mode;
    h->precision = precision;
    h->device = device;
    int rc = DSM_SGT_OK;
    do {
        if (mode != DSM_SGT_SH) {
            rc = fail(h, DSM_SGT_EUNSUPPORTED, "mode not available in this build");
            break;
        }
        if (precision != DSM_SGT_FP64) {
            rc = fail(h, DSM_SGT_EUNSUPPORTED, "precision not available in this build");
            break;
        }
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
            rc = fail(h, DSM_SGT_EHIP, "no HIP device " + std::to_string(device));
            break;
        }
        if (hipSetDevice(device) != hipSuccess) {
            rc = fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
            break;
        }
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(h, DSM_SGT_EHIP, "hipStreamCreate failed");
            break;
        }
        rc = sh_create(h);
    } while (0);
    *handle = h;  // returned even on failure so dsm_sgt_last_error() can report
    return rc;
}

int dsm_sgt_get_info(dsm_sgt_t *h, dsm_sgt_info_t *info)
{
    if (!h || !info) return DSM_SGT_EINVAL;
    if (h->sh.nn == 0) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    info->mode = h->mode;
    info->r_n = h->sh.r_n;
    info->theta_n = h->sh.theta_n;
    info->nn = h->sh.nn;
    info->num_tsgt = 10;
    info->num_rsgt = 5;
    info->num_synn = 5;
    info->intir0 = h->sh.intir0;
    info->r0 = h->sh.r0;
    return DSM_SGT_OK;
}

int dsm_sgt_get_grid(dsm_sgt_t *h, double *radii, double *theta_deg)
{
    if (!h) return DSM_SGT_EINVAL;
    if (h->sh.nn == 0) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    if (radii) std::memcpy(radii, h->sh.r_.data(), sizeof(double) * h->sh.r_n);
    if (theta_deg) std::memcpy(theta_deg, h->sh.theta_deg.data(), sizeof(double) * h->sh.theta_n);
    return DSM_SGT_OK;
}

int dsm_sgt_run(dsm_sgt_t *h, int32_t i_begin, int32_t i_end)
{
    if (!h) return DSM_SGT_EINVAL;
    if (h->sh.nn == 0) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    if (i_begin < h->ipos0 || i_end >= h->ipos0 + h->npos || i_begin > i_end)
        return fail(h, DSM_SGT_EINVAL, "frequency range outside [imin, imax] of the input");
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    // zig-zag order of the reference's rank distribution is irrelevant on one
    // device; batch consecutive frequencies (similar l cut-off)
    std::vector<int> all;
    for (int i = i_begin; i <= i_end; i++)
        if (i != 0) all.push_back(i);
    for (size_t b = 0; b < all.size(); b += h->nslot) {
        std::vector<int> batch(all.begin() + b, all.begin() + std::min(all.size(), b + (size_t)h->nslot));
        int rc = sh_run_batch(h, batch);
        if (rc) return rc;
    }
    return DSM_SGT_OK;
}

static int fetch_common(dsm_sgt_t *h, int32_t i, int32_t kind, int32_t ir, std::vector<double2> &buf, int &num)
{
    if (!h) return DSM_SGT_EINVAL;
    const int pos = i - h->ipos0;
    if (pos < 0 || pos >= h->npos || !h->done[pos])
        return fail(h, DSM_SGT_ESTATE, "frequency " + std::to_string(i) + " not computed");
    const int R = h->sh.r_n, T = h->sh.theta_n;
    const double2 *src;
    if (kind == DSM_SGT_TSGT) {
        num = 10;
        if (ir < 0 || ir >= R) return fail(h, DSM_SGT_EINVAL, "receiver index out of range");
        src = h->d_cat_t.as<double2>() + ((size_t)pos * R + ir) * T * num;
    } else if (kind == DSM_SGT_RSGT) {
        num = 5;
        if (ir < 0 || ir >= R) return fail(h, DSM_SGT_EINVAL, "receiver index out of range");
        src = h->d_cat_r.as<double2>() + ((size_t)pos * R + ir) * T * num;
    } else if (kind == DSM_SGT_SYNN) {
        num = 5;
        src = h->d_cat_s.as<double2>() + (size_t)pos * T * num;
    } else {
        return fail(h, DSM_SGT_EINVAL, "unknown catalog kind");
    }
    buf.resize((size_t)num * T);
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    HIPCHK(h, hipMemcpy(buf.data(), src, sizeof(double2) * buf.size(), hipMemcpyDeviceToHost));
    return DSM_SGT_OK;
}

int dsm_sgt_fetch_fp64(dsm_sgt_t *h, int32_t i, int32_t kind, int32_t ir, void *out)
{
    std::vector<double2> buf;
    int num = 0;
    int rc = fetch_common(h, i, kind, ir, buf, num);
    if (rc) return rc;
    std::memcpy(out, buf.data(), sizeof(double2) * buf.size());
    return DSM_SGT_OK;
}

int dsm_sgt_fetch(dsm_sgt_t *h, int32_t i, int32_t kind, int32_t ir, void *out)
{
    std::vector<double2> buf;
    int num = 0;
    int rc = fetch_common(h, i, kind, ir, buf, num);
    if (rc) return rc;
    float *o = (float *)out;
    for (size_t k = 0; k < buf.size(); k++) {
        o[2 * k] = (float)buf[k].x;
        o[2 * k + 1] = (float)buf[k].y;
    }
    return DSM_SGT_OK;
}

int dsm_sgt_lcut(dsm_sgt_t *h, int32_t i, int32_t *llog_minus_1)
{
    if (!h || !llog_minus_1) return DSM_SGT_EINVAL;
    const int pos = i - h->ipos0;
    if (pos < 0 || pos >= h->npos || !h->done[pos])
        return fail(h, DSM_SGT_ESTATE, "frequency " + std::to_string(i) + " not computed");
    *llog_minus_1 = h->llog[pos];
    return DSM_SGT_OK;
}

// File names of SGTsh-MPI.f90:570-606 / SGTpsv-MPI.f90:1106-1144:
// Fortran I7 fields joined by '.', blanks replaced by '0'.
static std::string i7(long v)
{
    char b[32];
    std::snprintf(b, sizeof b, "%7ld", v);
    std::string s(b);
    if (s.size() > 7) s = std::string(7, '*');
    for (char &c : s)
        if (c == ' ') c = '0';
    return s;
}

static int write_record(dsm_sgt_t *h, const std::string &path, int32_t i, int32_t kind, int32_t ir)
{
    const int T = h->sh.theta_n;
    const int num = kind == DSM_SGT_TSGT ? 10 : 5;
    std::vector<float> rec((size_t)2 * num * T);
    int rc = dsm_sgt_fetch(h, i, kind, ir, rec.data());
    if (rc) return rc;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(h, DSM_SGT_EIO, "cannot open " + path);
    size_t w = std::fwrite(rec.data(), sizeof(float), rec.size(), f);
    int c = std::fclose(f);
    if (w != rec.size() || c != 0) return fail(h, DSM_SGT_EIO, "short write " + path);
    return DSM_SGT_OK;
}

int dsm_sgt_write_catalog(dsm_sgt_t *h, int32_t i, const char *output_dir, const char *model_name)
{
    if (!h || !output_dir || !model_name) return DSM_SGT_EINVAL;
    const std::string od(output_dir), mn(model_name);
    const char *suf = h->mode == DSM_SGT_SH ? "SH" : "PSV";
    ::mkdir(od.c_str(), 0755);
    ::mkdir((od + "/TSGT").c_str(), 0755);
    ::mkdir((od + "/RSGT").c_str(), 0755);
    ::mkdir((od + "/log").c_str(), 0755);
    const int R = h->sh.r_n;
    for (int ir = 0; ir < R; ir++) {
        const long rr = (long)(h->sh.r_[ir] * 1.0e3);
        std::string name = od + "/TSGT/" + mn + "." + i7(h->sh.intir0) + "." + i7(rr) + "." + i7(i) + ".TSGT_" + suf;
        int rc = write_record(h, name, i, DSM_SGT_TSGT, ir);
        if (rc) return rc;
        if (h->in.rsgtswitch) {
            name = od + "/RSGT/" + mn + "." + i7(rr) + "." + i7(i) + ".RSGT_" + suf;
            rc = write_record(h, name, i, DSM_SGT_RSGT, ir);
            if (rc) return rc;
        }
    }
    if (h->in.synnswitch) {
        std::string name = od + "/RSGT/" + mn + "." + i7(h->sh.intir0) + "." + i7(i) + ".SYNN_" + suf;
        int rc = write_record(h, name, i, DSM_SGT_SYNN, 0);
        if (rc) return rc;
    }
    return DSM_SGT_OK;
}

int dsm_sgt_get_stats(dsm_sgt_t *h, dsm_sgt_stats_t *st)
{
    if (!h || !st) return DSM_SGT_EINVAL;
    *st = h->stats;
    return DSM_SGT_OK;
}

int dsm_sgt_reset_stats(dsm_sgt_t *h)
{
    if (!h) return DSM_SGT_EINVAL;
    h->stats = dsm_sgt_stats_t{};
    return DSM_SGT_OK;
}

int dsm_sgt_clear(dsm_sgt_t *h)
{
    if (!h) return DSM_SGT_EINVAL;
    std::fill(h->done.begin(), h->done.end(), 0);
    return DSM_SGT_OK;
}

const char *dsm_sgt_last_error(dsm_sgt_t *h) { return h ? h->err.c_str() : "null handle"; }

void dsm_sgt_destroy(dsm_sgt_t *h)
{
    if (!h) return;
    if (h->stream) {
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
        ev_collect(h);
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
}

}  // extern "C"
