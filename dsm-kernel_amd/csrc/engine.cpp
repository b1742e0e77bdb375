// engine.cpp -- the SGT engine behind the C-ABI of include/dsm_sgt.h.
//
// Pipeline for a batch of frequencies (one HIP stream):
//   host   : per-frequency band pieces (fe_setup_sh.cpp / fe_setup_psv.cpp)
//   [P-SV]   psv_l0    : the l = 0 system and its coefficient rows
//   loop over l chunks until every frequency of the batch is cut off:
//     *_solve  : (omega,l) band LDL^T + the distinct right-hand sides -> stencil values
//     *_coef   : interpolation, strain maps -> contraction rows, cut-off amplitudes
//     host     : replay of the reference's l cut-off (calamp) on the amplitudes
//   zero_tail  : drop the rows past each frequency's cut-off
//   synth      : FP64 MFMA contraction over l into the FP64 catalog in HBM
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstddef>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <map>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "dsm_km.h"
#include "dsm_sgt.h"
#include "fe_setup.h"
#include "km_stage.h"
#include "kernels.h"
#include "legendre_lane.h"
#include "psv_jobs.h"
#include "sh_jobs.h"

using namespace dsm;

namespace {

constexpr double PI = 3.1415926535897932;

// Device blocks released by destroyed engines, kept for the next engine of
// the process: an engine holds ~100 GB, and freeing it and allocating it
// again has stalled engine creation for seconds.  Per device; a released
// block is reused for a request of at least half its size.  Cached blocks
// of a device are freed, largest first, when a newly built engine leaves
// less than kHbmReserve free and before an allocation that would leave
// less, until that much is free (freeing every cached block there, as an
// earlier version did, cost the next engines seconds of fresh allocations
// after an autotuner's engines had filled the cache); an allocation that
// would still leave less fails (ENOMEM).  The runtime allocates kernel scratch (register spills) at
// launch from the same memory, and a queue whose scratch allocation fails
// aborts.  Never torn down (the process exit returns the memory).
class BufCache {
    std::mutex m_;
    std::multimap<std::pair<int, size_t>, void *> free_;

public:
    void *take(int dev, size_t bytes, size_t &got)
    {
        std::lock_guard<std::mutex> g(m_);
        auto it = free_.lower_bound({dev, bytes});
        if (it == free_.end() || it->first.first != dev || it->first.second / 2 > bytes) return nullptr;
        void *p = it->second;
        got = it->first.second;
        free_.erase(it);
        return p;
    }
    void give(int dev, void *p, size_t n)
    {
        std::lock_guard<std::mutex> g(m_);
        free_.insert({{dev, n}, p});
    }
    // free cached blocks of the device, largest first, until the device
    // has `want` bytes free (or the cache is empty)
    void trim(int dev, size_t want)
    {
        std::lock_guard<std::mutex> g(m_);
        for (;;) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr >= want) return;
            auto last = free_.end();
            for (auto it = free_.begin(); it != free_.end(); ++it)
                if (it->first.first == dev) last = it;  // the map is ordered by (device, size)
            if (last == free_.end()) return;
            (void)hipFree(last->second);
            free_.erase(last);
        }
    }
    // bytes of the device's cached blocks (free for the next engine)
    size_t cached(int dev)
    {
        std::lock_guard<std::mutex> g(m_);
        size_t n = 0;
        for (const auto &b : free_)
            if (b.first.first == dev) n += b.first.second;
        return n;
    }
    void flush(int dev)
    {
        std::lock_guard<std::mutex> g(m_);
        for (auto it = free_.begin(); it != free_.end();) {
            if (it->first.first == dev) {
                (void)hipFree(it->second);
                it = free_.erase(it);
            } else {
                ++it;
            }
        }
    }
};
BufCache &buf_cache()
{
    static BufCache *c = new BufCache();
    return *c;
}

constexpr size_t kHbmReserve = (size_t)8 << 30;

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    int dev = 0;
    ~DevBuf()
    {
        if (p) buf_cache().give(dev, p, n);
    }
    void swap(DevBuf &o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(dev, o.dev);
    }
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= n) return hipSuccess;
        if (p) buf_cache().give(dev, p, n);
        p = nullptr;
        n = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        size_t got = 0;
        if ((p = buf_cache().take(dev, bytes, got))) {
            n = got;
            return hipSuccess;
        }
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < bytes + kHbmReserve) {
            buf_cache().trim(dev, bytes + kHbmReserve);
            // still short of the reserve: refuse here (ENOMEM) rather than
            // let a later kernel launch abort the queue
            if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < bytes + kHbmReserve) return hipErrorOutOfMemory;
        }
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            buf_cache().flush(dev);
            e = hipMalloc(&p, bytes);
        }
        if (e == hipSuccess) n = bytes;
        else p = nullptr;
        return e;
    }
    template <class T> T *as() const { return (T *)p; }
};

struct TimedLaunch {
    hipEvent_t a, b;
    int kind;  // 0 solve, 1 legendre, 2 synth, 3 other
    hipStream_t s;
};

}  // namespace

// Host part of one P-SV batch: the per-frequency scalars of
// SGTpsv-MPI.f90:566-625 for every slot (the band pieces themselves are
// assembled on the device, psv_band.hip), prepared on a host thread while
// the previous batch runs on the device.
struct PsvPrep {
    std::vector<int> freqs;
    std::vector<double2> scal;  // [slot][1 + 3 nzone]: the band assembly's scalars (PsvBandArgs::scal)
    std::vector<double2> fk;
    std::vector<PsvSrc> src;
    std::vector<double> omg;
    std::vector<int> lsuf;
};

struct ShPrep {
    std::vector<int> freqs;
    std::vector<double> omg;    // [slot] real angular frequency (depth cut)
    std::vector<double2> band;  // [slot][a0 re/im pair, a2 pair][nn]
    std::vector<double2> om;    // [slot][omega^2, coef at the source]
    std::vector<int> lsuf;
};

struct dsm_sgt {
    int mode = 0, precision = 0, device = 0;
    hipStream_t stream = nullptr;
    // BF16 P-SV forms: the l = 0 systems run on a second stream beside the
    // (omega, l >= 1) solves (default priority, like the engine stream)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_l0in = nullptr, ev_l0out = nullptr;
    // the Legendre tables are computed on stream2 beside the host set-up and
    // the first batch's solves; the engine stream waits on ev_tab before its
    // first synthesis (tab_wait: that wait is still to be enqueued)
    hipEvent_t ev_tab = nullptr;
    bool tab_wait = false;
    // SH: the l chunks of a batch, recorded after each chunk's cut-off state
    // is copied to the host (two in flight)
    hipEvent_t ev_chunk[2] = {nullptr, nullptr};
    // overlap plan (BF16 P-SV forms): the synthesis of batch b runs on
    // stream_syn beside the solves of batch b+1.  The buffers between the
    // two stages are held twice (the *_alt set) and swapped before every
    // batch; ev_ready (engine stream) releases a batch's synthesis, and
    // ev_syn[set] marks the end of the synthesis that last read a set, which
    // the engine stream waits for before it writes that set again.
    bool ovl = false;
    int bset = 0;
    hipStream_t stream_syn = nullptr;
    hipEvent_t ev_ready = nullptr, ev_syn[2] = {nullptr, nullptr};
    std::string err;
    dsm_sgt_input_t in{};
    ShSetup sh;
    PsvSetup psv;
    PsvGeom pg;
    // run geometry (either mode)
    bool ready = false;
    int r_n = 0, theta_n = 0, nn = 0, intir0 = 0;
    int ncomp[3] = {0, 0, 0};
    // the device catalogs hold complex64 (the files' precision) in the
    // mixed forms, complex FP64 in the plain form
    int cat64 = 0;
    size_t cat_el() const { return cat64 ? sizeof(float2) : sizeof(double2); }
    double r0 = 0;
    std::vector<double> radii, theta_deg;
    // geometry of the device arrays
    int theta_pad = 0, lcap = 0, nneed = 0, nslot = 0, lchunk = 0;
    // launch plan: the caller's request (0 = default) and the tile grouping
    // and graph switch in force (xg 0: default grouping, -1: no tile list)
    dsm_sgt_tuning_t tune{};
    int xg = 0, graph = 0;
    int ipos0 = 0, npos = 0;  // the run covers frequencies [ipos0, ipos0+npos)
    // device catalog: ncat frequency positions; catpos[i - ipos0] is the
    // position of frequency i (i - ipos0 unless a catalog window is set:
    // then the frequencies of the last run, in ascending order; -1 when
    // not resident)
    int ncat = 0;
    std::vector<int> catpos;
    std::vector<int> need, rec_slot;
    std::vector<int> llog;    // per position: the reference's logged llog-1
    std::vector<char> done;   // per position: catalog resident on the device
    std::vector<char> logged; // per position: l cut-off known
    // device buffers
    DevBuf d_theta, d_tab, d_tab32, d_need, d_recslot, d_w0, d_w1, d_rinv, d_src;
    // BF16 forms: split operands, distance symmetry
    int ns = 0;                // bf16 terms per operand (0: not a BF16 form)
    int mirror = 0, nh = 0, nhp = 0;
    DevBuf d_tb, d_eb, d_stash, d_nlr, d_tiles;
    DevBuf d_kpart;  // split-K partial sums of the SYNN synthesis
    // overlap plan: the second set of the buffers the synthesis reads
    DevBuf d_eb_alt, d_Es_alt, d_nlr_alt, d_nl_alt, d_slotout_alt, d_tiles_alt, d_kpart_alt;
    void swap_sets()
    {
        d_eb.swap(d_eb_alt);
        d_Es.swap(d_Es_alt);
        d_nlr.swap(d_nlr_alt);
        d_nl.swap(d_nl_alt);
        d_slotout.swap(d_slotout_alt);
        d_tiles.swap(d_tiles_alt);
        d_kpart.swap(d_kpart_alt);
        bset ^= 1;
    }
    DevBuf d_diag;  // [0] small pivots, [1] non-finite catalog values of the current run
    std::vector<int> tiles_main, tiles_acc;  // host copies of the tile orders (launch sizes)
    int tiles_main_n = 0, tiles_acc_n = 0;
    DevBuf d_bfe, d_bzbase, d_bzsolid, d_bznl, d_bzkkdr, d_bcolz, d_bovrii, d_bovr, d_bscal;  // device band assembly
    DevBuf d_plm;  // P_l^m scratch of the P-SV Legendre table
    DevBuf d_band, d_om, d_slots, d_nl, d_slotout, d_scratch, d_xn, d_E, d_Es, d_amp, d_jobs;
    DevBuf d_cat_t, d_cat_r, d_cat_s;
    // P-SV only
    DevBuf d_kind, d_fluid, d_irf, d_fluidK, d_noderow, d_nodec1, d_jsta0, d_l0work, d_gexc, d_ends;
    // depth cut of the batched forms (0: off): per-row radius / wave speed,
    // per-slot angular frequency, first solved row per (slot, l block)
    double dcut = 0.0;
    DevBuf d_rrow, d_vrow, d_omg, d_kstart, d_rlow, d_rneed, d_zs, d_loff, d_lsuf, d_llog, d_wsave;
    // SH l cut-off replayed on the device: [slot] running maximum (double),
    // then [slot] count, lstop, lsuf (int)
    DevBuf d_shcut;
    // fused FP64 family (mixed forms): compact band and its selector bits
    DevBuf d_cband, d_csel;
    // l = 0 receiver rows of the BF16 forms, written beside the solves and
    // copied in after the coefficient kernel: [nq][ns][m] (one degree), slot list
    DevBuf d_l0row, d_l0slots;
    // graph mode: the executable graphs of the current batch (destroyed once
    // the batch has synchronised)
    std::vector<hipGraphExec_t> gexec;
    // l bound of the batched P-SV forms' solve launch: the previous batch's
    // longest cut-off and highest frequency (0: none yet)
    int lb_nl = 0, lb_f = 0;
    std::vector<SynthJob> jobs;
    int njobs_main = 0, njobs_syn = 0, njobs_acc = 0;  // launch order: main, accumulate, SYNN
    int syn_rows = 0;  // coefficient rows per receiver the TSGT/RSGT jobs read (P-SV 63 / 42, SH 36 / 26)
    PsvPrep prep[2];
    ShPrep shprep[2];
    // catalog output: the records of a run packed on the device into
    // complex64 blocks (d_pack, one per frequency of packlist), copied to
    // the host on stream_io and written by the writer thread, beside the
    // engine's next run
    hipStream_t stream_io = nullptr;
    hipEvent_t ev_pack = nullptr;
    DevBuf d_pack, d_packpos;
    std::vector<int> packlist, packpos;
    std::thread writer;
    int io_rc = 0;
    std::string io_err;
    float2 *pin = nullptr;  // pinned staging of the writer (NPIN blocks)
    // pinned host buffer of the read-backs inside a batch: the amplitudes of
    // the P-SV FP64 form's host cut-off replay, SH's per-chunk cut-off state
    // and first solved rows (a strided read-back of every l chunk into
    // pageable memory took milliseconds with the device idle)
    void *pin_amp = nullptr;
    size_t pin_amp_n = 0;
    hipError_t amp_buf(size_t bytes)
    {
        if (bytes <= pin_amp_n) return hipSuccess;
        if (pin_amp) (void)hipHostFree(pin_amp);
        pin_amp = nullptr;
        pin_amp_n = 0;
        const hipError_t e = hipHostMalloc(&pin_amp, bytes, hipHostMallocDefault);
        if (e == hipSuccess) pin_amp_n = bytes;
        return e;
    }
    size_t pin_n = 0;
    // KernelMaker catalog stage (dsm_km_points): one radius's catalog, its
    // phase tables, the points and their outputs
    DevBuf d_km_omt, d_km_omr, d_km_ph, d_km_pos, d_km_thd, d_km_pts, d_km_php, d_km_out;
    DevBuf d_km_fin, d_km_fout, d_km_tw;  // dsm_km_vectorfft: spectra, windows, twiddles
    int km_tw_n = 0;
    double2 *km_static_tw = nullptr;  // the device's static twiddle storage (dsm_km_vectorfft_noalloc)
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

static void ev_begin(dsm_sgt_t *h, TimedLaunch &t, int kind, hipStream_t s = nullptr)
{
    t.kind = kind;
    t.s = s ? s : h->stream;
    if (!h->timing) return;
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
    (void)hipEventRecord(t.a, t.s);
}
static void ev_end(dsm_sgt_t *h, TimedLaunch &t)
{
    if (!h->timing) return;
    (void)hipEventRecord(t.b, t.s);
    h->pending.push_back(t);
}
// Host wall time since t0, in ms.
static double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// hipStreamSynchronize of the engine stream inside a batch, its wait
// counted in stats.device_wait_ms.
static hipError_t dev_sync(dsm_sgt_t *h)
{
    const auto t0 = std::chrono::steady_clock::now();
    const hipError_t e = hipStreamSynchronize(h->stream);
    h->stats.device_wait_ms += ms_since(t0);
    return e;
}

// Adds the finished timed launches to the stats (launches on another
// stream that are still running stay pending).
static void ev_collect(dsm_sgt_t *h)
{
    std::vector<TimedLaunch> left;
    for (auto &t : h->pending) {
        if (hipEventQuery(t.b) == hipErrorNotReady) {
            left.push_back(t);
            continue;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        if (t.kind == 0) h->stats.solve_ms += ms;
        else if (t.kind == 1) h->stats.legendre_ms += ms;
        else if (t.kind == 2) h->stats.synth_ms += ms;
        else h->stats.other_ms += ms;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    h->pending.swap(left);
}

// End of a create: the table launches are on stream2; ev_tab marks their
// end.  The engine's own uploads are complete on return.
static int tables_enqueued(dsm_sgt_t *h)
{
    HIPCHK(h, hipEventRecord(h->ev_tab, h->stream2));
    h->tab_wait = true;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ev_collect(h);
    return DSM_SGT_OK;
}

// The engine stream's wait for the Legendre tables, enqueued once before
// the first synthesis (outside any graph capture).
static hipError_t wait_tables(dsm_sgt_t *h)
{
    if (!h->tab_wait) return hipSuccess;
    h->tab_wait = false;
    return hipStreamWaitEvent(h->stream, h->ev_tab, 0);
}

// Synthesis launch in the engine's precision; Eb is the bf16 split of
// ga.E (BF16 forms).
static void launch_synth(dsm_sgt_t *h, const SynthArgs &ga, const DevBuf &Eb, const int *nlr = nullptr,
                         const int *tiles = nullptr, int ntiles = 0, hipStream_t st = nullptr)
{
    if (!st) st = h->stream;
    if (h->ns) {
        SynthB16Args b{};
        b.Eb = Eb.as<unsigned short>();
        b.Tb = h->d_tb.as<unsigned short>();
        b.nq = ga.nq;
        b.r_eff = ga.r_eff;
        b.nslot = ga.nslot;
        b.lcap = ga.lcap;
        b.nplanes = (h->mode == DSM_SGT_PSV) ? PSV_NPLANES : 12;
        b.theta_n = ga.theta_n;
        b.nh = h->nh;
        b.nhp = h->nhp;
        b.mirror = h->mirror;
        b.nlr = nlr ? nlr : ga.nl;  // [r][slot]; [slot] when r_eff = 1
        b.slot_out = ga.slot_out;
        b.jobs = ga.jobs;
        b.njobs = ga.njobs;
        b.stash = h->d_stash.as<double>();
        b.tiles = tiles;
        b.ntiles = ntiles;
        b.nonfinite = ga.nonfinite;
        launch_synth_b16(b, h->ns, st);
    } else if (h->precision == DSM_SGT_FP32) {
        launch_synth_f32(ga, h->d_tab32.as<float>(), st);
    } else {
        launch_synth_f64(ga, st);
    }
}

// SYNN synthesis in split K: its launch has a few workgroups per job (two
// row tiles of 64 frequencies' surface rows), fewer than the device's CUs,
// each with the whole l series.  Eight parts of the series per theta
// function, fixed by lcap alone, so that every launch plan gives the same
// bits.
static hipError_t synn_split(dsm_sgt_t *h, SynthArgs &gs)
{
    const int nk = (h->lcap + 15) / 16;  // chunks of 16 degrees
    gs.kchunk = (nk + 7) / 8;
    gs.ksplit = (nk + gs.kchunk - 1) / gs.kchunk;
    gs.knb = 0;
    for (int j = 0; j < gs.njobs; j++) {
        const SynthJob &d = h->jobs[(size_t)(gs.jobs - h->d_jobs.as<SynthJob>()) + j];
        gs.knb = std::max(gs.knb, d.nb);
    }
    gs.kpart = nullptr;
    if (gs.ksplit <= 1) return hipSuccess;
    const hipError_t e = h->d_kpart.ensure(sizeof(double) * synth_f64_kpart_size(gs));
    if (e == hipSuccess) gs.kpart = h->d_kpart.as<double>();
    return e;
}

// bf16 split of the coefficient rows E [slot][nq][r_eff][lcap] into Eb:
// degrees below each slot's cut-off d_nl, zeros up to what the synthesis
// reads
static int split_rows(dsm_sgt_t *h, const DevBuf &E, int nslot, int nq, int r_eff, DevBuf &Eb)
{
    if (!h->ns) return DSM_SGT_OK;
    launch_split_rows(E.as<double2>(), nslot, nq, r_eff, h->lcap, h->ns, Eb.as<unsigned short>(), h->d_nl.as<int>(),
                      h->stream);
    HIPCHK(h, hipGetLastError());
    return DSM_SGT_OK;
}

// Distance symmetry of the BF16 and FP64 synthesis kernels: the grid is
// symmetric about 90 degrees and every job's functions share one parity
// under theta -> pi - theta.  par[b]: parity of table plane b without (-1)^l.
static void set_mirror(dsm_sgt_t *h, const int *par)
{
    h->mirror = 0;
    const int n = h->theta_n;
    bool sym = (h->ns > 0 || h->precision == DSM_SGT_FP64) && n >= 2;
    for (int i = 0; sym && i < n; i++)
        if (std::fabs(h->theta_deg[i] + h->theta_deg[n - 1 - i] - 180.0) > 1e-9) sym = false;
    for (auto &j : h->jobs) {
        j.msign = par[j.b[0]];
        for (int t = 1; t < j.nb; t++)
            if (par[j.b[t]] != par[j.b[0]]) sym = false;
    }
    if (const char *e = std::getenv("DSM_SGT_MIRROR"))
        if (std::atoi(e) == 0) sym = false;
    h->mirror = sym ? 1 : 0;
    h->nh = sym ? (n + 1) / 2 : n;
    h->nhp = ((h->nh + 127) / 128) * 128;
}

// Tables of the mixed forms: FP32 copy, or the bf16 split (after set_mirror),
// on stream st.
static int make_mixed_tables(dsm_sgt_t *h, size_t nplanes, hipStream_t st)
{
    if (h->precision == DSM_SGT_FP64) return DSM_SGT_OK;
    if (h->ns) {
        HIPCHK(h, h->d_tb.ensure(sizeof(unsigned short) * h->ns * nplanes * (size_t)h->nhp * h->lcap));
        launch_split_table(h->d_tab.as<double>(), (int)nplanes, h->lcap, h->theta_pad, h->nh, h->nhp, h->ns,
                           h->d_tb.as<unsigned short>(), st);
        HIPCHK(h, hipGetLastError());
        return DSM_SGT_OK;
    }
    const size_t n = nplanes * (size_t)h->lcap * h->theta_pad;
    HIPCHK(h, h->d_tab32.ensure(sizeof(float) * n));
    launch_tab_to_f32(h->d_tab.as<double>(), h->d_tab32.as<float>(), n, st);
    HIPCHK(h, hipGetLastError());
    return DSM_SGT_OK;
}

// Tile order of the main and accumulate launches of the BF16 forms: tiles
// sorted by their l extent (the longest series among the tile's rows),
// longest first, so the long tiles start first and spread over all XCDs.
static int b16_tile_order(dsm_sgt_t *h, const std::vector<int> &nlr, int S, int R)
{
    int nx, ny;
    synth_b16_tiles(S, R, h->nhp, nx, ny);
    const int M = S * R * 2;
    std::vector<int> kt(ny, 0);
    for (int ty = 0; ty < ny; ty++)
        for (int pr = (ty * 128) >> 1; pr <= (std::min(M - 1, ty * 128 + 127) >> 1); pr++)
            kt[ty] = std::max(kt[ty], nlr[pr]);
    // Tile order.  A workgroup's blockIdx goes to XCD blockIdx % 8, and the
    // XCDs have separate L2s.  The nx column tiles of one (job, row block)
    // group read the same coefficient rows, and row blocks of one job read
    // the same table columns: chunks of XG groups (sorted longest series
    // first, so a chunk's groups are of one job and finish together) go to
    // one XCD, XG nx workgroups that share their operands through its L2.
    // Chunks are dealt to the least loaded XCD; slots past an XCD's queue
    // hold -1 (no tile).
    constexpr int NXCD = 8;
    const int XG = h->xg > 0 ? h->xg : std::max(1, 32 / nx);
    // column tiles of a group in full-tile work (the last one may be narrow)
    const double wx = nx - (h->nh - (nx - 1) * 128 <= SYNTH_B16_NARROW ? 0.75 : 0.0);
    auto order = [&](int j0, int nj, std::vector<int> &out) {
        std::vector<int> groups;
        std::vector<double> w(nj * ny);
        for (int j = 0; j < nj; j++)
            for (int ty = 0; ty < ny; ty++) {
                groups.push_back(j * ny + ty);
                const int ce = ((kt[ty] + 1) / 2 + 31) / 32, co = (kt[ty] / 2 + 31) / 32;
                w[j * ny + ty] = (double)h->jobs[j0 + j].nb * (ce + co);
            }
        std::stable_sort(groups.begin(), groups.end(), [&](int a, int b) { return w[a] > w[b]; });
        std::vector<std::vector<int>> q(NXCD);
        double load[NXCD] = {};
        for (size_t c = 0; c < groups.size(); c += XG) {
            int x = 0;
            for (int y = 1; y < NXCD; y++)
                if (load[y] < load[x]) x = y;
            for (size_t g = c; g < std::min(groups.size(), c + XG); g++) {
                load[x] += w[groups[g]] * wx;
                for (int tx = 0; tx < nx; tx++) q[x].push_back(groups[g] * nx + tx);
            }
        }
        size_t len = 0;
        for (auto &v : q) len = std::max(len, v.size());
        out.assign(len * NXCD, -1);
        for (int x = 0; x < NXCD; x++)
            for (size_t i = 0; i < q[x].size(); i++) out[i * NXCD + x] = q[x][i];
    };
    order(0, h->njobs_main, h->tiles_main);
    order(h->njobs_main, h->njobs_acc, h->tiles_acc);
    h->tiles_main_n = (int)h->tiles_main.size();
    h->tiles_acc_n = (int)h->tiles_acc.size();
    HIPCHK(h, h->d_tiles.ensure(sizeof(int) * (h->tiles_main.size() + h->tiles_acc.size())));
    HIPCHK(h, hipMemcpyAsync(h->d_tiles.p, h->tiles_main.data(), sizeof(int) * h->tiles_main.size(),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_tiles.as<int>() + h->tiles_main.size(), h->tiles_acc.data(),
                             sizeof(int) * h->tiles_acc.size(), hipMemcpyHostToDevice, h->stream));
    return DSM_SGT_OK;
}

// MFMA flops the bf16 synthesis issued for the batch, from the rows' l
// extents the kernels used (nlr, device values): per tile and K step of 32
// degrees 2 x 128 x 128 x 32 per cross product (2 x 128 x 32 x 32 for a
// narrow last distance tile).
static double b16_issued_flops(const dsm_sgt_t *h, const std::vector<int> &nlr, int S, int R)
{
    int nx, ny;
    synth_b16_tiles(S, R, h->nhp, nx, ny);
    const double wx = nx - (h->nh - (nx - 1) * 128 <= SYNTH_B16_NARROW ? 0.75 : 0.0);
    const int M = S * R * 2;
    const double per = 2.0 * 128 * 128 * 32 * (h->ns == 3 ? 6 : 3);
    double fl = 0.0;
    for (int ty = 0; ty < ny; ty++) {
        int kt = 0;
        for (int pr = (ty * 128) >> 1; pr <= (std::min(M - 1, ty * 128 + 127) >> 1); pr++) kt = std::max(kt, nlr[pr]);
        const int ce = ((kt + 1) / 2 + 31) / 32, co = (kt / 2 + 31) / 32;
        for (int j = 0; j < h->njobs_main + h->njobs_acc; j++) fl += per * wx * h->jobs[j].nb * (ce + co);
    }
    return fl;
}

// Device buffers of the BF16 forms for nslot frequencies: split receiver
// rows and the stash of the largest launch (the SYNN rows are FP64).
// lpad >= 0: only the degrees l >= lpad are zeroed (the caller writes every
// lower degree a synthesis tile reads in each batch); < 0: all
static int alloc_b16(dsm_sgt_t *h, int nslot, int nq, int R, int lpad)
{
    if (!h->ns) return DSM_SGT_OK;
    const size_t M = (size_t)nslot * R * 2;
    // the synthesis kernel addresses a split's rows with 32-bit offsets
    if (M * h->lcap >= (size_t)1 << 32)
        return fail(h, DSM_SGT_EUNSUPPORTED, "too many receivers x frequencies per batch for the bf16 synthesis");
    HIPCHK(h, h->d_eb.ensure(sizeof(unsigned short) * nq * h->ns * M * h->lcap));
    if (lpad >= 0) {
        launch_zero_pad_b16(h->d_eb.as<unsigned short>(), (size_t)nq * h->ns * M, M, h->lcap, lpad, h->stream);
        HIPCHK(h, hipGetLastError());
    } else {
        HIPCHK(h, hipMemset(h->d_eb.p, 0, h->d_eb.n));
    }
    HIPCHK(h, h->d_nlr.ensure(sizeof(int) * (size_t)nslot * R));
    const size_t g = synth_b16_grid(nslot, R, h->nhp, std::max(h->njobs_main, h->njobs_acc));
    HIPCHK(h, h->d_stash.ensure(sizeof(double) * g * 8 * 32 * 64));
    return DSM_SGT_OK;
}

// Synthesis jobs of the SH contraction (table in sh_jobs.h).
// BF16 forms: a component whose theta functions have both parities under
// theta -> pi - theta (Y0 with Y1; Y2, Z33 with Z23, Y3) becomes two jobs,
// one per parity, the second accumulating, so that every job can use the
// distance symmetry.  Launch order: main (writes), accumulate, SYNN.
static void sh_synth_jobs(dsm_sgt_t *h)
{
    static const int fpar[6] = {1, -1, -1, 1, 1, -1};  // parity of Y0..Y5 without (-1)^(l+mu)
    h->jobs.clear();
    void *outs[3] = {h->d_cat_t.p, h->d_cat_r.p, h->d_cat_s.p};
    const int ncomp[3] = {10, 5, 5};
    std::vector<SynthJob> main, acc, syn;
    h->syn_rows = 0;
    for (int k = 0; k < SH_NJOBS; k++) {
        const ShJobDef &d = SH_JOBS[k];
        // the reference computes RSGT only with rsgtswitch, SYNN only with
        // synnswitch (SGTsh-MPI.f90:517-550)
        if ((d.kind == 1 && !h->in.rsgtswitch) || (d.kind == 2 && !h->in.synnswitch)) continue;
        if (d.kind != 2) h->syn_rows += d.nb;
        SynthJob j{};
        j.comp[0] = d.comp;
        j.comp[1] = -1;
        j.sign[0] = 1.0;
        j.sign[1] = 0.0;
        j.ncomp = ncomp[d.kind];
        j.out = outs[d.kind];
        j.c64 = h->cat64;
        // split point: the first plane whose parity differs from the first's
        // (the planes of one parity are contiguous in SH_JOBS); the FP32
        // kernel has no distance symmetry
        int cut = d.nb;
        if (h->precision != DSM_SGT_FP32)
            for (int t = 1; t < d.nb; t++)
                if (fpar[d.b[t]] != fpar[d.b[0]]) {
                    cut = t;
                    break;
                }
        for (int part = 0; part < 2; part++) {
            const int t0 = part ? cut : 0, t1 = part ? d.nb : cut;
            if (t1 <= t0) continue;
            SynthJob p = j;
            p.q0 = d.q0 + t0;
            p.nb = t1 - t0;
            for (int t = 0; t < 4; t++) p.b[t] = (t < p.nb) ? (d.mu - 1) * 6 + d.b[t0 + t] : (d.mu - 1) * 6;
            p.accumulate = part;
            (d.kind == 2 ? syn : part ? acc : main).push_back(p);
        }
    }
    h->jobs = main;
    h->jobs.insert(h->jobs.end(), acc.begin(), acc.end());
    h->jobs.insert(h->jobs.end(), syn.begin(), syn.end());
    h->njobs_main = (int)main.size();
    h->njobs_acc = (int)acc.size();
    h->njobs_syn = (int)syn.size();
}

// Synthesis jobs of the P-SV contraction (table in psv_jobs.h), ordered
// main (pass 0, TSGT/RSGT), accumulate (pass 1), SYNN.
static void psv_synth_jobs(dsm_sgt_t *h)
{
    h->jobs.clear();
    void *outs[3] = {h->d_cat_t.p, h->d_cat_r.p, h->d_cat_s.p};
    const int ncomp[3] = {20, 10, 10};
    int counts[3] = {0, 0, 0};
    h->syn_rows = 0;
    for (int grp = 0; grp < 3; grp++)
        for (int k = 0; k < PSV_NJOBS; k++) {
            const PsvJobDef &d = PSV_JOBS[k];
            const int g = (d.kind == 2) ? 2 : d.pass;
            if (g != grp) continue;
            // RSGT only with rsgtswitch, SYNN only with synnswitch
            // (SGTpsv-MPI.f90:917-1084)
            if ((d.kind == 1 && !h->in.rsgtswitch) || (d.kind == 2 && !h->in.synnswitch)) continue;
            if (d.kind != 2) h->syn_rows += d.nb;
            SynthJob j{};
            j.q0 = d.q0;
            j.nb = d.nb;
            for (int t = 0; t < 4; t++) j.b[t] = psv_plane(d.mu, d.f[t]);
            j.comp[0] = d.comp0;
            j.comp[1] = d.comp1;
            j.sign[0] = (double)d.sign0;
            j.sign[1] = (double)d.sign1;
            j.accumulate = d.pass;
            j.ncomp = ncomp[d.kind];
            j.out = outs[d.kind];
            j.c64 = h->cat64;
            h->jobs.push_back(j);
            counts[grp]++;
        }
    h->njobs_main = counts[0];
    h->njobs_acc = counts[1];
    h->njobs_syn = counts[2];
}

// The device catalog of ncat frequencies (the kinds the switches keep) must
// fit in the device's free memory with the reserve: without a window, say
// which window would fit rather than fail in an allocation.
static int check_catalog_fits(dsm_sgt_t *h, int R, int T, const int ncomp[3])
{
    const double per = (double)h->cat_el() * T * ((double)R * (ncomp[0] + (h->in.rsgtswitch ? ncomp[1] : 0)) +
                                                  (h->in.synnswitch ? ncomp[2] : 0));
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return DSM_SGT_OK;
    if (per * h->ncat <= (double)fr - (double)kHbmReserve) return DSM_SGT_OK;
    buf_cache().flush(h->device);  // blocks kept from destroyed engines count as free
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return DSM_SGT_OK;
    const double avail = (double)fr - (double)kHbmReserve;
    if (per * h->ncat <= avail) return DSM_SGT_OK;
    const int w = avail > 0 ? (int)(0.5 * avail / per) : 0;  // (half: the batch buffers need the rest)
    return fail(h, DSM_SGT_ENOMEM,
                "the device catalog of " + std::to_string(h->ncat) + " frequencies needs " +
                    std::to_string((long long)(per * h->ncat / 1e9)) + " GB of the " +
                    std::to_string((long long)(fr / 1000000000ULL)) + " GB free: run in catalog windows (" +
                    "dsm_sgt_tuning_t.catalog_window" + (w > 0 ? " <= " + std::to_string(w) : std::string()) + ")");
}

// Catalog positions without a window: frequency i at i - ipos0.
static void init_catpos(dsm_sgt_t *h)
{
    h->catpos.assign(h->npos, -1);
    if (h->ncat >= h->npos)
        for (int p = 0; p < h->npos; p++) h->catpos[p] = p;
}

static int sh_create(dsm_sgt_t *h)
{
    int rc = sh_setup(h->in, h->sh, h->err);
    if (rc) return rc;
    const ShSetup &s = h->sh;
    const int R = s.r_n;
    h->r_n = R;
    h->theta_n = s.theta_n;
    h->nn = s.nn;
    h->intir0 = s.intir0;
    h->r0 = s.r0;
    h->radii = s.r_;
    h->theta_deg = s.theta_deg;
    h->ncomp[0] = 10;
    h->ncomp[1] = 5;
    h->ncomp[2] = 5;
    h->theta_pad = ((s.theta_n + 127) / 128) * 128;  // synthesis N tile
    h->lcap = ((s.maxlmax + 1 + 63) / 64) * 64;
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
    h->nslot = h->tune.nslot > 0 ? h->tune.nslot : 32;
    // (whole blocks of 64 degrees: the solve's scratch is laid out per block)
    h->lchunk = h->tune.lchunk > 0 ? (std::max(64, h->tune.lchunk) + 63) / 64 * 64 : 512;
    h->ipos0 = std::max(0, h->in.imin);
    h->npos = std::max(0, h->in.imax - h->ipos0 + 1);
    h->llog.assign(h->npos, 0);
    h->done.assign(h->npos, 0);
    h->logged.assign(h->npos, 0);
    h->ncat = h->tune.catalog_window > 0 ? std::min(h->tune.catalog_window, std::max(1, h->npos)) : h->npos;
    init_catpos(h);
    {
        const int nc[3] = {10, 5, 5};
        if (const int rc = check_catalog_fits(h, R, s.theta_n, nc)) return rc;
    }

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
    {
        // the FP64 table on stream2, beside the rest of the set-up
        TimedLaunch t;
        ev_begin(h, t, 1, h->stream2);
        launch_legendre_sh(h->d_theta.as<double>(), s.theta_n, h->theta_pad, s.maxlmax, h->lcap,
                           h->d_tab.as<double>(), h->stream2);
        HIPCHK(h, hipGetLastError());
        ev_end(h, t);
    }

    const int S = h->nslot, nn = s.nn;
    HIPCHK(h, h->d_band.ensure(sizeof(double2) * 4 * (size_t)nn * S));
    HIPCHK(h, h->d_om.ensure(sizeof(double2) * 2 * S));
    HIPCHK(h, h->d_slots.ensure(sizeof(int) * S));
    HIPCHK(h, h->d_nl.ensure(sizeof(int) * S));
    HIPCHK(h, h->d_slotout.ensure(sizeof(int) * S));
    const size_t nthreads = (size_t)S * h->lchunk;
    HIPCHK(h, h->d_scratch.ensure(sizeof(double2) * (size_t)(2 * ((nn + SH_SEG - 1) / SH_SEG) + 20) * nthreads));
    HIPCHK(h, h->d_xn.ensure(sizeof(double2) * (size_t)S * h->nneed * SH_NSOL * h->lcap));
    HIPCHK(h, h->d_E.ensure(sizeof(double2) * (size_t)S * 36 * R * h->lcap));
    HIPCHK(h, h->d_Es.ensure(sizeof(double2) * (size_t)S * 5 * h->lcap));
    HIPCHK(h, h->d_amp.ensure(sizeof(double) * (size_t)S * h->lcap));
    HIPCHK(h, h->d_shcut.ensure((sizeof(double) + 3 * sizeof(int)) * (size_t)S));
    const size_t np = (size_t)std::max(1, h->ncat);
    HIPCHK(h, h->d_cat_t.ensure(h->cat_el() * np * R * s.theta_n * 10));
    if (h->in.rsgtswitch) HIPCHK(h, h->d_cat_r.ensure(h->cat_el() * np * R * s.theta_n * 5));
    if (h->in.synnswitch) HIPCHK(h, h->d_cat_s.ensure(h->cat_el() * np * s.theta_n * 5));
    // depth cut of the mixed forms (DSM_SGT_DEPTH_CUT overrides; 0 disables):
    // node radius and smaller shear speed per band row
    h->dcut = (h->precision == DSM_SGT_FP64) ? 0.0 : 40.0;
    if (const char *e = std::getenv("DSM_SGT_DEPTH_CUT")) h->dcut = std::max(0.0, std::atof(e));
    if (h->dcut > 0.0) {
        std::vector<double> rr(s.nn), vv(s.nn);
        for (int j = 0; j < s.nn; j++) {
            const double r = s.ra[j];
            int iz = 0;
            while (iz + 1 < s.nzone && r > s.vrmax[iz]) iz++;
            double x = 1.0, a = 0.0, c = 0.0;
            for (int q = 0; q < 4; q++) {
                a += s.vsv[4 * iz + q] * x;
                c += s.vsh[4 * iz + q] * x;
                x *= r / s.rmax;
            }
            rr[j] = r;
            vv[j] = std::min(a, c);
        }
        const int nlbmax = (h->lchunk + 63) / 64 + 1;
        HIPCHK(h, h->d_rrow.ensure(sizeof(double) * s.nn));
        HIPCHK(h, h->d_vrow.ensure(sizeof(double) * s.nn));
        HIPCHK(h, h->d_omg.ensure(sizeof(double) * S));
        HIPCHK(h, h->d_kstart.ensure(sizeof(int) * (size_t)S * nlbmax));
        HIPCHK(h, hipMemcpy(h->d_rrow.p, rr.data(), sizeof(double) * s.nn, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->d_vrow.p, vv.data(), sizeof(double) * s.nn, hipMemcpyHostToDevice));
    }
    sh_synth_jobs(h);
    {
        // parity of Y0..Y5 (legendre_lane.h) times (-1)^mu
        static const int fpar[6] = {1, -1, -1, 1, 1, -1};
        int par[12];
        for (int mu = 1; mu <= 2; mu++)
            for (int b = 0; b < 6; b++) par[(mu - 1) * 6 + b] = fpar[b] * ((mu & 1) ? -1 : 1);
        set_mirror(h, par);
    }
    // the split of every batch writes each degree a tile reads (the terms
    // below each cut-off, zeros up to the longest cut-off + 128): only the
    // padding degrees above lmax are zeroed here
    rc = alloc_b16(h, S, 36, R, s.maxlmax + 1);
    if (rc) return rc;
    HIPCHK(h, h->d_jobs.ensure(sizeof(SynthJob) * h->jobs.size()));
    HIPCHK(h, hipMemcpy(h->d_jobs.p, h->jobs.data(), sizeof(SynthJob) * h->jobs.size(), hipMemcpyHostToDevice));

    TimedLaunch t;
    ev_begin(h, t, 1, h->stream2);
    rc = make_mixed_tables(h, 12, h->stream2);
    if (rc) return rc;
    ev_end(h, t);
    rc = tables_enqueued(h);
    if (rc) return rc;
    h->ready = true;
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

// Host part of one SH batch: the band pieces a0, a2 and the scalars of
// every frequency (SGTsh-MPI.f90:392-420), prepared on a host thread while
// the previous batch runs on the device.
static void sh_prepare(const ShSetup &s, const std::vector<int> &freqs, ShPrep &p)
{
    const int S = (int)freqs.size(), nn = s.nn;
    p.freqs = freqs;
    p.band.resize(4 * (size_t)nn * S);
    p.om.resize(2 * (size_t)S);
    p.lsuf.resize(S);
    p.omg.resize(S);
    for (int k = 0; k < S; k++) {
        ShOmega o;
        sh_omega_setup(s, freqs[k], o);
        p.omg[k] = o.omega;
        for (int j = 0; j < nn; j++) {
            p.band[(size_t)(k * 4 + 0) * nn + j] = make_double2(o.a0[2 * j].real(), o.a0[2 * j].imag());
            p.band[(size_t)(k * 4 + 1) * nn + j] = make_double2(o.a0[2 * j + 1].real(), o.a0[2 * j + 1].imag());
            p.band[(size_t)(k * 4 + 2) * nn + j] = make_double2(o.a2[2 * j].real(), o.a2[2 * j].imag());
            p.band[(size_t)(k * 4 + 3) * nn + j] = make_double2(o.a2[2 * j + 1].real(), o.a2[2 * j + 1].imag());
        }
        p.om[2 * k] = make_double2(o.comega2.real(), o.comega2.imag());
        p.om[2 * k + 1] = make_double2(o.coef_spn.real(), o.coef_spn.imag());
        p.lsuf[k] = o.lsuf;
    }
}

static int sh_run_batch(dsm_sgt_t *h, const ShPrep &p)
{
    const ShSetup &s = h->sh;
    const std::vector<int> &freqs = p.freqs;
    const std::vector<double2> &band = p.band, &om = p.om;
    const int S = (int)freqs.size(), nn = s.nn, R = s.r_n, LC = h->lcap;
    std::vector<int> slotout(S);
    for (int k = 0; k < S; k++) slotout[k] = h->catpos[freqs[k] - h->ipos0];  // catalog position
    HIPCHK(h, hipMemcpyAsync(h->d_band.p, band.data(), sizeof(double2) * band.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_om.p, om.data(), sizeof(double2) * om.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_slotout.p, slotout.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    if (h->dcut > 0.0)
        HIPCHK(h, hipMemcpyAsync(h->d_omg.p, p.omg.data(), sizeof(double) * S, hipMemcpyHostToDevice, h->stream));

    // every slot runs the l chunks until its cut-off, which the device
    // replays after each chunk (sh_cut_kernel); later chunks skip the slots
    // whose cut-off is known.  The host reads each chunk's state one chunk
    // behind: chunk c is queued before the state after chunk c-1 is read, so
    // the device does not wait for the host, and the chunk queued after the
    // last slot finished does no work.
    std::vector<int> all(S);
    for (int k = 0; k < S; k++) all[k] = k;
    HIPCHK(h, hipMemcpyAsync(h->d_slots.p, all.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    std::vector<unsigned char> st((sizeof(double) + 3 * sizeof(int)) * (size_t)S, 0);  // carry 0, count 0
    {
        int *lst = (int *)(st.data() + sizeof(double) * S) + S;
        for (int k = 0; k < S; k++) {
            lst[k] = INT_MAX;
            lst[S + k] = p.lsuf[k];
        }
    }
    HIPCHK(h, hipMemcpyAsync(h->d_shcut.p, st.data(), st.size(), hipMemcpyHostToDevice, h->stream));
    double *cut_carry = h->d_shcut.as<double>();
    int *cut_count = (int *)(cut_carry + S), *cut_lstop = cut_count + S, *cut_lsuf = cut_lstop + S;
    const int nlbmax = (h->lchunk + 63) / 64;
    const int nchunk = s.maxlmax / h->lchunk + 1;
    // pinned: per chunk [S] lstop, then [S][nlbmax] first solved rows
    const size_t per_chunk = (size_t)S * (1 + nlbmax);
    HIPCHK(h, h->amp_buf(sizeof(int) * per_chunk * nchunk));
    int *pin = (int *)h->pin_amp;
    auto all_cut = [&](int c) {
        for (int k = 0; k < S; k++)
            if (pin[per_chunk * c + k] == INT_MAX) return false;
        return true;
    };
    int c = 0, last = -1;  // last: the chunk whose state holds every cut-off
    for (; c < nchunk; c++) {
        const int l0 = c * h->lchunk, lend = std::min(l0 + h->lchunk - 1, s.maxlmax);
        const int nlb = (lend - l0 + 64) / 64;
        const int nblocks = S * nlb;
        ShSolveArgs sa{};
        sa.small_pivots = h->d_diag.as<unsigned long long>();
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
        sa.lstop = cut_lstop;
        TimedLaunch t0;
        ev_begin(h, t0, 0);
        if (h->dcut > 0.0) {  // mixed forms: first solved row per block of 64 degrees
            ShDepthArgs da{h->d_rrow.as<double>(), h->d_vrow.as<double>(), h->d_omg.as<double>(),
                           h->d_slots.as<int>(), S, nlb, l0, s.ns, h->dcut, 0.35,
                           h->d_kstart.as<int>(), cut_lstop};
            launch_sh_depth(da, h->stream);
            sa.jstart = h->d_kstart.as<int>();
        }
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
        ca.lstop = cut_lstop;
        launch_sh_coef(ca, nblocks, h->stream);
        HIPCHK(h, hipGetLastError());
        ShCutArgs cu{h->d_amp.as<double>(), cut_lsuf, LC, l0, lend, s.maxlmax, s.ratl, cut_carry, cut_count, cut_lstop};
        launch_sh_cut(cu, S, h->stream);
        HIPCHK(h, hipGetLastError());
        ev_end(h, t0);
        h->stats.solve_launches++;
        int *pc = pin + per_chunk * c;
        if (h->dcut > 0.0)
            HIPCHK(h, hipMemcpyAsync(pc + S, h->d_kstart.p, sizeof(int) * nblocks, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(pc, cut_lstop, sizeof(int) * S, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipEventRecord(h->ev_chunk[c & 1], h->stream));
        if (c > 0) {
            HIPCHK(h, hipEventSynchronize(h->ev_chunk[(c - 1) & 1]));
            if (all_cut(c - 1)) {
                last = c - 1;
                c++;
                break;
            }
        }
    }
    const int nqueued = c;
    if (last < 0) {  // the last chunk ends at lmax: its state holds every cut-off
        last = nqueued - 1;
        HIPCHK(h, hipEventSynchronize(h->ev_chunk[last & 1]));
    }
    std::vector<int> nl(S);
    for (int k = 0; k < S; k++) {
        nl[k] = pin[per_chunk * last + k];
        h->llog[freqs[k] - h->ipos0] = std::min(s.maxlmax, nl[k]) - 1;
        h->done[freqs[k] - h->ipos0] = 1;
        h->logged[freqs[k] - h->ipos0] = 1;
    }
    // systems solved: every degree of the chunks up to the one whose state
    // found the slot's cut-off (that chunk is solved whole); rows per system
    // from the depth cut's first row of its block of 64 degrees
    for (int k = 0; k < S; k++) {
        const int ck = std::min(nl[k], s.maxlmax) / h->lchunk;
        const int nsys = std::min(s.maxlmax, (ck + 1) * h->lchunk - 1) + 1;
        h->stats.systems += nsys;
        // algorithmic bytes: band read once, 14 nonzero right-hand sides'
        // (value, derivative) at every receiver written once (w = 2)
        h->stats.solve_bytes += (double)nsys * 16.0 * (2.0 * nn + 4.0 * 14 * R);
        // algorithmic bytes of the reference's systems l = 1 .. nl-1: the
        // band (w = 2) of the rows solved and the (value, derivative) of 14
        // right-hand sides at every receiver
        for (int lb0 = 0; lb0 < nl[k]; lb0 += 64) {
            const int la = std::max(1, lb0), lb = std::min(nl[k], lb0 + 64);
            if (lb <= la) continue;
            int ne = nn;
            if (h->dcut > 0.0) {
                const int cc = lb0 / h->lchunk, nlbc = (std::min((cc + 1) * h->lchunk - 1, s.maxlmax) - cc * h->lchunk + 64) / 64;
                ne = nn - pin[per_chunk * cc + S + (size_t)k * nlbc + (lb0 - cc * h->lchunk) / 64] + 1;
            }
            h->stats.solve_alg_bytes += (double)(lb - la) * 16.0 * (2.0 * ne + 4.0 * 14 * R);
        }
    }
    HIPCHK(h, hipMemcpyAsync(h->d_nl.p, nl.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    TimedLaunch tz;
    ev_begin(h, tz, 3);
    // (the bf16 split writes the zeros of the receiver rows of the BF16
    // forms; the SYNN rows stay FP64)
    if (!h->ns)
        launch_zero_tail(h->d_E.as<double2>(), S, 36 * R, h->d_slots.as<int>(), h->d_nl.as<int>(), LC, h->stream);
    launch_zero_tail(h->d_Es.as<double2>(), S, 5, h->d_slots.as<int>(), h->d_nl.as<int>(), LC, h->stream);
    HIPCHK(h, hipGetLastError());
    int rc = split_rows(h, h->d_E, S, 36, R, h->d_eb);
    if (rc) return rc;
    std::vector<int> nlr;  // [r][slot] l extents of the bf16 rows
    if (h->ns) {
        nlr.resize((size_t)R * S);
        for (int r = 0; r < R; r++)
            for (int k = 0; k < S; k++) nlr[(size_t)r * S + k] = nl[k];
        HIPCHK(h, hipMemcpyAsync(h->d_nlr.p, nlr.data(), sizeof(int) * nlr.size(), hipMemcpyHostToDevice, h->stream));
        h->stats.synth_mfma_flops += b16_issued_flops(h, nlr, S, R);
    }
    ev_end(h, tz);

    SynthArgs ga{};
    ga.nonfinite = h->d_diag.as<unsigned long long>() + 1;
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
    ga.mirror = h->mirror;
    ga.nh = h->nh;
    HIPCHK(h, wait_tables(h));
    // SYNN rows (surface node, no receiver dimension): FP64 in every form
    // (as P-SV)
    SynthArgs gs = ga;
    gs.E = h->d_Es.as<double2>();
    gs.nq = SH_NQS;
    gs.r_eff = 1;
    gs.jobs = h->d_jobs.as<SynthJob>() + h->njobs_main + h->njobs_acc;
    gs.njobs = h->njobs_syn;
    if (h->njobs_syn) HIPCHK(h, synn_split(h, gs));
    TimedLaunch tg;
    ev_begin(h, tg, 2);
    launch_synth(h, ga, h->d_eb, h->d_nlr.as<int>());
    HIPCHK(h, hipGetLastError());
    if (h->njobs_acc) {  // the second parity of split components
        ga.jobs = h->d_jobs.as<SynthJob>() + h->njobs_main;
        ga.njobs = h->njobs_acc;
        launch_synth(h, ga, h->d_eb, h->d_nlr.as<int>());
        HIPCHK(h, hipGetLastError());
    }
    ev_end(h, tg);
    if (h->njobs_syn) {
        TimedLaunch tg2;
        ev_begin(h, tg2, 3);
        launch_synth_f64(gs, h->stream);
        HIPCHK(h, hipGetLastError());
        ev_end(h, tg2);
    }
    h->stats.synth_launches += h->njobs_acc ? 2 : 1;
    double sumL = 0.0;
    for (int k = 0; k < S; k++) sumL += nl[k];
    // 4 flops per (complex coefficient x real basis) per (r,theta); 36 rows
    // per receiver (26 without RSGT)
    h->stats.synth_flops += 4.0 * sumL * s.theta_n * (double)h->syn_rows * R;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ev_collect(h);
    return DSM_SGT_OK;
}

// ------------------------------------------------------------------------
// P-SV
// ------------------------------------------------------------------------
static int psv_create(dsm_sgt_t *h)
{
    // the FP64 distance table first, on stream2: it needs only the distances
    // and the degree range, and runs beside the host set-up below (whose
    // distances are the same theta_grid)
    bool tab_early = false;
    if ((h->in.theta_list_n > 0 || h->in.thetadelta > 0) && h->in.maxlmax >= 1) {
        std::vector<double> deg;
        const int tn = theta_grid(h->in, deg);
        if (tn >= 1 && tn <= (1 << 20)) {
            const int LC = ((h->in.maxlmax + 1 + 63) / 64) * 64, tp = ((tn + 127) / 128) * 128;
            std::vector<double> th(tn);
            for (int i = 0; i < tn; i++) th[i] = deg[i] / 180.0 * PI;
            HIPCHK(h, h->d_theta.ensure(sizeof(double) * tn));
            HIPCHK(h, h->d_tab.ensure(sizeof(double) * PSV_NPLANES * (size_t)LC * tp));
            HIPCHK(h, h->d_plm.ensure(sizeof(double) * 4 * (size_t)LC * tp));
            HIPCHK(h, hipMemcpy(h->d_theta.p, th.data(), sizeof(double) * tn, hipMemcpyHostToDevice));
            TimedLaunch t;
            ev_begin(h, t, 1, h->stream2);
            launch_legendre_psv(h->d_theta.as<double>(), tn, tp, h->in.maxlmax, LC, h->d_plm.as<double>(),
                                h->d_tab.as<double>(), h->stream2);
            HIPCHK(h, hipGetLastError());
            ev_end(h, t);
            tab_early = true;
        }
    }
    int rc = psv_setup(h->in, h->psv, h->err);
    if (rc) return rc;
    const PsvSetup &s = h->psv;
    h->pg = psv_geometry(s);
    const PsvGeom &g = h->pg;
    const int R = s.r_n;
    h->r_n = R;
    h->theta_n = s.theta_n;
    h->nn = s.nn;
    h->intir0 = s.intir0;
    h->r0 = s.r0;
    h->radii = s.r_;
    h->theta_deg = s.theta_deg;
    h->ncomp[0] = 20;
    h->ncomp[1] = 10;
    h->ncomp[2] = 10;
    h->theta_pad = ((s.theta_n + 127) / 128) * 128;
    h->lcap = ((s.maxlmax + 1 + 63) / 64) * 64;
    h->nneed = (int)g.need.size();
    // the batched forms store the factors of the depth-cut systems only,
    // which leaves room for four times more frequencies per batch
    h->nslot = (h->precision == DSM_SGT_FP64) ? 32 : 128;
    if (const char *e = std::getenv("DSM_SGT_NSLOT")) h->nslot = std::max(1, std::atoi(e));
    if (h->tune.nslot > 0) h->nslot = h->tune.nslot;
    // one solve launch covers the whole l range unless told otherwise (the
    // batched forms need it: the cut-off is replayed on the device)
    h->lchunk = s.maxlmax + 1;
    if (const char *e = std::getenv("DSM_SGT_LCHUNK")) h->lchunk = std::max(64, std::atoi(e));
    if (h->tune.lchunk > 0 && h->precision == DSM_SGT_FP64) h->lchunk = std::max(64, h->tune.lchunk);
    h->ipos0 = std::max(0, h->in.imin);
    h->npos = std::max(0, h->in.imax - h->ipos0 + 1);
    h->llog.assign(h->npos, 0);
    h->done.assign(h->npos, 0);
    h->logged.assign(h->npos, 0);
    h->ncat = h->tune.catalog_window > 0 ? std::min(h->tune.catalog_window, std::max(1, h->npos)) : h->npos;
    init_catpos(h);
    {
        const int nc[3] = {20, 10, 10};
        if (const int rc = check_catalog_fits(h, R, s.theta_n, nc)) return rc;
    }

    std::vector<double> th(s.theta_n);
    for (int i = 0; i < s.theta_n; i++) th[i] = s.theta_deg[i] / 180.0 * PI;
    std::vector<int> kind(s.nn);
    psv_pack_kind(s, kind.data());
    const int S = h->nslot, nn = s.nn, LC = h->lcap, N0 = s.nn0;
    HIPCHK(h, h->d_theta.ensure(sizeof(double) * s.theta_n));
    HIPCHK(h, h->d_tab.ensure(sizeof(double) * PSV_NPLANES * (size_t)LC * h->theta_pad));
    HIPCHK(h, h->d_need.ensure(sizeof(int) * h->nneed));
    HIPCHK(h, h->d_recslot.ensure(sizeof(int) * 6 * R));
    HIPCHK(h, h->d_w0.ensure(sizeof(double) * 3 * R));
    HIPCHK(h, h->d_w1.ensure(sizeof(double) * 3 * R));
    HIPCHK(h, h->d_irf.ensure(sizeof(double) * R));
    HIPCHK(h, h->d_fluid.ensure(sizeof(int) * R));
    HIPCHK(h, h->d_kind.ensure(sizeof(int) * nn));
    HIPCHK(h, h->d_noderow.ensure(sizeof(int) * N0));
    HIPCHK(h, h->d_nodec1.ensure(sizeof(int) * N0));
    HIPCHK(h, h->d_jsta0.ensure(sizeof(int) * R));
    if (!tab_early) HIPCHK(h, hipMemcpy(h->d_theta.p, th.data(), sizeof(double) * s.theta_n, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_need.p, g.need.data(), sizeof(int) * h->nneed, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_recslot.p, g.rec_row.data(), sizeof(int) * 6 * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_w0.p, s.w0.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_w1.p, s.w1.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_irf.p, g.irf.data(), sizeof(double) * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_fluid.p, s.fluid_rec.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_kind.p, kind.data(), sizeof(int) * nn, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_noderow.p, s.node_row.data(), sizeof(int) * N0, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_nodec1.p, s.node_c1.data(), sizeof(int) * N0, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_jsta0.p, s.jsta0.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    if (!tab_early) {
        // the FP64 table on stream2, beside the rest of the set-up
        TimedLaunch t;
        ev_begin(h, t, 1, h->stream2);
        HIPCHK(h, h->d_plm.ensure(sizeof(double) * 4 * (size_t)LC * h->theta_pad));
        launch_legendre_psv(h->d_theta.as<double>(), s.theta_n, h->theta_pad, s.maxlmax, LC, h->d_plm.as<double>(),
                            h->d_tab.as<double>(), h->stream2);
        HIPCHK(h, hipGetLastError());
        ev_end(h, t);
    }

    HIPCHK(h, h->d_band.ensure(sizeof(double2) * 16 * (size_t)nn * S));
    {
        // resident tables of the device band assembly
        PsvBandTables bt;
        psv_band_tables(s, bt);
        for (int j = 0; j < nn; j++) {
            int z = 0;
            for (int iz = 0; iz < s.nzone; iz++) {
                const int ncol = s.fe[iz].solid ? 2 * s.fe[iz].nlayer + 2 : s.fe[iz].nlayer + 1;
                if (j + 1 >= s.kkdr[iz] && j + 1 < s.kkdr[iz] + ncol) z++;
            }
            if (z > 2) return fail(h, DSM_SGT_EUNSUPPORTED, "a band column shared by more than two zones");
        }
        auto up = [&](DevBuf &d, const void *src, size_t bytes) -> int {
            HIPCHK(h, d.ensure(std::max<size_t>(bytes, 16)));
            HIPCHK(h, hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
            return DSM_SGT_OK;
        };
        int rc2 = up(h->d_bfe, bt.fe.data(), sizeof(double) * bt.fe.size());
        if (!rc2) rc2 = up(h->d_bzbase, bt.zbase.data(), sizeof(long long) * bt.zbase.size());
        if (!rc2) rc2 = up(h->d_bzsolid, bt.zsolid.data(), sizeof(int) * bt.zsolid.size());
        if (!rc2) rc2 = up(h->d_bznl, bt.znl.data(), sizeof(int) * bt.znl.size());
        if (!rc2) rc2 = up(h->d_bzkkdr, bt.zkkdr.data(), sizeof(int) * bt.zkkdr.size());
        if (!rc2) rc2 = up(h->d_bcolz, bt.colz.data(), sizeof(int) * bt.colz.size());
        if (!rc2) rc2 = up(h->d_bovrii, bt.ovr_ii.data(), sizeof(int) * bt.ovr_ii.size());
        if (!rc2) rc2 = up(h->d_bovr, bt.ovr.data(), sizeof(double) * bt.ovr.size());
        if (rc2) return rc2;
        HIPCHK(h, h->d_bscal.ensure(sizeof(double2) * (size_t)S * (1 + 3 * s.nzone)));
    }
    HIPCHK(h, h->d_src.ensure(sizeof(PsvSrc) * S));
    HIPCHK(h, h->d_fluidK.ensure(sizeof(double2) * (size_t)S * R));
    HIPCHK(h, h->d_slots.ensure(sizeof(int) * S));
    HIPCHK(h, h->d_nl.ensure(sizeof(int) * S));
    HIPCHK(h, h->d_slotout.ensure(sizeof(int) * S));
    HIPCHK(h, h->d_l0work.ensure(sizeof(double2) * 5 * (size_t)N0 * S));
    const int n = nn - s.row0 + 1, klo = g.rowlo - s.row0 + 1, fs = s.jtmp - s.row0 + 1;
    const int lch = std::min(h->lchunk, s.maxlmax + 1);
    const size_t nthreads = (size_t)S * (psv_solve_blocks(0, lch - 1) + 1) * 256;
    HIPCHK(h, h->d_zs.ensure(sizeof(double2) * (size_t)4 * (n - fs + 1) * nthreads));
    // batched forms: depth cut (DSM_SGT_DEPTH_CUT overrides; 0 disables)
    h->dcut = (h->precision == DSM_SGT_FP64) ? 0.0 : 40.0;
    if (const char *e = std::getenv("DSM_SGT_DEPTH_CUT")) h->dcut = std::max(0.0, std::atof(e));
    // factor storage of full systems: the FP64 form, and every form without
    // the depth cut (the cut systems of the batched forms are sized per batch)
    if (h->precision == DSM_SGT_FP64 || h->dcut <= 0.0)
        HIPCHK(h, h->d_scratch.ensure(sizeof(double2) * (size_t)3 * (n - klo + 1) * nthreads));
    HIPCHK(h, h->d_loff.ensure(sizeof(long long) * (size_t)S * (psv_solve_blocks(0, lch - 1) + 1) + 16));
    HIPCHK(h, h->d_gexc.ensure(sizeof(double2) * (size_t)S * 16 * LC));
    HIPCHK(h, h->d_ends.ensure(sizeof(double2) * 3 * nthreads));
    HIPCHK(h, h->d_wsave.ensure(sizeof(double2) * 6 * nthreads));
    {
        std::vector<double> rr, vv;
        psv_row_profile(s, rr, vv);
        HIPCHK(h, h->d_rrow.ensure(sizeof(double) * nn));
        HIPCHK(h, h->d_vrow.ensure(sizeof(double) * nn));
        HIPCHK(h, hipMemcpy(h->d_rrow.p, rr.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->d_vrow.p, vv.data(), sizeof(double) * nn, hipMemcpyHostToDevice));
        HIPCHK(h, h->d_omg.ensure(sizeof(double) * S));
        if (h->precision != DSM_SGT_FP64) {
            std::vector<int> sel(nn, 0);
            for (int j = 0; j < nn; j++)
                for (int ii = 0; ii < 4; ii++)
                    if (s.term_kind[4 * (size_t)j + ii] == 2) sel[j] |= 1 << ii;
            HIPCHK(h, h->d_csel.ensure(sizeof(int) * nn));
            HIPCHK(h, hipMemcpy(h->d_csel.p, sel.data(), sizeof(int) * nn, hipMemcpyHostToDevice));
            HIPCHK(h, h->d_cband.ensure(sizeof(double2) * 8 * (size_t)nn * S));
        }
        HIPCHK(h, h->d_kstart.ensure(sizeof(int) * (size_t)S * (psv_solve_blocks(0, lch - 1) + 1)));
        std::vector<int> rlow(R);
        for (int r = 0; r < R; r++) rlow[r] = s.ksta[r] - s.row0;
        HIPCHK(h, h->d_rlow.ensure(sizeof(int) * R));
        HIPCHK(h, hipMemcpy(h->d_rlow.p, rlow.data(), sizeof(int) * R, hipMemcpyHostToDevice));
        HIPCHK(h, h->d_rneed.ensure((size_t)R * (psv_solve_blocks(0, lch - 1) + 1)));
        HIPCHK(h, h->d_lsuf.ensure(sizeof(int) * S));
        HIPCHK(h, h->d_llog.ensure(sizeof(int) * S));
    }
    HIPCHK(h, h->d_xn.ensure(sizeof(double2) * (size_t)S * h->nneed * PSV_NSOL * LC));
    // the bf16 forms write the receiver rows split (alloc_b16); the SYNN
    // rows are FP64 in every form
    if (!h->ns) HIPCHK(h, h->d_E.ensure(sizeof(double2) * (size_t)S * PSV_NQ * R * LC));
    HIPCHK(h, h->d_Es.ensure(sizeof(double2) * (size_t)S * PSV_NQS * LC));
    HIPCHK(h, h->d_amp.ensure(sizeof(double2) * 2 * (size_t)S * LC));
    const size_t np = (size_t)std::max(1, h->ncat);
    HIPCHK(h, h->d_cat_t.ensure(h->cat_el() * np * R * s.theta_n * 20));
    if (h->in.rsgtswitch) HIPCHK(h, h->d_cat_r.ensure(h->cat_el() * np * R * s.theta_n * 10));
    if (h->in.synnswitch) HIPCHK(h, h->d_cat_s.ensure(h->cat_el() * np * s.theta_n * 10));
    psv_synth_jobs(h);
    {
        // parity of F0..F9 (legendre_lane.h) times (-1)^mu
        static const int fpar[10] = {1, -1, 1, -1, 1, 1, -1, -1, 1, 1};
        int par[PSV_NPLANES];
        for (int mu = 0; mu <= 2; mu++)
            for (int f = 0; f < 10; f++) {
                if (mu == 0 && !(f <= 2 || f == 9)) continue;
                par[psv_plane(mu, f)] = fpar[f] * ((mu & 1) ? -1 : 1);
            }
        set_mirror(h, par);
    }
    // P-SV batches write every degree up to maxlmax that a tile reads:
    // coefficient rows, l = 0 rows, cut-off tails, zero rows below the cut
    rc = alloc_b16(h, S, PSV_NQ, R, s.maxlmax + 1);
    if (rc) return rc;
    // overlap plan (BF16 forms, eager launches): the second set of the
    // buffers the synthesis reads, initialised as the first
    h->ovl = h->tune.overlap == 1 && h->ns && !h->graph;
    if (h->ovl) {
        h->swap_sets();
        HIPCHK(h, h->d_Es.ensure(h->d_Es_alt.n));
        HIPCHK(h, h->d_nl.ensure(h->d_nl_alt.n));
        HIPCHK(h, h->d_slotout.ensure(h->d_slotout_alt.n));
        rc = alloc_b16(h, S, PSV_NQ, R, s.maxlmax + 1);
        h->swap_sets();
        if (rc) return rc;
    }
    if (h->ns) {
        HIPCHK(h, h->d_l0row.ensure(sizeof(unsigned short) * PSV_NQ * h->ns * (size_t)S * R * 2));
        HIPCHK(h, h->d_l0slots.ensure(sizeof(int) * S));
    }
    HIPCHK(h, h->d_jobs.ensure(sizeof(SynthJob) * h->jobs.size()));
    HIPCHK(h, hipMemcpy(h->d_jobs.p, h->jobs.data(), sizeof(SynthJob) * h->jobs.size(), hipMemcpyHostToDevice));

    TimedLaunch t;
    ev_begin(h, t, 1, h->stream2);
    rc = make_mixed_tables(h, PSV_NPLANES, h->stream2);
    if (rc) return rc;
    ev_end(h, t);
    rc = tables_enqueued(h);
    if (rc) return rc;
    h->ready = true;
    return DSM_SGT_OK;
}

// Host set-up of the frequencies `freqs` into `p`, split over host threads
// (psv_omega_setup only reads the set-up).  Returns false if the pinned
// buffer could not be allocated.
static void psv_prepare(const PsvSetup &s, const std::vector<int> &freqs, PsvPrep &p)
{
    const int S = (int)freqs.size(), R = s.r_n, nz = s.nzone;
    p.freqs = freqs;
    p.scal.assign((size_t)S * (1 + 3 * nz), cz());
    p.fk.assign((size_t)S * R, cz());
    p.src.assign(S, PsvSrc{});
    p.omg.assign(S, 0.0);
    p.lsuf.assign(S, 0);
    auto d2 = [](const cd &c) { return make_double2(c.real(), c.imag()); };
    for (int k = 0; k < S; k++) {
        PsvOmega o;
        psv_omega_scalars(s, freqs[k], o);
        p.omg[k] = o.omega;
        double2 *sc = &p.scal[(size_t)k * (1 + 3 * nz)];
        sc[0] = d2(o.comega2);
        for (int iz = 0; iz < nz; iz++) {
            sc[1 + 3 * iz] = d2(o.zc1[iz]);
            sc[2 + 3 * iz] = d2(o.zc2[iz]);
            sc[3 + 3 * iz] = d2(o.zcf[iz]);
        }
        for (int q = 0; q < 160; q++) {
            p.src[k].anum[q] = d2(o.anum[q]);
            p.src[k].bnum[q] = d2(o.bnum[q]);
        }
        p.src[k].coef1 = d2(o.coef1_spn);
        p.src[k].coef2 = d2(o.coef2_spn);
        for (int r = 0; r < R; r++) p.fk[(size_t)k * R + r] = d2(o.fluidK[r]);
        p.lsuf[k] = o.lsuf;
    }
}

// Band pieces of S frequencies into d_band slots slot0.. (psv_band_kernel).
static int psv_device_band(dsm_sgt_t *h, const std::vector<double2> &scal, int S, int slot0)
{
    const PsvSetup &s = h->psv;
    HIPCHK(h, hipMemcpyAsync(h->d_bscal.p, scal.data(), sizeof(double2) * scal.size(), hipMemcpyHostToDevice,
                             h->stream));
    PsvBandArgs ba{};
    ba.fe = h->d_bfe.as<double>();
    ba.zbase = h->d_bzbase.as<long long>();
    ba.zsolid = h->d_bzsolid.as<int>();
    ba.znl = h->d_bznl.as<int>();
    ba.zkkdr = h->d_bzkkdr.as<int>();
    ba.colz = h->d_bcolz.as<int>();
    ba.ovr_ii = h->d_bovrii.as<int>();
    ba.ovr = h->d_bovr.as<double2>();
    ba.nz = s.nzone;
    ba.nn = s.nn;
    ba.scal = h->d_bscal.as<double2>();
    ba.slot0 = slot0;
    ba.band = h->d_band.as<double2>();
    launch_psv_band(ba, S, h->stream);
    HIPCHK(h, hipGetLastError());
    return DSM_SGT_OK;
}

// Graph mode: the launches between begin_capture and end_capture_launch are
// recorded on the engine stream (kernels only: no memset or copy nodes) and
// run as one graph.  The executable graph lives until the batch has
// synchronised (free_graphs).
struct CaptureGuard {  // ends a capture left open by an early return
    hipStream_t s = nullptr;
    bool on = false;
    ~CaptureGuard()
    {
        if (!on) return;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
    }
};

static int begin_capture(dsm_sgt_t *h, CaptureGuard &cg)
{
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    cg.s = h->stream;
    cg.on = true;
    return DSM_SGT_OK;
}

// t (may be null): timing events around the graph's launch only, after
// its instantiation
static int end_capture_launch(dsm_sgt_t *h, CaptureGuard &cg, TimedLaunch *t, int kind)
{
    hipGraph_t g = nullptr;
    cg.on = false;
    const hipError_t le = hipGetLastError();
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    HIPCHK(h, le);
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(h, e);
    h->gexec.push_back(x);
    if (t) ev_begin(h, *t, kind);
    HIPCHK(h, hipGraphLaunch(x, h->stream));
    if (t) ev_end(h, *t);
    return DSM_SGT_OK;
}

static void free_graphs(dsm_sgt_t *h)
{
    for (hipGraphExec_t x : h->gexec) (void)hipGraphExecDestroy(x);
    h->gexec.clear();
}

// l extent per (receiver, slot) row from the host's view of a batch: the
// cut-off nl, and the leading blocks of 256 degrees whose solved rows
// (first solved row kst per block) reach the receiver's stencil.  The
// tile order is built from it; the kernels use the device's own values.
static void psv_nlr_estimate(const PsvSetup &s, const std::vector<int> &nl, const std::vector<int> &kst_host,
                             int kst_nblk, int S, std::vector<int> &nlr)
{
    const int R = s.r_n;
    nlr.assign((size_t)R * S, 0);
    for (int r = 0; r < R; r++) {
        const int rlow = s.ksta[r] - s.row0;  // local row of ksta-1
        for (int k = 0; k < S; k++) {
            int lr = nl[k];
            if (!kst_host.empty()) {
                int nb = 0;
                while (nb < kst_nblk && kst_host[(size_t)k * kst_nblk + nb] <= rlow) nb++;
                lr = std::min(lr, 256 * nb);
            }
            nlr[(size_t)r * S + k] = lr;
        }
    }
}

// psv_run_batch's answer when a batch's cut-off was not reached within
// its l bound: run the batch again with the full l range
constexpr int PSV_RETRY_FULL = -100;

static int psv_run_batch(dsm_sgt_t *h, const PsvPrep &p, bool full_range)
{
    const PsvSetup &s = h->psv;
    const PsvGeom &g = h->pg;
    const std::vector<int> &freqs = p.freqs;
    const int S = (int)freqs.size(), nn = s.nn, R = s.r_n, LC = h->lcap;
    const std::vector<PsvSrc> &src = p.src;
    const std::vector<double2> &fk = p.fk;
    const std::vector<double> &omg = p.omg;
    std::vector<CutState> cut(S);
    std::vector<int> slotout(S);
    for (int k = 0; k < S; k++) {
        cut[k].lsuf = p.lsuf[k];
        cut[k].llog = s.maxlmax;
        slotout[k] = h->catpos[freqs[k] - h->ipos0];  // catalog position
    }
    // overlap plan: this batch's set of synthesis buffers was last read by
    // the synthesis two batches back
    if (h->ovl) HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_syn[h->bset], 0));
    {
        const int rc = psv_device_band(h, p.scal, S, 0);
        if (rc) return rc;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_src.p, src.data(), sizeof(PsvSrc) * S, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_fluidK.p, fk.data(), sizeof(double2) * fk.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_omg.p, omg.data(), sizeof(double) * S, hipMemcpyHostToDevice, h->stream));
    if (h->d_lsuf.p)
        HIPCHK(h, hipMemcpyAsync(h->d_lsuf.p, p.lsuf.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_slotout.p, slotout.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    std::vector<int> active(S);
    for (int k = 0; k < S; k++) active[k] = k;
    HIPCHK(h, hipMemcpyAsync(h->d_slots.p, active.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    const PsvSrcConst sc{s.ecC0, s.ecF0, s.ecL0, s.r0, s.ra[s.mtmp - 1], s.ra[s.mtmp]};
    // l = 0 of every slot (bf16 forms: after the coefficient kernel, whose
    // first block stages zeros at l = 0)
    PsvL0Args l0args{};
    {
        PsvL0Args za{};
        za.band = h->d_band.as<double2>();
        za.kind = h->d_kind.as<int>();
        za.slot_list = h->d_slots.as<int>();
        za.nslots = S;
        za.nn = nn;
        za.nn0 = s.nn0;
        za.row0_0 = s.row0_0;
        za.jtmp = s.jtmp;
        za.js = s.js0;
        za.node_row = h->d_noderow.as<int>();
        za.node_c1 = h->d_nodec1.as<int>();
        za.work = h->d_l0work.as<double2>();
        za.r_n = R;
        za.jsta0 = h->d_jsta0.as<int>();
        za.fluid = h->d_fluid.as<int>();
        za.w0 = h->d_w0.as<double>();
        za.w1 = h->d_w1.as<double>();
        za.irf = h->d_irf.as<double>();
        za.fluidK = h->d_fluidK.as<double2>();
        za.src = h->d_src.as<PsvSrc>();
        za.sc = sc;
        za.E = h->d_E.as<double2>();
        za.Es = h->d_Es.as<double2>();
        za.lcap = LC;
        if (h->ns) {
            // one-degree rows (lcap 1) beside the solves, on the second
            // stream once the band and the scalars are in place
            za.eb = B16Rows{h->d_l0row.as<unsigned short>(), h->ns, S, R, 1};
            za.slot_list = h->d_l0slots.as<int>();
            HIPCHK(h, hipMemcpyAsync(h->d_l0slots.p, active.data(), sizeof(int) * S, hipMemcpyHostToDevice,
                                     h->stream));
            HIPCHK(h, hipEventRecord(h->ev_l0in, h->stream));
            HIPCHK(h, hipStreamWaitEvent(h->stream2, h->ev_l0in, 0));
            // rows the l = 0 system leaves unwritten stay zero (not captured)
            HIPCHK(h, hipMemsetAsync(h->d_l0row.p, 0, h->d_l0row.n, h->stream2));
            launch_psv_l0(za, h->stream2);
            HIPCHK(h, hipGetLastError());
            HIPCHK(h, hipEventRecord(h->ev_l0out, h->stream2));
        }
        l0args = za;
        if (!h->ns) {
            TimedLaunch t0;
            ev_begin(h, t0, 0);
            launch_psv_l0(za, h->stream);
            HIPCHK(h, hipGetLastError());
            ev_end(h, t0);
        }
    }
    const int n = nn - s.row0 + 1, klo = g.rowlo - s.row0 + 1, fs = s.jtmp - s.row0 + 1;
    // graph mode (BF16 forms with the depth cut, whole l range in one
    // launch): everything after the factor-storage read-back runs as two
    // graphs, the solve-to-row-extent sequence and the synthesis pair
    const bool gmode = h->graph && h->ns && h->dcut > 0.0 && h->d_lsuf.p && h->lchunk > s.maxlmax;
    // Batched forms (one solve launch, cut-off replayed on the device): the
    // l range of the launch is bounded by the previous batch's longest
    // cut-off, scaled to this batch's highest frequency, plus 10% and 128
    // degrees.  A slot whose cut-off is not reached within the bound sends
    // the batch back for the full range (the same systems, the same bits):
    // deep sources stop at a small fraction of lmax.
    // (DSM_SGT_LBOUND sets the 1.1 factor; 0 turns the bound off)
    int lmaxb = s.maxlmax;
    double lbf = 1.1;
    if (const char *e = std::getenv("DSM_SGT_LBOUND")) lbf = std::atof(e);
    if (!full_range && lbf > 0.0 && h->precision != DSM_SGT_FP64 && h->d_lsuf.p && h->lchunk > s.maxlmax &&
        h->lb_nl > 0 && h->lb_f > 0) {
        const double g = lbf * h->lb_nl * ((double)freqs.back() / h->lb_f) + 128.0 * (lbf >= 1.0);
        if (g < s.maxlmax) lmaxb = std::min(s.maxlmax, ((int)g + 64) / 64 * 64 - 1);
    }
    CaptureGuard cg;
    TimedLaunch t_solve;  // events of the current solve launch
    std::vector<int> nlr_est;
    HIPCHK(h, h->amp_buf(sizeof(double2) * (size_t)S * LC * 2));
    double2 *ampuv = (double2 *)h->pin_amp;
    std::vector<int> kst_host;  // depth cut per (slot, block of 256 l), single-launch case
    int kst_nblk = 0;
    int l0 = 0;
    bool dev_cut = false;
    while (!active.empty()) {
        const int lend = std::min(l0 + h->lchunk - 1, lmaxb);
        const int nlb = (lend - l0 + 64) / 64;
        const int nblocks = (int)active.size() * nlb;
        // batched forms in one launch: the cut-off is replayed on the device
        dev_cut = h->precision != DSM_SGT_FP64 && h->d_lsuf.p && l0 == 0 && lend == lmaxb;
        HIPCHK(h, hipMemcpyAsync(h->d_slots.p, active.data(), sizeof(int) * active.size(), hipMemcpyHostToDevice,
                                 h->stream));
        PsvSolveArgs sa{};
        sa.band = h->d_band.as<double2>();
        sa.kind = h->d_kind.as<int>();
        sa.slot_list = h->d_slots.as<int>();
        sa.nlb = nlb;
        sa.l0 = l0;
        sa.lend = lend;
        sa.nn = nn;
        sa.row0 = s.row0;
        sa.jtmp = s.jtmp;
        sa.rowlo = g.rowlo;
        sa.nneed = h->nneed;
        sa.need = h->d_need.as<int>();
        sa.scratch = h->d_scratch.as<double2>();
        sa.sstride = (long)active.size() * psv_solve_blocks(l0, lend) * 256;
        sa.xn = h->d_xn.as<double2>();
        sa.lcap = LC;
        sa.src = h->d_src.as<PsvSrc>();
        sa.sc = sc;
        sa.gexc = h->d_gexc.as<double2>();
        sa.ends = h->d_ends.as<double2>();
        sa.kstart = (h->dcut > 0.0) ? h->d_kstart.as<int>() : nullptr;
        sa.cband = h->d_cband.as<double2>();
        sa.csel = h->d_csel.as<int>();
        sa.zs = h->d_zs.as<double2>();
        sa.wsave = h->d_wsave.as<double2>();
        sa.loff = nullptr;
        TimedLaunch &t0 = t_solve;
        ev_begin(h, t0, 0);
        if (h->precision != DSM_SGT_FP64 && l0 == 0)
            launch_psv_compact(h->d_band.as<double2>(), h->d_kind.as<int>(), nn, h->d_slots.as<int>(),
                               (int)active.size(), h->d_cband.as<double2>(), h->stream);
        if (h->dcut > 0.0) {
            PsvDepthArgs da{};
            da.rrow = h->d_rrow.as<double>();
            da.vrow = h->d_vrow.as<double>();
            da.omega = h->d_omg.as<double>();
            da.slot_list = h->d_slots.as<int>();
            da.nslots = (int)active.size();
            da.nblk = psv_solve_blocks(l0, lend);
            da.l0 = l0;
            da.row0 = s.row0;
            da.jtmp = s.jtmp;
            da.dcut = h->dcut;
            da.dcap = 0.35;
            da.kstart = h->d_kstart.as<int>();
            launch_psv_depth(da, h->stream);
            if (l0 == 0 && lend == lmaxb) {
                // factor storage of the depth-cut systems: offsets per block, total
                const int nbt = (int)active.size() * da.nblk;
                long long *loff = h->d_loff.as<long long>();
                // mixed forms: checkpoints of the factorization below the source
                sa.ckpt = h->precision != DSM_SGT_FP64 ? 1 : 0;
                launch_psv_loff(h->d_kstart.as<int>(), nbt, n, klo, fs, sa.ckpt, loff, loff + nbt, h->stream);
                long long total = 0;
                HIPCHK(h, hipMemcpyAsync(&total, loff + nbt, sizeof(long long), hipMemcpyDeviceToHost, h->stream));
                kst_nblk = da.nblk;
                kst_host.resize((size_t)S * kst_nblk);
                HIPCHK(h, hipMemcpyAsync(kst_host.data(), h->d_kstart.p, sizeof(int) * kst_host.size(),
                                         hipMemcpyDeviceToHost, h->stream));
                HIPCHK(h, dev_sync(h));
                HIPCHK(h, h->d_scratch.ensure(sizeof(double2) * (size_t)std::max(total, 1LL)));
                sa.scratch = h->d_scratch.as<double2>();
                sa.loff = loff;
            } else if (h->precision != DSM_SGT_FP64) {
                return fail(h, DSM_SGT_EUNSUPPORTED, "the depth cut needs one solve launch per batch (DSM_SGT_LCHUNK)");
            }
            if (gmode) {
                // the synthesis tile order needs only the read-back depth
                // cut: built and copied before the capture
                psv_nlr_estimate(s, std::vector<int>(S, s.maxlmax + 1), kst_host, kst_nblk, S, nlr_est);
                if (h->xg >= 0) {
                    const int rc = b16_tile_order(h, nlr_est, S, R);
                    if (rc) return rc;
                }
                ev_end(h, t0);  // the eager part of the solve stage
                // (the capture takes no event of another stream: the l = 0
                // stream is joined before it)
                HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_l0out, 0));
                const int rc = begin_capture(h, cg);
                if (rc) return rc;
            }
            if (h->ns && l0 == 0 && lend == lmaxb) {
                PsvNeedArgs na{h->d_kstart.as<int>(), h->d_rlow.as<int>(), (int)active.size(), da.nblk, R,
                               h->d_rneed.as<unsigned char>()};
                launch_psv_need(na, h->stream);
            }
        }
        // every solve path stores factors (or checkpoints) in the scratch
        // buffer: refuse the launch rather than let it address null
        if (!sa.scratch || (sa.ckpt && !sa.loff))
            return fail(h, DSM_SGT_ESTATE, "band solve launched without factor storage");
        launch_psv_excite(sa, (int)active.size(), h->stream);
        launch_psv_solve(sa, (int)active.size(), h->precision != DSM_SGT_FP64, h->stream);
        HIPCHK(h, hipGetLastError());
        PsvCoefArgs ca{};
        ca.nq = h->in.rsgtswitch ? PSV_NQ : PSV_NQT;  // the RSGT rows only with rsgtswitch
        ca.xn = h->d_xn.as<double2>();
        ca.slot_list = h->d_slots.as<int>();
        ca.nlb = nlb;
        ca.l0 = l0;
        ca.lend = lend;
        ca.lcap = LC;
        ca.nneed = h->nneed;
        ca.r_n = R;
        ca.rec_row = h->d_recslot.as<int>();
        ca.fluid = h->d_fluid.as<int>();
        ca.w0 = h->d_w0.as<double>();
        ca.w1 = h->d_w1.as<double>();
        ca.irf = h->d_irf.as<double>();
        ca.fluidK = h->d_fluidK.as<double2>();
        ca.amp_i0 = g.amp_i0;
        ca.amp_i1 = g.amp_i1;
        ca.syn_i0 = g.syn_i0;
        ca.syn_i1 = g.syn_i1;
        ca.E = h->d_E.as<double2>();
        ca.Es = h->d_Es.as<double2>();
        ca.ampuv = h->d_amp.as<double2>();
        if (h->dcut > 0.0) {  // receivers below a block's solved rows: zero rows
            ca.kstart = h->d_kstart.as<int>();
            ca.rlow = h->d_rlow.as<int>();
            ca.nblk = psv_solve_blocks(l0, lend);
        }
        if (h->ns) {
            ca.eb = B16Rows{h->d_eb.as<unsigned short>(), h->ns, S, R, LC};
            if (h->dcut > 0.0 && l0 == 0 && lend == lmaxb) ca.rneed = h->d_rneed.as<unsigned char>();
        }
        launch_psv_coef(ca, nblocks, h->stream);
        HIPCHK(h, hipGetLastError());
        if (h->ns && l0 == 0) {  // the l = 0 rows into degree 0 (the coefficient kernel staged zeros there)
            if (!gmode) HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_l0out, 0));
            launch_b16_put_l0(h->d_l0row.as<unsigned short>(), h->d_eb.as<unsigned short>(),
                              (size_t)PSV_NQ * h->ns * S * R * 2, (size_t)S * R * 2, LC, h->stream);
            HIPCHK(h, hipGetLastError());
        }
        if (dev_cut) {
            PsvCutArgs cu{h->d_amp.as<double2>(), h->d_lsuf.as<int>(), S, LC, lmaxb, s.ratl, h->d_nl.as<int>(),
                          h->d_llog.as<int>()};
            launch_psv_cut(cu, h->stream);
            HIPCHK(h, hipGetLastError());
        }
        if (!gmode) ev_end(h, t0);
        h->stats.solve_launches++;
        const long nsys = (long)active.size() * (lend - std::max(l0, 1) + 1);
        h->stats.systems += nsys;
        // bytes of the design: factor rows L (3 complex) written and read
        // back from the deepest receiver row up, forward values of the 4
        // source solutions from the source row up, 6 solutions at the
        // recorded rows
        h->stats.solve_bytes += (double)nsys * 16.0 * (2.0 * 3 * (n - klo + 1) + 2.0 * 4 * (n - fs + 1) +
                                                       (double)PSV_NSOL * h->nneed);
        if (dev_cut) {  // the whole l range in one launch, cut-off replayed on the device
            active.clear();
            break;
        }
        HIPCHK(h, hipMemcpy2DAsync(&ampuv[2 * (size_t)l0], sizeof(double2) * 2 * LC,
                                   h->d_amp.as<double2>() + 2 * (size_t)l0, sizeof(double2) * 2 * LC,
                                   sizeof(double2) * 2 * (lend - l0 + 1), S, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, dev_sync(h));
        // replay the cut-off in the reference's order (SGTpsv-MPI.f90:616-634, 884-888)
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
                if (l == 0) continue;
                const double2 u = ampuv[((size_t)k * LC + l) * 2], v = ampuv[((size_t)k * LC + l) * 2 + 1];
                const double a1 = std::hypot(u.x, u.y), a2 = std::hypot(v.x, v.y);
                const double a0 = std::sqrt(a1 * a1 + a2 * a2);
                for (int m = -2; m <= 2; m++) {
                    if (std::abs(m) > l) continue;
                    calamp(c, m == 0 ? a0 : 0.0, l, s.ratl, false);
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
        nl[k] = dev_cut ? s.maxlmax + 1 : cut[k].nl;  // device cut-off: known after the batch
        all[k] = k;
        if (!dev_cut) {
            h->llog[freqs[k] - h->ipos0] = cut[k].llog - 1;
            h->done[freqs[k] - h->ipos0] = 1;
            h->logged[freqs[k] - h->ipos0] = 1;
        }
    }
    if (!dev_cut) HIPCHK(h, hipMemcpyAsync(h->d_nl.p, nl.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    // (one launch over the whole l range: the slot list is already all slots)
    if (!gmode) HIPCHK(h, hipMemcpyAsync(h->d_slots.p, all.data(), sizeof(int) * S, hipMemcpyHostToDevice, h->stream));
    TimedLaunch tz;
    if (!gmode) ev_begin(h, tz, 3);
    int rc = DSM_SGT_OK;
    if (h->ns) {
        // bf16 rows written by the coefficient kernels: zero the cut-off
        // tails; l extent per (receiver, slot): the l cut-off, and the
        // blocks of l whose solved rows reach the receiver's stencil
        launch_zero_tail_b16(h->d_eb.as<unsigned short>(), PSV_NQ, h->ns, S, R, LC, h->d_nl.as<int>(), s.maxlmax,
                             h->stream);
        launch_zero_tail(h->d_Es.as<double2>(), S, PSV_NQS, h->d_slots.as<int>(), h->d_nl.as<int>(), LC, h->stream);
        if (h->dcut > 0.0 && !kst_host.empty()) {
            launch_psv_nlr(h->d_kstart.as<int>(), h->d_rlow.as<int>(), h->d_nl.as<int>(), S, kst_nblk, R,
                           h->d_nlr.as<int>(), h->stream);
        } else {
            psv_nlr_estimate(s, nl, kst_host, kst_nblk, S, nlr_est);
            HIPCHK(h, hipMemcpyAsync(h->d_nlr.p, nlr_est.data(), sizeof(int) * nlr_est.size(), hipMemcpyHostToDevice,
                                     h->stream));
        }
    } else {
        launch_zero_tail(h->d_E.as<double2>(), S, PSV_NQ * R, h->d_slots.as<int>(), h->d_nl.as<int>(), LC, h->stream);
        launch_zero_tail(h->d_Es.as<double2>(), S, PSV_NQS, h->d_slots.as<int>(), h->d_nl.as<int>(), LC, h->stream);
    }
    HIPCHK(h, hipGetLastError());
    if (gmode) {  // solve graph: excitation .. row extents
        TimedLaunch tga;
        rc = end_capture_launch(h, cg, &tga, 0);
        if (rc) return rc;
    } else {
        ev_end(h, tz);
    }
    if (dev_cut && lmaxb < s.maxlmax) {
        // the bound holds if every slot's cut-off fell within it
        HIPCHK(h, hipMemcpyAsync(nl.data(), h->d_nl.p, sizeof(int) * S, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, dev_sync(h));
        for (int k = 0; k < S; k++)
            if (nl[k] > lmaxb) {
                free_graphs(h);
                ev_collect(h);
                return PSV_RETRY_FULL;
            }
    }
    if (h->ns && !gmode) {  // host estimate of the l extents for the tile order
        psv_nlr_estimate(s, nl, kst_host, kst_nblk, S, nlr_est);
        if (h->xg >= 0) rc = b16_tile_order(h, nlr_est, S, R);
        if (rc) return rc;
    }
    const bool tiled = h->ns && h->xg >= 0;

    SynthArgs ga{};
    ga.nonfinite = h->d_diag.as<unsigned long long>() + 1;
    ga.E = h->d_E.as<double2>();
    ga.nq = PSV_NQ;
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
    ga.mirror = h->mirror;
    ga.nh = h->nh;
    HIPCHK(h, wait_tables(h));
    // SYNN rows (10 per frequency, a few percent of the synthesis): FP64
    // MFMA in every form.  The surface displacement of a shallow source at
    // far distances is a near-field series that cancels over l, where the
    // FP32 sums of the mixed forms lose up to 1e-4 of it.
    SynthArgs gs = ga;
    gs.E = h->d_Es.as<double2>();
    gs.nq = PSV_NQS;
    gs.r_eff = 1;
    gs.jobs = h->d_jobs.as<SynthJob>() + h->njobs_main + h->njobs_acc;
    gs.njobs = h->njobs_syn;
    if (h->njobs_syn) HIPCHK(h, synn_split(h, gs));
    // overlap plan: the synthesis on its own stream, once the engine stream
    // has written this batch's rows, extents and tile order
    const hipStream_t ss = h->ovl ? h->stream_syn : h->stream;
    if (h->ovl) {
        HIPCHK(h, hipEventRecord(h->ev_ready, h->stream));
        HIPCHK(h, hipStreamWaitEvent(ss, h->ev_ready, 0));
    }
    TimedLaunch tg;
    if (gmode) {
        rc = begin_capture(h, cg);
        if (rc) return rc;
    } else {
        ev_begin(h, tg, 2, ss);
    }
    launch_synth(h, ga, h->d_eb, h->d_nlr.as<int>(), tiled ? h->d_tiles.as<int>() : nullptr, h->tiles_main_n, ss);
    HIPCHK(h, hipGetLastError());
    ga.jobs = h->d_jobs.as<SynthJob>() + h->njobs_main;
    ga.njobs = h->njobs_acc;
    launch_synth(h, ga, h->d_eb, h->d_nlr.as<int>(), tiled ? h->d_tiles.as<int>() + h->tiles_main_n : nullptr,
                 h->tiles_acc_n, ss);
    HIPCHK(h, hipGetLastError());
    if (gmode) {  // synthesis graph: the main and the accumulate launch
        rc = end_capture_launch(h, cg, &tg, 2);
        if (rc) return rc;
    } else {
        ev_end(h, tg);
    }
    if (h->njobs_syn) {
        TimedLaunch tg2;
        ev_begin(h, tg2, 3, ss);
        launch_synth_f64(gs, ss);
        HIPCHK(h, hipGetLastError());
        ev_end(h, tg2);
    }
    if (h->ovl) HIPCHK(h, hipEventRecord(h->ev_syn[h->bset], ss));
    h->stats.synth_launches += 2;
    std::vector<int> nlr_dev;
    if (h->ns) {  // the l extents the synthesis used, for its issued-flop count
        nlr_dev.resize((size_t)R * S);
        HIPCHK(h, hipMemcpyAsync(nlr_dev.data(), h->d_nlr.p, sizeof(int) * nlr_dev.size(), hipMemcpyDeviceToHost,
                                 h->stream));
    }
    if (dev_cut) {  // the cut-off replayed on the device
        std::vector<int> llog(S);
        HIPCHK(h, hipMemcpyAsync(nl.data(), h->d_nl.p, sizeof(int) * S, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(llog.data(), h->d_llog.p, sizeof(int) * S, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, dev_sync(h));
        for (int k = 0; k < S; k++) {
            h->llog[freqs[k] - h->ipos0] = llog[k] - 1;
            h->done[freqs[k] - h->ipos0] = 1;
            h->logged[freqs[k] - h->ipos0] = 1;
        }
        h->lb_nl = *std::max_element(nl.begin(), nl.end());
        h->lb_f = freqs.back();
    }
    double sumL = 0.0;
    for (int k = 0; k < S; k++) sumL += nl[k];
    // 4 flops per (complex coefficient x real function) per (r,theta)
    h->stats.synth_flops += 4.0 * sumL * s.theta_n * (double)h->syn_rows * R;
    // algorithmic bytes of the reference's systems l = 1 .. nl-1: the band
    // (w = 4 entries per row) of the rows solved -- from the depth cut's
    // first row of the system's block of 256 degrees down to the surface --
    // and (U, V, U', V') of 18 right-hand sides at every receiver
    // (summed per block of 256 degrees, whose systems share the cut: this
    // runs between two batches, with the device idle)
    for (int k = 0; k < S; k++)
        for (int l0b = 0; l0b < nl[k]; l0b += 256) {
            const int la = std::max(1, l0b), lb = std::min(nl[k], l0b + 256);
            if (lb <= la) continue;
            int ne = n;
            if (!kst_host.empty() && (l0b >> 8) < kst_nblk) ne = n - kst_host[(size_t)k * kst_nblk + (l0b >> 8)] + 1;
            h->stats.solve_alg_bytes += (double)(lb - la) * 16.0 * (4.0 * ne + 4.0 * 18 * R);
        }
    HIPCHK(h, dev_sync(h));
    free_graphs(h);
    if (h->ns) h->stats.synth_mfma_flops += b16_issued_flops(h, nlr_dev, S, R);
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
    return dsm_sgt_create_tuned(in, mode, precision, device, nullptr, handle);
}

// The inputs the set-up reads, checked before anything is computed: zone
// count, frequency and degree ranges, and every model and grid value finite
// (the reference reads them unchecked, and a non-finite value would reach
// the catalogs as NaN).  Empty string: valid.
static std::string check_input(const dsm_sgt_input_t &in)
{
    if (in.nzone < 1 || in.nzone > DSM_SGT_MAXZONE)
        return "nzone " + std::to_string(in.nzone) + " outside 1.." + std::to_string(DSM_SGT_MAXZONE);
    if (in.maxlmax < 0) return "maxlmax is negative";
    if (in.imin < 0 || in.imax < in.imin) return "frequency range imin..imax is empty or negative";
    auto finite = [](double v) { return std::isfinite(v); };
    for (int z = 0; z < in.nzone; z++) {
        bool ok = finite(in.vrmin[z]) && finite(in.vrmax[z]) && finite(in.qmu[z]) && finite(in.qkappa[z]);
        for (int c = 0; c < 4; c++)
            ok = ok && finite(in.rho[z][c]) && finite(in.vpv[z][c]) && finite(in.vph[z][c]) && finite(in.vsv[z][c]) &&
                 finite(in.vsh[z][c]) && finite(in.eta[z][c]);
        if (!ok) return "zone " + std::to_string(z + 1) + " of the model holds a non-finite value";
    }
    const double grid[] = {in.re, in.ratc, in.ratl, in.omegai_seed, in.tlen, in.rmin_, in.rmax_,
                           in.rdelta_, in.r0min, in.thetamin, in.thetamax, in.thetadelta};
    for (double v : grid)
        if (!finite(v)) return "a configuration or grid value is not finite";
    if (!(in.tlen > 0.0)) return "tlen must be positive";
    if (!(in.omegai_seed > 0.0)) return "the omegai seed must be positive (omegai = -log(seed) / tlen)";
    // receiver lists: present when counted, finite, strictly ascending
    auto list_ok = [&](int32_t n, const double *v, double lo, double hi) {
        if (n < 0 || n > (1 << 20) || (n > 0 && !v)) return false;
        for (int32_t i = 0; i < n; i++)
            if (!finite(v[i]) || v[i] < lo || v[i] > hi || (i > 0 && !(v[i] > v[i - 1]))) return false;
        return true;
    };
    if (!list_ok(in.r_list_n, in.r_list, 0.0, 1.0e9))
        return "the receiver radius list must be finite, positive and strictly ascending";
    if (!list_ok(in.theta_list_n, in.theta_list, 0.0, 180.0))
        return "the distance list must be finite, within 0..180 degrees and strictly ascending";
    return "";
}

int dsm_sgt_create_tuned(const dsm_sgt_input_t *in, int32_t mode, int32_t precision, int32_t device,
                         const dsm_sgt_tuning_t *tuning, dsm_sgt_t **handle)
{
    if (!in || !handle) return DSM_SGT_EINVAL;
    *handle = nullptr;
    dsm_sgt_t *h = new (std::nothrow) dsm_sgt_t();
    if (!h) return DSM_SGT_ENOMEM;
    if (tuning) {
        h->tune = *tuning;
        h->xg = tuning->xcd_group < 0 ? -1 : tuning->xcd_group;
        h->graph = tuning->graph == 1 ? 1 : 0;
    }
    h->in = *in;
    h->mode = mode;
    h->precision = precision;
    h->cat64 = precision != DSM_SGT_FP64;
    h->device = device;
    int rc = DSM_SGT_OK;
    do {
        if (mode != DSM_SGT_SH && mode != DSM_SGT_PSV) {
            rc = fail(h, DSM_SGT_EINVAL, "unknown mode " + std::to_string(mode));
            break;
        }
        if (precision < DSM_SGT_FP64 || precision > DSM_SGT_BF16X6) {
            rc = fail(h, DSM_SGT_EINVAL, "unknown precision " + std::to_string(precision));
            break;
        }
        h->ns = (precision == DSM_SGT_BF16X3) ? 2 : (precision == DSM_SGT_BF16X6) ? 3 : 0;
        if (const std::string bad = check_input(*in); !bad.empty()) {
            rc = fail(h, DSM_SGT_EINVAL, bad);
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
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&h->stream_io, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_pack, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_l0in, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_l0out, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_tab, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_chunk[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_chunk[1], hipEventDisableTiming) != hipSuccess ||
            hipStreamCreateWithFlags(&h->stream_syn, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_ready, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_syn[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_syn[1], hipEventDisableTiming) != hipSuccess) {
            rc = fail(h, DSM_SGT_EHIP, "hipStreamCreate failed");
            break;
        }
        if (const hipError_t e = km_static_twiddles(device, &h->km_static_tw); e != hipSuccess) {
            rc = fail(h, DSM_SGT_EHIP, std::string("static twiddle storage: ") + hipGetErrorString(e));
            break;
        }
        if (h->d_diag.ensure(2 * sizeof(unsigned long long)) != hipSuccess ||
            hipMemset(h->d_diag.p, 0, h->d_diag.n) != hipSuccess) {
            rc = fail(h, DSM_SGT_ENOMEM, "diagnostic counters");
            break;
        }
        rc = (mode == DSM_SGT_SH) ? sh_create(h) : psv_create(h);
        // the cache gives its blocks back when the engine leaves less than
        // the reserve free (kernel scratch is allocated at launch)
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < kHbmReserve) buf_cache().trim(device, kHbmReserve);
    } while (0);
    *handle = h;  // returned even on failure so dsm_sgt_last_error() can report
    return rc;
}

int dsm_sgt_get_tuning(dsm_sgt_t *h, dsm_sgt_tuning_t *t)
{
    if (!h || !t) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    *t = dsm_sgt_tuning_t{};
    t->nslot = h->nslot;
    t->lchunk = h->lchunk;
    t->xcd_group = h->xg;
    t->graph = h->graph;
    t->catalog_window = h->ncat < h->npos ? h->ncat : 0;
    t->overlap = h->ovl ? 1 : 0;
    return DSM_SGT_OK;
}

int dsm_sgt_autotune(const dsm_sgt_input_t *in, int32_t mode, int32_t precision, int32_t device,
                     const char *cache_path, dsm_sgt_tuning_t *best, double *probe_ms)
{
    if (!in || !best) return DSM_SGT_EINVAL;
    if (probe_ms) *probe_ms = 0.0;
    // cache key: FNV-1a of the input record, mode, precision and device name
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DSM_SGT_EHIP;
    uint64_t hv = 1469598103934665603ULL;
    auto mix = [&hv](const void *p, size_t n) {
        for (size_t i = 0; i < n; i++) hv = (hv ^ ((const unsigned char *)p)[i]) * 1099511628211ULL;
    };
    // the record up to the list pointers, then the lists' values (not their addresses)
    mix(in, offsetof(dsm_sgt_input_t, r_list));
    if (in->r_list_n > 0 && in->r_list) mix(in->r_list, sizeof(double) * (size_t)in->r_list_n);
    if (in->theta_list_n > 0 && in->theta_list) mix(in->theta_list, sizeof(double) * (size_t)in->theta_list_n);
    mix(&mode, sizeof(mode));
    mix(&precision, sizeof(precision));
    mix(prop.gcnArchName, std::strlen(prop.gcnArchName));
    // (the key's suffix names the line format: nslot lchunk xcd_group graph
    // overlap, then the probe time; lines of other formats never match)
    char key[32];
    std::snprintf(key, sizeof(key), "%016llx-o", (unsigned long long)hv);
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "r")) {
            char line[256], k[64];
            bool hit = false;
            while (std::fgets(line, sizeof(line), f)) {
                dsm_sgt_tuning_t t{};
                double ms = 0.0;
                if (std::sscanf(line, "%63s %d %d %d %d %d %lf", k, &t.nslot, &t.lchunk, &t.xcd_group, &t.graph,
                                &t.overlap, &ms) == 7 &&
                    std::strcmp(k, key) == 0) {
                    *best = t;  // the last entry of a key wins
                    hit = true;
                }
            }
            std::fclose(f);
            if (hit) return DSM_SGT_OK;
        }
    }
    // candidates: batch size and l chunk need an engine each; the tile
    // grouping and the graph switch are changed on a live engine.  Largest
    // batch first: the smaller engines reuse its released device blocks.
    const bool psv = mode == DSM_SGT_PSV, bf16 = precision == DSM_SGT_BF16X3 || precision == DSM_SGT_BF16X6;
    std::vector<int> nslots, lchunks = {0}, xgs = {0}, graphs = {0};
    if (psv && precision != DSM_SGT_FP64) {
        nslots = {128, 96, 64};
        if (bf16) {
            xgs = {0, 2, 8, -1};
            graphs = {0, 1};
        }
    } else if (psv) {
        nslots = {48, 32, 16};
    } else {
        // (SH's solve and coefficient kernels hold one wave per 64 degrees of
        // a slot: small batches and chunks leave the device under-occupied)
        nslots = {128, 64, 32};
        lchunks = {4096, 2048, 1024};
    }
    // the probe: the whole frequency range when it holds at most 1024
    // frequencies (the plan that wins on the top frequencies alone was
    // measured slower on the whole 5 s catalog); else the top frequencies,
    // a run's short first batch (16) and two full batches of the largest
    // candidate
    const int first = std::max(1, in->imin), ihi = in->imax;
    const int nprobe = (ihi - first + 1 <= 1024) ? ihi - first + 1 : 16 + 2 * nslots[0];
    const int ilo = std::max(first, ihi - nprobe + 1);
    if (ihi < ilo) return DSM_SGT_EINVAL;
    double best_ms = -1.0, tight_ms = -1.0;
    dsm_sgt_tuning_t tight{};
    int last_rc = DSM_SGT_OK;
    std::vector<std::pair<dsm_sgt_tuning_t, double>> tried;
    // headroom a plan must leave: the device's free memory plus the cached
    // blocks of earlier engines, with the plan's engine alive after its
    // runs.  Catalogs produced in windows build one engine after another
    // (several source depths), and plans that left the device nearly full
    // had their blocks freed and allocated afresh between those engines
    // (seconds each); such plans are taken only when no other runs.
    hipDeviceProp_t dp{};
    size_t room = (size_t)64 << 30;
    if (hipGetDeviceProperties(&dp, device) == hipSuccess) room = std::min(room, dp.totalGlobalMem / 5);
    // time plan `t` (tile grouping and graph mode set on the live engine):
    // two runs, the faster counted; a later plan must beat the best by more
    // than 1% (timing noise), and the candidates come simplest first
    auto probe = [&](dsm_sgt_t *h, const dsm_sgt_tuning_t &t) -> int {
        h->xg = t.xcd_group;
        h->graph = t.graph;
        double ms = -1.0;
        for (int rep = 0; rep < 2; rep++) {
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = dsm_sgt_run(h, ilo, ihi);
            if (rc) return rc;
            const double dt = ms_since(t0);
            if (ms < 0.0 || dt < ms) ms = dt;
        }
        size_t fr = 0, tot = 0;
        const bool roomy = hipMemGetInfo(&fr, &tot) != hipSuccess || fr + buf_cache().cached(device) >= room;
        if (!roomy) {
            if (tight_ms < 0.0 || ms < 0.99 * tight_ms) {
                tight_ms = ms;
                tight = t;
                tight.catalog_window = 0;
            }
            return DSM_SGT_OK;
        }
        tried.emplace_back(t, ms);
        if (best_ms < 0.0 || ms < 0.99 * best_ms) {
            best_ms = ms;
            *best = t;
            best->catalog_window = 0;
        }
        return DSM_SGT_OK;
    };
    auto engine = [&](const dsm_sgt_tuning_t &t, dsm_sgt_t **h) -> int {
        dsm_sgt_tuning_t c = t;
        c.catalog_window = ihi - ilo + 1;  // the probe's frequencies only
        const int rc = dsm_sgt_create_tuned(in, mode, precision, device, &c, h);
        if (rc) {  // e.g. a batch that does not fit in device memory
            dsm_sgt_destroy(*h);
            *h = nullptr;
        }
        return rc;
    };
    // stage 1: batch size, l chunk and tile grouping, eager launches
    for (int ns : nslots)
        for (int lc : lchunks) {
            dsm_sgt_tuning_t t{};
            t.nslot = ns;
            t.lchunk = lc;
            dsm_sgt_t *h = nullptr;
            int rc = engine(t, &h);
            for (int xg : xgs) {
                if (rc) break;
                t.xcd_group = xg;
                rc = probe(h, t);
            }
            if (rc) last_rc = rc;
            dsm_sgt_destroy(h);
        }
    // SH: the smallest batch x chunk among the plans within 3% of the
    // fastest.  The solve's scratch grows with both, and a plan that is only
    // as fast on the probe can take most of the device (on the 2.5 s
    // multi-depth grid, 128 x 4096 ran no faster per depth than 64 x 1024
    // and its blocks had to be allocated afresh after other engines).
    if (!psv && !tried.empty()) {
        double fast = tried[0].second;
        for (const auto &x : tried) fast = std::min(fast, x.second);
        const std::pair<dsm_sgt_tuning_t, double> *pick = nullptr;
        auto foot = [](const dsm_sgt_tuning_t &t) { return (double)t.nslot * t.lchunk; };
        for (const auto &x : tried)
            if (x.second <= 1.03 * fast && (!pick || foot(x.first) < foot(pick->first))) pick = &x;
        *best = pick->first;
        best->catalog_window = 0;
        best_ms = pick->second;
    }
    // stage 2: graph mode for the winner
    if (best_ms >= 0.0 && graphs.size() > 1) {
        dsm_sgt_tuning_t t = *best;
        t.graph = 1;
        dsm_sgt_t *h = nullptr;
        int rc = engine(t, &h);
        if (!rc) rc = probe(h, t);
        if (rc) last_rc = rc;
        dsm_sgt_destroy(h);
    }
    // (the overlap plan is not a candidate: it runs the synthesis beside the
    // solves, whose launches it then stretches -- 9.1 -> 13.4 ms on the 5 s
    // grid for a 1.5-2.5% shorter catalog -- and holds a second set of the
    // synthesis buffers; a caller asks for it)
    // the probe engines' blocks, of every candidate's sizes, are freed: left
    // in the cache they split the device between sizes the chosen plan's
    // engines do not all reuse, and engines built one after another then
    // had blocks freed and allocated afresh between them (seconds each)
    buf_cache().flush(device);
    if (best_ms < 0.0 && tight_ms >= 0.0) {  // every plan left the device nearly full
        *best = tight;
        best_ms = tight_ms;
    }
    if (best_ms < 0.0) return last_rc ? last_rc : DSM_SGT_EHIP;
    if (probe_ms) *probe_ms = best_ms;
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "a")) {
            std::fprintf(f, "%s %d %d %d %d %d %.3f\n", key, best->nslot, best->lchunk, best->xcd_group, best->graph,
                         best->overlap, best_ms);
            std::fclose(f);
        }
    }
    return DSM_SGT_OK;
}

int dsm_sgt_get_info(dsm_sgt_t *h, dsm_sgt_info_t *info)
{
    if (!h || !info) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    info->mode = h->mode;
    info->r_n = h->r_n;
    info->theta_n = h->theta_n;
    info->nn = h->nn;
    info->num_tsgt = h->ncomp[0];
    info->num_rsgt = h->ncomp[1];
    info->num_synn = h->ncomp[2];
    info->intir0 = h->intir0;
    info->r0 = h->r0;
    return DSM_SGT_OK;
}

int dsm_sgt_get_grid(dsm_sgt_t *h, double *radii, double *theta_deg)
{
    if (!h) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    if (radii) std::memcpy(radii, h->radii.data(), sizeof(double) * h->r_n);
    if (theta_deg) std::memcpy(theta_deg, h->theta_deg.data(), sizeof(double) * h->theta_n);
    return DSM_SGT_OK;
}

// Run the frequencies `all` (ascending, distinct, no 0) in batches of
// consecutive entries.
static int run_batches(dsm_sgt_t *h, const std::vector<int> &all)
{
    // (the l bound keeps the last batch's cut-off across runs: a run of the
    // next catalog window continues from it; a wrong guess only costs a
    // full-range re-run of the batch)
    if (h->ncat < h->npos) {
        // catalog window: this run's frequencies replace the resident ones
        if ((int)all.size() > h->ncat)
            return fail(h, DSM_SGT_EINVAL, "more frequencies than the catalog window holds");
        std::fill(h->done.begin(), h->done.end(), 0);
        std::fill(h->catpos.begin(), h->catpos.end(), -1);
        for (size_t k = 0; k < all.size(); k++) h->catpos[all[k] - h->ipos0] = (int)k;
    }
    // the reference's rank distribution (SGTpsv-MPI.f90:550) balances
    // ranks; on one device, batches of consecutive frequencies of the set
    // share similar l cut-offs
    std::vector<std::vector<int>> batches;
    {
        // a short first batch when batches are large: its host set-up is the
        // only one not hidden behind device work
        size_t b = 0, first = (h->mode == DSM_SGT_PSV && h->nslot > 32) ? 16 : (size_t)h->nslot;
        while (b < all.size()) {
            const size_t len = batches.empty() ? first : (size_t)h->nslot;
            batches.emplace_back(all.begin() + b, all.begin() + std::min(all.size(), b + len));
            b += len;
        }
    }
    // the host set-up of batch b+1 runs while batch b is on the device
    const int nb = (int)batches.size();
    if (nb == 0) return DSM_SGT_OK;
    if (h->mode == DSM_SGT_SH) {
        auto prepare = [h, &batches](int b) { sh_prepare(h->sh, batches[b], h->shprep[b & 1]); };
        std::future<void> next = std::async(std::launch::async, prepare, 0);
        for (int b = 0; b < nb; b++) {
            const auto tw = std::chrono::steady_clock::now();
            next.get();
            h->stats.prep_wait_ms += ms_since(tw);
            if (b + 1 < nb) next = std::async(std::launch::async, prepare, b + 1);
            const int rc = sh_run_batch(h, h->shprep[b & 1]);
            if (rc) {
                if (b + 1 < nb) next.wait();
                return rc;
            }
        }
        return DSM_SGT_OK;
    }
    auto prepare = [h, &batches](int b) { return psv_prepare(h->psv, batches[b], h->prep[b & 1]); };
    std::future<void> next = std::async(std::launch::async, prepare, 0);
    for (int b = 0; b < nb; b++) {
        const auto tw = std::chrono::steady_clock::now();
        next.get();
        h->stats.prep_wait_ms += ms_since(tw);
        if (b + 1 < nb) next = std::async(std::launch::async, prepare, b + 1);
        if (h->ovl) h->swap_sets();  // the other set of synthesis buffers
        int rc = psv_run_batch(h, h->prep[b & 1], false);
        if (rc == PSV_RETRY_FULL) {
            h->stats.solve_retries++;
            rc = psv_run_batch(h, h->prep[b & 1], true);
        }
        if (rc) {
            if (b + 1 < nb) next.wait();
            return rc;
        }
    }
    return DSM_SGT_OK;
}

// A run of the frequency set `all`: the batches, then the counts of small
// pivots (stats) and of non-finite catalog values, which fail the run with
// DSM_SGT_ENUMERIC (the catalog is kept, as the reference writes it).
static int run_frequencies(dsm_sgt_t *h, const std::vector<int> &all)
{
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    HIPCHK(h, hipMemsetAsync(h->d_diag.p, 0, h->d_diag.n, h->stream));
    const int rc = run_batches(h, all);
    // (overlap plan: the last batches' synthesis is still on its stream)
    const hipError_t es = hipStreamSynchronize(h->stream_syn);
    if (rc) return rc;
    HIPCHK(h, es);
    ev_collect(h);
    unsigned long long diag[2] = {0, 0};
    HIPCHK(h, hipMemcpyAsync(diag, h->d_diag.p, sizeof(diag), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->stats.small_pivots += (int64_t)diag[0];
    h->stats.nonfinite += (int64_t)diag[1];
    if (diag[1])
        return fail(h, DSM_SGT_ENUMERIC,
                    std::to_string(diag[1]) + " non-finite catalog values in frequencies " + std::to_string(all.front()) +
                        ".." + std::to_string(all.back()) + " (a zero or non-finite pivot, or non-finite input)");
    return DSM_SGT_OK;
}

int dsm_sgt_run(dsm_sgt_t *h, int32_t i_begin, int32_t i_end)
{
    if (!h) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    if (i_begin < h->ipos0 || i_end >= h->ipos0 + h->npos || i_begin > i_end)
        return fail(h, DSM_SGT_EINVAL, "frequency range outside [imin, imax] of the input");
    std::vector<int> all;
    for (int i = i_begin; i <= i_end; i++)
        if (i != 0) all.push_back(i);
    return run_frequencies(h, all);
}

int dsm_sgt_run_list(dsm_sgt_t *h, const int32_t *idx, int32_t n)
{
    if (!h) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    if (n < 0 || (n > 0 && !idx)) return fail(h, DSM_SGT_EINVAL, "bad frequency list");
    std::vector<int> all;
    for (int k = 0; k < n; k++) {
        if (idx[k] < h->ipos0 || idx[k] >= h->ipos0 + h->npos)
            return fail(h, DSM_SGT_EINVAL, "frequency " + std::to_string(idx[k]) + " outside [imin, imax] of the input");
        if (idx[k] != 0) all.push_back(idx[k]);
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    return run_frequencies(h, all);
}

static int fetch_common(dsm_sgt_t *h, int32_t i, int32_t kind, int32_t ir, std::vector<double2> &buf, int &num)
{
    if (!h) return DSM_SGT_EINVAL;
    if (i - h->ipos0 < 0 || i - h->ipos0 >= h->npos || !h->done[i - h->ipos0])
        return fail(h, DSM_SGT_ESTATE, "frequency " + std::to_string(i) + " not computed (or not in the catalog window)");
    const int pos = h->catpos[i - h->ipos0];  // device catalog position
    if (pos < 0 || pos >= h->ncat) return fail(h, DSM_SGT_ESTATE, "frequency " + std::to_string(i) + " not resident");
    const int R = h->r_n, T = h->theta_n;
    // device catalogs: [pos][comp][r][theta] (r only for TSGT, RSGT), of
    // complex64 or complex FP64 elements
    const char *src;
    const size_t el = h->cat_el();
    int rstride = R;
    if (kind == DSM_SGT_TSGT) {
        num = h->ncomp[0];
        if (ir < 0 || ir >= R) return fail(h, DSM_SGT_EINVAL, "receiver index out of range");
        src = h->d_cat_t.as<char>() + el * (((size_t)pos * num * R + ir) * T);
    } else if (kind == DSM_SGT_RSGT) {
        num = h->ncomp[1];
        if (!h->in.rsgtswitch) return fail(h, DSM_SGT_ESTATE, "RSGT not computed (rsgtswitch = 0)");
        if (ir < 0 || ir >= R) return fail(h, DSM_SGT_EINVAL, "receiver index out of range");
        src = h->d_cat_r.as<char>() + el * (((size_t)pos * num * R + ir) * T);
    } else if (kind == DSM_SGT_SYNN) {
        num = h->ncomp[2];
        if (!h->in.synnswitch) return fail(h, DSM_SGT_ESTATE, "SYNN not computed (synnswitch = 0)");
        rstride = 1;
        src = h->d_cat_s.as<char>() + el * ((size_t)pos * num * T);
    } else {
        return fail(h, DSM_SGT_EINVAL, "unknown catalog kind");
    }
    std::vector<char> cm(el * num * T);
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    HIPCHK(h, hipMemcpy2D(cm.data(), el * T, src, el * T * rstride, el * T, num, hipMemcpyDeviceToHost));
    // records per distance: [theta][comp]
    buf.resize((size_t)num * T);
    const float2 *c32 = (const float2 *)cm.data();
    const double2 *c64 = (const double2 *)cm.data();
    for (int c = 0; c < num; c++)
        for (int t = 0; t < T; t++) {
            const size_t k = (size_t)c * T + t;
            buf[(size_t)t * num + c] = h->cat64 ? make_double2((double)c32[k].x, (double)c32[k].y) : c64[k];
        }
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

int dsm_sgt_psv_band(dsm_sgt_t *h, int32_t i, double *band)
{
    if (!h || !band) return DSM_SGT_EINVAL;
    if (h->mode != DSM_SGT_PSV || !h->ready) return fail(h, DSM_SGT_ESTATE, "needs a created P-SV handle");
    if (i < 0) return fail(h, DSM_SGT_EINVAL, "negative frequency index");
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    PsvPrep p;
    psv_prepare(h->psv, std::vector<int>{i}, p);
    int rc = psv_device_band(h, p.scal, 1, 0);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(band, h->d_band.p, sizeof(double2) * 16 * (size_t)h->psv.nn, hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return DSM_SGT_OK;
}

int dsm_sgt_lcut(dsm_sgt_t *h, int32_t i, int32_t *llog_minus_1)
{
    if (!h || !llog_minus_1) return DSM_SGT_EINVAL;
    const int pos = i - h->ipos0;
    if (pos < 0 || pos >= h->npos || !h->logged[pos])
        return fail(h, DSM_SGT_ESTATE, "frequency " + std::to_string(i) + " not computed");
    *llog_minus_1 = h->llog[pos];
    return DSM_SGT_OK;
}

// complex64 values of one frequency's records: TSGT, RSGT, SYNN blocks
static size_t record_block(const dsm_sgt_t *h)
{
    return (size_t)h->theta_n * ((size_t)h->r_n * (h->ncomp[0] + h->ncomp[1]) + h->ncomp[2]);
}

// The computed frequencies of [i_begin, i_end] (i = 0 skipped); fails if
// one of them is not resident.
static int resident_list(dsm_sgt_t *h, int32_t i_begin, int32_t i_end, std::vector<int> &freqs)
{
    freqs.clear();
    if (i_begin > i_end || i_begin < h->ipos0 || i_end >= h->ipos0 + h->npos)
        return fail(h, DSM_SGT_EINVAL, "frequency range outside [imin, imax] of the input");
    for (int i = i_begin; i <= i_end; i++) {
        if (i == 0) continue;
        const int p = i - h->ipos0;
        if (!h->done[p] || h->catpos[p] < 0)
            return fail(h, DSM_SGT_ESTATE, "frequency " + std::to_string(i) + " not computed (or not in the catalog window)");
        freqs.push_back(i);
    }
    return DSM_SGT_OK;
}

// Join the writer thread and the copies of the last asynchronous output;
// returns their status.
static int io_join(dsm_sgt_t *h)
{
    if (h->writer.joinable()) h->writer.join();
    if (h->stream_io) {
        const hipError_t e = hipStreamSynchronize(h->stream_io);
        if (e != hipSuccess && !h->io_rc) {
            h->io_rc = DSM_SGT_EHIP;
            h->io_err = std::string("catalog copy: ") + hipGetErrorString(e);
        }
    }
    const int rc = h->io_rc;
    if (rc) h->err = h->io_err;
    h->io_rc = 0;
    h->io_err.clear();
    return rc;
}

// Records of `freqs` (resident) cast and transposed into d_pack on the
// engine stream, block k for freqs[k]; ev_pack marks their completion.
// The caller has joined the previous output (d_pack is free).
static int pack_catalog(dsm_sgt_t *h, const std::vector<int> &freqs)
{
    const size_t blk = record_block(h), n = freqs.size();
    if (!n) return DSM_SGT_OK;
    HIPCHK(h, h->d_pack.ensure(sizeof(float2) * blk * n));
    HIPCHK(h, h->d_packpos.ensure(sizeof(int) * n));
    h->packlist = freqs;
    h->packpos.resize(n);
    for (size_t k = 0; k < n; k++) h->packpos[k] = h->catpos[freqs[k] - h->ipos0];
    HIPCHK(h, hipMemcpyAsync(h->d_packpos.p, h->packpos.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->stream));
    const int R = h->r_n, T = h->theta_n;
    const DevBuf *src[3] = {&h->d_cat_t, &h->d_cat_r, &h->d_cat_s};
    size_t off = 0;
    TimedLaunch t;
    ev_begin(h, t, 3);
    for (int kind = 0; kind < 3; kind++) {
        PackArgs a{};
        a.src = src[kind]->p;
        a.c64 = h->cat64;
        a.ncomp = h->ncomp[kind];
        a.nrec = kind == DSM_SGT_SYNN ? 1 : R;
        if (!a.src) {  // a kind the switches left out: its part of the block is not written
            off += (size_t)a.nrec * T * a.ncomp;
            continue;
        }
        a.theta_n = T;
        a.pos_list = h->d_packpos.as<int>();
        a.dst = h->d_pack.as<float2>();
        a.dst_stride = blk;
        a.dst_off = off;
        launch_pack_records(a, (int)n, h->stream);
        off += (size_t)a.nrec * T * a.ncomp;
    }
    ev_end(h, t);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev_pack, h->stream));
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

static bool write_file(const std::string &path, const float2 *data, size_t n, std::string &err)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    const size_t w = std::fwrite(data, sizeof(float2), n, f);
    const int c = std::fclose(f);
    if (w != n || c != 0) {
        err = "short write " + path;
        return false;
    }
    return true;
}

// The files of frequency i from its packed block `blk` (record_block
// values): TSGT of every receiver, RSGT if rsgtswitch, SYNN if synnswitch
// (tsgtswitch is not read, as in SGTpsv-MPI.f90:1103-1150).
static bool write_block_files(const dsm_sgt_t *h, const std::string &od, const std::string &mn, int i,
                              const float2 *blk, std::string &err)
{
    const char *suf = h->mode == DSM_SGT_SH ? "SH" : "PSV";
    const int R = h->r_n, T = h->theta_n;
    const size_t nt = (size_t)T * h->ncomp[0], nr = (size_t)T * h->ncomp[1], ns = (size_t)T * h->ncomp[2];
    for (int ir = 0; ir < R; ir++) {
        const long rr = (long)(h->radii[ir] * 1.0e3);
        std::string name = od + "/TSGT/" + mn + "." + i7(h->intir0) + "." + i7(rr) + "." + i7(i) + ".TSGT_" + suf;
        if (!write_file(name, blk + ir * nt, nt, err)) return false;
        if (h->in.rsgtswitch) {
            name = od + "/RSGT/" + mn + "." + i7(rr) + "." + i7(i) + ".RSGT_" + suf;
            if (!write_file(name, blk + R * nt + ir * nr, nr, err)) return false;
        }
    }
    if (h->in.synnswitch) {
        const std::string name = od + "/RSGT/" + mn + "." + i7(h->intir0) + "." + i7(i) + ".SYNN_" + suf;
        if (!write_file(name, blk + R * (nt + nr), ns, err)) return false;
    }
    return true;
}

static void make_dirs(const std::string &od)
{
    ::mkdir(od.c_str(), 0755);
    ::mkdir((od + "/TSGT").c_str(), 0755);
    ::mkdir((od + "/RSGT").c_str(), 0755);
    ::mkdir((od + "/log").c_str(), 0755);
}

constexpr int NPIN = 3;  // pinned staging blocks of the writer

// Writer thread: copies the packed blocks to pinned staging on stream_io
// (NPIN - 1 ahead) and writes each frequency's files while the next
// blocks copy.  Touches only d_pack, the staging, stream_io and the
// handle's read-only geometry; its status goes to io_rc / io_err.
static void writer_body(dsm_sgt_t *h, std::string od, std::string mn)
{
    (void)hipSetDevice(h->device);
    const size_t blk = record_block(h);
    const int n = (int)h->packlist.size();
    hipEvent_t ev[NPIN] = {};
    for (auto &e : ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    std::string err;
    int rc = DSM_SGT_OK;
    auto issue = [&](int k) {
        const hipError_t e = hipMemcpyAsync(h->pin + (size_t)(k % NPIN) * blk, h->d_pack.as<float2>() + (size_t)k * blk,
                                            sizeof(float2) * blk, hipMemcpyDeviceToHost, h->stream_io);
        if (e == hipSuccess) (void)hipEventRecord(ev[k % NPIN], h->stream_io);
        return e;
    };
    for (int k = 0; k < std::min(NPIN - 1, n) && !rc; k++)
        if (issue(k) != hipSuccess) rc = DSM_SGT_EHIP, err = "catalog copy failed";
    for (int k = 0; k < n && !rc; k++) {
        if (k + NPIN - 1 < n && issue(k + NPIN - 1) != hipSuccess) {
            rc = DSM_SGT_EHIP, err = "catalog copy failed";
            break;
        }
        if (hipEventSynchronize(ev[k % NPIN]) != hipSuccess) {
            rc = DSM_SGT_EHIP, err = "catalog copy failed";
            break;
        }
        if (!write_block_files(h, od, mn, h->packlist[k], h->pin + (size_t)(k % NPIN) * blk, err)) rc = DSM_SGT_EIO;
    }
    (void)hipStreamSynchronize(h->stream_io);
    for (auto &e : ev) (void)hipEventDestroy(e);
    h->io_rc = rc;
    h->io_err = err;
}

int dsm_sgt_write_catalogs_async(dsm_sgt_t *h, int32_t i_begin, int32_t i_end, const char *output_dir,
                                 const char *model_name)
{
    if (!h || !output_dir || !model_name) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    int rc = io_join(h);
    if (rc) return rc;
    std::vector<int> freqs;
    rc = resident_list(h, i_begin, i_end, freqs);
    if (rc || freqs.empty()) return rc;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    const size_t blk = record_block(h);
    if (h->pin_n < NPIN * blk) {
        if (h->pin) (void)hipHostFree(h->pin);
        h->pin = nullptr;
        h->pin_n = 0;
        HIPCHK(h, hipHostMalloc((void **)&h->pin, sizeof(float2) * NPIN * blk, hipHostMallocDefault));
        h->pin_n = NPIN * blk;
    }
    make_dirs(output_dir);
    rc = pack_catalog(h, freqs);
    if (rc) return rc;
    HIPCHK(h, hipStreamWaitEvent(h->stream_io, h->ev_pack, 0));
    try {
        h->writer = std::thread(writer_body, h, std::string(output_dir), std::string(model_name));
    } catch (...) {
        return fail(h, DSM_SGT_ENOMEM, "cannot start the catalog writer thread");
    }
    return DSM_SGT_OK;
}

int dsm_sgt_copy_catalog_async(dsm_sgt_t *h, int32_t i_begin, int32_t i_end, void *dst, size_t dst_bytes)
{
    if (!h || !dst) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    int rc = io_join(h);
    if (rc) return rc;
    std::vector<int> freqs;
    rc = resident_list(h, i_begin, i_end, freqs);
    if (rc || freqs.empty()) return rc;
    const size_t bytes = sizeof(float2) * record_block(h) * freqs.size();
    if (dst_bytes < bytes)
        return fail(h, DSM_SGT_EINVAL, "destination holds " + std::to_string(dst_bytes) + " bytes, the records need " +
                                           std::to_string(bytes));
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    rc = pack_catalog(h, freqs);
    if (rc) return rc;
    HIPCHK(h, hipStreamWaitEvent(h->stream_io, h->ev_pack, 0));
    HIPCHK(h, hipMemcpyAsync(dst, h->d_pack.p, bytes, hipMemcpyDeviceToHost, h->stream_io));
    return DSM_SGT_OK;
}

int dsm_sgt_wait(dsm_sgt_t *h)
{
    if (!h) return DSM_SGT_EINVAL;
    return io_join(h);
}

int dsm_sgt_host_alloc(size_t bytes, void **ptr)
{
    if (!ptr) return DSM_SGT_EINVAL;
    *ptr = nullptr;
    if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        return DSM_SGT_ENOMEM;
    }
    return DSM_SGT_OK;
}

void dsm_sgt_host_free(void *ptr)
{
    if (ptr) (void)hipHostFree(ptr);
}

int dsm_sgt_release_cache(int32_t device)
{
    if (device < 0) return DSM_SGT_EINVAL;
    buf_cache().flush(device);
    return DSM_SGT_OK;
}

int dsm_sgt_write_catalog(dsm_sgt_t *h, int32_t i, const char *output_dir, const char *model_name)
{
    if (!h || !output_dir || !model_name) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    int rc = io_join(h);
    if (rc) return rc;
    std::vector<int> freqs;
    rc = resident_list(h, i, i, freqs);
    if (rc || freqs.empty()) return rc;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    rc = pack_catalog(h, freqs);
    if (rc) return rc;
    std::vector<float2> blk(record_block(h));
    HIPCHK(h, hipMemcpyAsync(blk.data(), h->d_pack.p, sizeof(float2) * blk.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    make_dirs(output_dir);
    std::string err;
    if (!write_block_files(h, output_dir, model_name, i, blk.data(), err)) return fail(h, DSM_SGT_EIO, err);
    return DSM_SGT_OK;
}

// ------------------------------------------------------------------------
// KernelMaker catalog stage
// ------------------------------------------------------------------------
// The reference's exp(cmplx(0.d0, x)): CMPLX of default kind, so x rounded
// to float and the factor's cos and sin taken in float, as the C library's
// cosf / sinf give them (flang's runtime calls the same functions).
static float2 km_phase(double x)
{
    const float xf = (float)x;
    return make_float2(std::cos(xf), std::sin(xf));
}

// catalog position of frequency i in engine e, -1 if i = 0; -2 if not resident
static int km_pos(const dsm_sgt_t *e, int i)
{
    if (i == 0) return -1;
    const int p = i - e->ipos0;
    if (p < 0 || p >= e->npos || !e->done[p] || e->catpos[p] < 0) return -2;
    return e->catpos[p];
}

// fn(k) for k in [0, n) split over host threads
extern "C++" template <class F> inline void km_host_parallel(long n, F fn)
{
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        for (long k = 0; k < n; k++) fn(k);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (long k = (long)t * n / nt; k < (long)(t + 1) * n / nt; k++) fn(k);
        });
    for (auto &x : th) x.join();
}

int dsm_km_points(dsm_sgt_t *h, dsm_sgt_t *sh, int32_t ir, const dsm_km_params_t *prm, int32_t npoints,
                  const double *deltar, const double *deltas, const double *trig, void *h3, void *h4)
{
    if (!h || !prm) return DSM_SGT_EINVAL;
    if (!h->ready || h->mode != DSM_SGT_PSV) return fail(h, DSM_SGT_ESTATE, "needs a created P-SV engine");
    if (npoints < 0 || (npoints > 0 && (!deltar || !deltas || !trig || !h3 || !h4)))
        return fail(h, DSM_SGT_EINVAL, "bad point arrays");
    if (ir < 0 || ir >= h->r_n) return fail(h, DSM_SGT_EINVAL, "receiver index out of range");
    if (!h->in.rsgtswitch || (sh && !sh->in.rsgtswitch))
        return fail(h, DSM_SGT_ESTATE, "the stage reads RSGT records: the engines need rsgtswitch = 1");
    if (prm->compo < 0 || prm->compo > 2) return fail(h, DSM_SGT_EINVAL, "compo must be 0 (Z), 1 (R) or 2 (T)");
    if (!(prm->ipdistance > 0.0)) return fail(h, DSM_SGT_EINVAL, "ipdistance must be positive");
    if (prm->fmin < 0 || prm->fmax < prm->fmin) return fail(h, DSM_SGT_EINVAL, "bad frequency range");
    if (h->in.theta_list_n > 0 || (sh && sh->in.theta_list_n > 0))
        return fail(h, DSM_SGT_EINVAL, "the sinc interpolation in distance needs the uniform distance grid, not a list");
    const int nf = prm->fmax - prm->fmin + 1, T = h->theta_n;
    std::vector<int> pos(nf), pos_sh(nf, -1);
    for (int j = 0; j < nf; j++) {
        pos[j] = km_pos(h, prm->fmin + j);
        if (pos[j] == -2)
            return fail(h, DSM_SGT_ESTATE, "frequency " + std::to_string(prm->fmin + j) + " not resident in the P-SV engine");
    }
    // the SH engine's records of the same radius, if it holds that radius
    int ir_sh = -1;
    if (sh) {
        if (!sh->ready || sh->mode != DSM_SGT_SH) return fail(h, DSM_SGT_EINVAL, "the second engine must be an SH engine");
        if (sh->theta_n != T || sh->in.tlen != h->in.tlen || sh->in.thetamin != h->in.thetamin ||
            sh->in.thetadelta != h->in.thetadelta)
            return fail(h, DSM_SGT_EINVAL, "the SH engine's distances or tlen differ from the P-SV engine's");
        for (int k = 0; k < sh->r_n; k++)
            if (sh->radii[k] == h->radii[ir]) ir_sh = k;
        if (ir_sh >= 0)
            for (int j = 0; j < nf; j++) {
                pos_sh[j] = km_pos(sh, prm->fmin + j);
                if (pos_sh[j] == -2)
                    return fail(h, DSM_SGT_ESTATE,
                                "frequency " + std::to_string(prm->fmin + j) + " not resident in the SH engine");
            }
    }
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    if (sh && ir_sh >= 0 && sh->device != h->device) return fail(h, DSM_SGT_EINVAL, "engines on different devices");
    const double tlen = h->in.tlen, cred = prm->c_red_reci;
    constexpr double PI_KM = 3.1415926535897932;  // KernelMaker's module parameters
    // phase tables: rdsgtomega's per (frequency, distance), clsgt's per
    // (point, frequency) for deltar and deltas
    std::vector<float2> ph((size_t)nf * T), php((size_t)2 * npoints * nf);
    km_host_parallel((long)nf * T, [&](long k) {
        const int j = (int)(k / T), t = (int)(k % T);
        const double redtime = -h->theta_deg[t] * cred;
        ph[k] = km_phase(-2.0 * PI_KM * (double)(prm->fmin + j) / tlen * redtime);
    });
    km_host_parallel((long)npoints * nf, [&](long k) {
        const int p = (int)(k / nf), j = (int)(k % nf);
        const double f = -2.0 * PI_KM * (double)(prm->fmin + j) / tlen;
        php[k] = km_phase(f * deltar[p] * cred);
        php[(size_t)npoints * nf + k] = km_phase(f * deltas[p] * cred);
    });
    std::vector<int> posall(pos);
    posall.insert(posall.end(), pos_sh.begin(), pos_sh.end());
    HIPCHK(h, h->d_km_omt.ensure(sizeof(double2) * (size_t)nf * T * 20));
    HIPCHK(h, h->d_km_omr.ensure(sizeof(double2) * (size_t)nf * T * 10));
    HIPCHK(h, h->d_km_ph.ensure(sizeof(float2) * ph.size()));
    HIPCHK(h, h->d_km_pos.ensure(sizeof(int) * posall.size()));
    HIPCHK(h, h->d_km_thd.ensure(sizeof(double) * T));
    HIPCHK(h, h->d_km_pts.ensure(sizeof(double) * std::max<size_t>(1, (size_t)npoints * 14)));
    HIPCHK(h, h->d_km_php.ensure(sizeof(float2) * std::max<size_t>(1, php.size())));
    const size_t nout = (size_t)npoints * nf * 6;
    HIPCHK(h, h->d_km_out.ensure(sizeof(double2) * std::max<size_t>(1, 2 * nout)));
    HIPCHK(h, hipMemcpyAsync(h->d_km_ph.p, ph.data(), sizeof(float2) * ph.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_km_pos.p, posall.data(), sizeof(int) * posall.size(), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_km_thd.p, h->theta_deg.data(), sizeof(double) * T, hipMemcpyHostToDevice,
                             h->stream));
    if (npoints > 0) {
        double *pts = h->d_km_pts.as<double>();
        HIPCHK(h, hipMemcpyAsync(pts, deltar, sizeof(double) * npoints, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(pts + npoints, deltas, sizeof(double) * npoints, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(pts + 2 * (size_t)npoints, trig, sizeof(double) * 12 * npoints,
                                 hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->d_km_php.p, php.data(), sizeof(float2) * php.size(), hipMemcpyHostToDevice,
                                 h->stream));
    }
    KmRdArgs ra{};
    ra.cat_t_psv = h->d_cat_t.p;
    ra.cat_r_psv = h->d_cat_r.p;
    ra.c64_psv = h->cat64;
    ra.r_psv = h->r_n;
    ra.ir_psv = ir;
    if (sh && ir_sh >= 0) {
        ra.cat_t_sh = sh->d_cat_t.p;
        ra.cat_r_sh = sh->d_cat_r.p;
        ra.c64_sh = sh->cat64;
        ra.r_sh = sh->r_n;
        ra.ir_sh = ir_sh;
    }
    ra.pos_psv = h->d_km_pos.as<int>();
    ra.pos_sh = h->d_km_pos.as<int>() + nf;
    ra.ph = h->d_km_ph.as<float2>();
    ra.nf = nf;
    ra.theta_n = T;
    ra.om_t = h->d_km_omt.as<double2>();
    ra.om_r = h->d_km_omr.as<double2>();
    // the SH engine's last run must be complete before its catalog is read
    if (sh && ir_sh >= 0) HIPCHK(sh, hipStreamSynchronize(sh->stream));
    TimedLaunch t;
    ev_begin(h, t, 3);
    launch_km_rd(ra, h->stream);
    HIPCHK(h, hipGetLastError());
    KmPointArgs pa{};
    pa.om_t = ra.om_t;
    pa.om_r = ra.om_r;
    pa.thetaD = h->d_km_thd.as<double>();
    pa.theta_n = T;
    pa.nf = nf;
    pa.npoints = npoints;
    pa.compo = prm->compo;
    pa.thetadelta = h->in.thetadelta;
    pa.ipdistance = prm->ipdistance;
    pa.deltar = h->d_km_pts.as<double>();
    pa.deltas = h->d_km_pts.as<double>() + npoints;
    pa.trig = h->d_km_pts.as<double>() + 2 * (size_t)npoints;
    pa.ph_r = h->d_km_php.as<float2>();
    pa.ph_s = h->d_km_php.as<float2>() + (size_t)npoints * nf;
    for (int k = 0; k < 6; k++) pa.mt[k] = prm->mt[k];
    pa.h3 = h->d_km_out.as<double2>();
    pa.h4 = h->d_km_out.as<double2>() + nout;
    launch_km_points(pa, h->stream);
    HIPCHK(h, hipGetLastError());
    ev_end(h, t);
    if (npoints > 0) {
        HIPCHK(h, hipMemcpyAsync(h3, pa.h3, sizeof(double2) * nout, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(h4, pa.h4, sizeof(double2) * nout, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ev_collect(h);
    return DSM_SGT_OK;
}

int dsm_km_vectorfft(dsm_sgt_t *h, int32_t nvec, int32_t imin, int32_t imax, int32_t np1, const void *ccvec,
                     double omegai, double tlen, int32_t iws, int32_t iwe, double *rvec)
{
    if (!h) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    if (np1 < 2 || np1 > KM_FFT_MAX_NP1 || (np1 & (np1 - 1)))
        return fail(h, DSM_SGT_EINVAL, "np1 must be a power of two in 2.." + std::to_string(KM_FFT_MAX_NP1));
    const int N = 2 * np1;
    if (imin < 0 || imax < imin || imax >= N || iws < 0 || iwe < iws || iwe >= N)
        return fail(h, DSM_SGT_EINVAL, "spectrum or window outside 0..2 np1 - 1");
    if (nvec < 0 || (nvec > 0 && (!ccvec || !rvec))) return fail(h, DSM_SGT_EINVAL, "bad arrays");
    if (!(tlen > 0.0)) return fail(h, DSM_SGT_EINVAL, "tlen must be positive");
    if (nvec == 0) return DSM_SGT_OK;
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    const size_t nin = (size_t)nvec * (imax - imin + 1), nout = (size_t)nvec * (iwe - iws + 1);
    HIPCHK(h, h->d_km_fin.ensure(sizeof(double2) * nin));
    HIPCHK(h, h->d_km_fout.ensure(sizeof(double) * nout));
    if (h->km_tw_n != N) {  // e^{2 pi i k / N}, k < N / 2
        HIPCHK(h, h->d_km_tw.ensure(sizeof(double2) * (N / 2)));
        launch_km_twiddles(h->d_km_tw.as<double2>(), N, h->stream);
        HIPCHK(h, hipGetLastError());
        h->km_tw_n = N;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_km_fin.p, ccvec, sizeof(double2) * nin, hipMemcpyHostToDevice, h->stream));
    KmFftArgs fa{};
    fa.ccvec = h->d_km_fin.as<double2>();
    fa.imin = imin;
    fa.imax = imax;
    fa.np1 = np1;
    fa.iws = iws;
    fa.iwe = iwe;
    fa.omegai = omegai;
    fa.tlen = tlen;
    fa.tw2 = h->d_km_tw.as<double2>();
    fa.rvec = h->d_km_fout.as<double>();
    TimedLaunch t;
    ev_begin(h, t, 3);
    launch_km_fft(fa, nvec, h->stream);
    HIPCHK(h, hipGetLastError());
    ev_end(h, t);
    HIPCHK(h, hipMemcpyAsync(rvec, fa.rvec, sizeof(double) * nout, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    ev_collect(h);
    return DSM_SGT_OK;
}

// A device array of `bytes` bytes at p on `device`: within one allocation
// of that device (not host memory, not running past the allocation's end).
static bool device_span(const void *p, size_t bytes, int device, std::string &why)
{
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != device) {
        (void)hipGetLastError();
        why = "is not device memory of the engine's GPU";
        return false;
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        why = "is not device memory of the engine's GPU";
        return false;
    }
    const uintptr_t b = (uintptr_t)base, q = (uintptr_t)p;
    if (q < b || q - b > size || bytes > size - (q - b)) {
        why = "extends past the end of its device allocation";
        return false;
    }
    return true;
}

int dsm_km_vectorfft_noalloc(dsm_sgt_t *h, void *stream, int32_t nvec, int32_t imin, int32_t imax, int32_t np1,
                             const void *ccvec, double omegai, double tlen, int32_t iws, int32_t iwe, double *rvec)
{
    if (!h) return DSM_SGT_EINVAL;
    if (!h->ready) return fail(h, DSM_SGT_ESTATE, "engine not initialised");
    if (np1 > KM_FFT_MAX_NP1)
        return fail(h, DSM_SGT_EINVAL, "np1 above " + std::to_string(KM_FFT_MAX_NP1) +
                                           ", the size of the static twiddle storage");
    if (np1 < 2 || (np1 & (np1 - 1))) return fail(h, DSM_SGT_EINVAL, "np1 must be a power of two");
    const int N = 2 * np1;
    if (imin < 0 || imax < imin || imax >= N || iws < 0 || iwe < iws || iwe >= N)
        return fail(h, DSM_SGT_EINVAL, "spectrum or window outside 0..2 np1 - 1");
    if (nvec < 0) return fail(h, DSM_SGT_EINVAL, "bad arrays");
    if (!(tlen > 0.0)) return fail(h, DSM_SGT_EINVAL, "tlen must be positive");
    if (nvec == 0) return DSM_SGT_OK;
    if (!h->km_static_tw) return fail(h, DSM_SGT_ESTATE, "no static twiddle storage on the engine's device");
    if (hipSetDevice(h->device) != hipSuccess) return fail(h, DSM_SGT_EHIP, "hipSetDevice failed");
    const size_t nin = (size_t)nvec * (imax - imin + 1), nout = (size_t)nvec * (iwe - iws + 1);
    std::string why;
    if (!ccvec || !device_span(ccvec, sizeof(double2) * nin, h->device, why))
        return fail(h, DSM_SGT_EINVAL, "ccvec " + (ccvec ? why : std::string("is NULL")));
    if (!rvec || !device_span(rvec, sizeof(double) * nout, h->device, why))
        return fail(h, DSM_SGT_EINVAL, "rvec " + (rvec ? why : std::string("is NULL")));
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    if (stream) {  // the caller's stream must belong to the engine's GPU
        hipDevice_t sd = -1;
        if (hipStreamGetDevice(s, &sd) != hipSuccess || sd != h->device) {
            (void)hipGetLastError();
            return fail(h, DSM_SGT_EINVAL, "stream is not a stream of the engine's GPU");
        }
    }
    launch_km_twiddles(h->km_static_tw, N, s);
    KmFftArgs fa{};
    fa.ccvec = (const double2 *)ccvec;
    fa.imin = imin;
    fa.imax = imax;
    fa.np1 = np1;
    fa.iws = iws;
    fa.iwe = iwe;
    fa.omegai = omegai;
    fa.tlen = tlen;
    fa.tw2 = h->km_static_tw;
    fa.rvec = rvec;
    launch_km_fft(fa, nvec, s);
    HIPCHK(h, hipGetLastError());
    if (!stream) HIPCHK(h, hipStreamSynchronize(h->stream));
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
    std::fill(h->logged.begin(), h->logged.end(), 0);
    return DSM_SGT_OK;
}

const char *dsm_sgt_last_error(dsm_sgt_t *h) { return h ? h->err.c_str() : "null handle"; }

void dsm_sgt_destroy(dsm_sgt_t *h)
{
    if (!h) return;
    (void)io_join(h);
    if (h->stream_io) (void)hipStreamDestroy(h->stream_io);
    if (h->ev_pack) (void)hipEventDestroy(h->ev_pack);
    if (h->pin) (void)hipHostFree(h->pin);
    if (h->stream) {
        (void)hipSetDevice(h->device);
        (void)hipStreamSynchronize(h->stream);
        if (h->stream2) (void)hipStreamSynchronize(h->stream2);
        free_graphs(h);
        ev_collect(h);
        (void)hipStreamDestroy(h->stream);
    }
    if (h->stream2) {
        (void)hipStreamSynchronize(h->stream2);
        (void)hipStreamDestroy(h->stream2);
    }
    if (h->stream_syn) {
        (void)hipStreamSynchronize(h->stream_syn);
        (void)hipStreamDestroy(h->stream_syn);
    }
    for (hipEvent_t e : {h->ev_ready, h->ev_syn[0], h->ev_syn[1]})
        if (e) (void)hipEventDestroy(e);
    if (h->pin_amp) (void)hipHostFree(h->pin_amp);  // (after the streams: copies into it are stream work)
    if (h->ev_tab) (void)hipEventDestroy(h->ev_tab);
    for (hipEvent_t e : h->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    if (h->ev_l0in) (void)hipEventDestroy(h->ev_l0in);
    if (h->ev_l0out) {
        (void)hipEventDestroy(h->ev_l0out);
    }
    delete h;
}

}  // extern "C"
