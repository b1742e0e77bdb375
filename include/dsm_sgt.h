/*
 * dsm_sgt.h -- C-ABI of the MI355X strain-Green's-tensor (SGT) engine.
 *
 * Drop-in for the per-frequency loop body of IPGP/DSM-Kernel's SGTsh and
 * SGTpsv drivers.  The reference has no plugin/operator API for this path:
 * the work is inlined in the driver loops
 *     src/SGTsh/SGTsh-MPI.f90:389-613     (SH, one frequency per iteration)
 *     src/SGTpsv/SGTpsv-MPI.f90:535-1153  (P-SV, one frequency per iteration)
 * and the host side around them (input parsing, grid, element integrals)
 * is the set-up code SGTsh-MPI.f90:27-348 / SGTpsv-MPI.f90:37-477.  These
 * entry points replace that loop body at frequency-batch granularity; the
 * caller keeps the reference's stdin grammar, files and catalog layout.
 *
 * All arrays are plain pointers; complex numbers are interleaved
 * (re, im) pairs, i.e. Fortran complex(4)/complex(8), C float/double
 * _Complex.  No exceptions or process exits cross this boundary: every
 * call returns a status (0 = ok, see DSM_SGT_E*) and
 * dsm_sgt_last_error() describes the last failure.  A handle is used by
 * one host thread at a time.
 */
#ifndef DSM_SGT_H
#define DSM_SGT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSM_SGT_MAXZONE 64

/* Parsed inputs.  Fields mirror what the reference reads:
 *   parameter file (stdin)  src/SGTsh/others.f90:1-60  (pinput)
 *   DSM.config              src/SGTsh/others.f90:62-92 (readDSMconf)
 *   model file              src/SGTpsv/SGTpsv-MPI.f90:100-105 (28 values per zone)
 * Polynomials are stored coefficient-fastest like the reference's
 * rrho(4,nzone); omegai_seed is the DSM.config value, converted by the
 * engine as -log(seed)/tlen (SGTpsv-MPI.f90:125). */
typedef struct dsm_sgt_input {
    int32_t nzone;
    int32_t maxlmax;
    int32_t imin, imax;
    int32_t rsgtswitch, tsgtswitch, synnswitch, psgtswitch;
    double vrmin[DSM_SGT_MAXZONE], vrmax[DSM_SGT_MAXZONE];
    double rho[DSM_SGT_MAXZONE][4], vpv[DSM_SGT_MAXZONE][4], vph[DSM_SGT_MAXZONE][4];
    double vsv[DSM_SGT_MAXZONE][4], vsh[DSM_SGT_MAXZONE][4], eta[DSM_SGT_MAXZONE][4];
    double qmu[DSM_SGT_MAXZONE], qkappa[DSM_SGT_MAXZONE];
    double re, ratc, ratl, omegai_seed;
    double tlen;
    double rmin_, rmax_, rdelta_;
    double r0min;
    double thetamin, thetamax, thetadelta;
    /* Receiver geometry as lists instead of the uniform grids above, as
     * DSM3D_forward reads it (src/DSM3D_forward/mpi-tipsv.f90:163-210: a
     * file of receiver radii and a file of distances).  r_list_n > 0: the
     * receiver radii are r_list[0..r_list_n) (km, ascending), and rmin_,
     * rmax_, rdelta_ are not used; theta_list_n > 0: the distances are
     * theta_list[0..theta_list_n) (degrees, ascending, within [0, 180]), and
     * thetamin, thetamax, thetadelta are not used.  Zero counts (and null
     * pointers) keep the grids.  The arrays are read by dsm_sgt_create only. */
    int32_t r_list_n, theta_list_n;
    const double *r_list;
    const double *theta_list;
} dsm_sgt_input_t;

enum { DSM_SGT_SH = 1, DSM_SGT_PSV = 2 };
/* Arithmetic of the synthesis contraction.  FP64 is the parity form; the
 * mixed forms are FP32 (FP32 MFMA), BF16X3 (operands split into two bf16
 * terms, three products) and BF16X6 (three terms, six products), all with
 * FP32 MFMA accumulation over 32 degrees added into FP64.  The band solve
 * and the Legendre table stay FP64 in all forms. */
enum { DSM_SGT_FP64 = 0, DSM_SGT_FP32 = 1, DSM_SGT_BF16X3 = 2, DSM_SGT_BF16X6 = 3 };
/* Catalog kinds (file suffixes TSGT_*, RSGT_*, SYNN_*). */
enum { DSM_SGT_TSGT = 0, DSM_SGT_RSGT = 1, DSM_SGT_SYNN = 2 };

enum {
    DSM_SGT_OK = 0,
    DSM_SGT_EINVAL = 1,    /* bad argument or input (e.g. source outside the model) */
    DSM_SGT_ESTATE = 2,    /* frequency not computed, or handle misuse */
    DSM_SGT_EHIP = 3,      /* HIP runtime error (no device, launch failure) */
    DSM_SGT_ENOMEM = 4,    /* device or host allocation failed */
    DSM_SGT_EIO = 5,       /* catalog file could not be written */
    DSM_SGT_EUNSUPPORTED = 6,
    DSM_SGT_ENUMERIC = 7   /* the run finished but the catalog holds non-finite values
                              (stats.nonfinite); the catalog is kept */
};

typedef struct dsm_sgt dsm_sgt_t;

/* Sizes of a run after the reference's own set-up (receiver trimming for
 * SH at SGTsh-MPI.f90:152-162, source clamping in calspo). */
typedef struct dsm_sgt_info {
    int32_t mode;
    int32_t r_n;        /* receiver radii */
    int32_t theta_n;    /* epicentral distances */
    int32_t nn;         /* band rows of the (omega,l) systems */
    int32_t num_tsgt, num_rsgt, num_synn;  /* components per catalog record */
    int32_t intir0;     /* int(r0*1e3) before calspo clamps r0 (file names) */
    double r0;          /* source radius after calspo */
} dsm_sgt_info_t;

/* Per-stage device time and work since the last reset (HIP events on the
 * engine's stream). */
typedef struct dsm_sgt_stats {
    double solve_ms, legendre_ms, synth_ms, other_ms;
    int64_t solve_launches, synth_launches;
    int64_t systems;          /* (omega,l) systems factored */
    double solve_bytes;       /* algorithmic bytes of the band solves */
    double synth_flops;       /* algorithmic flops of the synthesis contraction */
    double synth_bytes;       /* operand bytes streamed by the synthesis */
    double synth_mfma_flops;  /* MFMA flops the synthesis kernels issued (bf16 forms: all cross products) */
    double prep_wait_ms;      /* host time the run waited for a batch's host set-up */
    double device_wait_ms;    /* host time the run blocked on the device inside batches */
    int64_t solve_retries;    /* batched P-SV forms: batches run again with the full l range because a
                                 cut-off fell past the batch's l bound */
    double solve_alg_bytes;   /* algorithmic bytes of the reference's (omega, l >= 1) systems below the
                                 cut-off: 16 (w n + 4 R r_n) each, w = 4 / 2 band entries per row, n the rows
                                 solved (after the depth cut), R = 18 / 14 right-hand sides */
    int64_t small_pivots;     /* SH: pivots with |d| <= 1e-14 in the band factorizations, the test of
                                 dclisb0 (src/SGTsh/dclib.f90:37-86), which reports and continues; P-SV's
                                 dcsymbdl0 receives eps = 0 from its driver and tests nothing */
    int64_t nonfinite;        /* non-finite values the synthesis stored into the catalogs */
} dsm_sgt_stats_t;

/* Create an engine on HIP device `device` for `mode` and `precision`.
 * Runs the omega-independent set-up (grid, element integrals, Legendre
 * table).  Replaces SGTsh-MPI.f90:146-348 / SGTpsv-MPI.f90:123-477. */
int dsm_sgt_create(const dsm_sgt_input_t *in, int32_t mode, int32_t precision, int32_t device,
                   dsm_sgt_t **handle);

/* Launch plan of the batched pipeline.  No counterpart in the reference,
 * whose MPI driver gives one frequency at a time to each rank
 * (SGTpsv-MPI.f90:550).  A field left 0 takes the built-in default. */
typedef struct dsm_sgt_tuning {
    int32_t nslot;      /* frequencies per batch: a solve launch factors nslot x (l range) systems */
    int32_t lchunk;     /* degrees per solve launch between cut-off checks (SH, FP64 P-SV);
                           the batched P-SV forms always take the whole l range */
    int32_t xcd_group;  /* BF16 synthesis: tile groups dealt to one XCD at a time; -1: no tile list
                           (the kernel's own XCD-sliced order) */
    int32_t graph;      /* 1 (batched P-SV forms): each batch's launches after the factor-storage
                           read-back run as two HIP graphs, the solve sequence and the synthesis
                           pair; 0: eager launches */
    int32_t catalog_window; /* > 0: the device catalog holds this many frequencies, the frequencies of
                               the last dsm_sgt_run only (at most this many), so that catalogs larger
                               than device memory are produced run by run; 0: all of [imin, imax] */
    int32_t overlap;    /* 1 (BF16 P-SV forms, eager launches): the synthesis of batch b runs on a
                           stream of its own beside the solves of batch b+1, with the buffers between
                           them held twice (about 47 GB more for 128 frequencies of the 5 s grid);
                           0: one stream.  Not chosen by dsm_sgt_autotune */
    int32_t reserved[2];
} dsm_sgt_tuning_t;

/* dsm_sgt_create with a launch plan (NULL: the defaults). */
int dsm_sgt_create_tuned(const dsm_sgt_input_t *in, int32_t mode, int32_t precision, int32_t device,
                         const dsm_sgt_tuning_t *tuning, dsm_sgt_t **handle);
/* The launch plan an engine runs with (defaults filled in). */
int dsm_sgt_get_tuning(dsm_sgt_t *h, dsm_sgt_tuning_t *tuning);
/* Autotuner: times the candidate launch plans for this input, mode and
 * precision on a probe run (all of [imin, imax] when it holds at most 1024
 * frequencies, else its top 16 + twice the largest candidate batch), each
 * plan run twice, the faster run counted, and returns the fastest plan in
 * `best` and its probe time in `probe_ms`.  Candidates: the batch size,
 * the l chunk (SH, FP64 P-SV) and the synthesis tile grouping (BF16
 * forms); then graph mode for the winner (batched P-SV forms).  The overlap
 * plan is not a candidate (a caller sets it).  With `cache_path` (may be NULL) the result is looked up in and
 * appended to that text file, keyed by a hash of the input, mode,
 * precision and device name; a cached plan returns with probe_ms = 0.
 * The probe engines' catalogs are discarded and their device memory freed
 * (the first engine built afterwards allocates its own). */
int dsm_sgt_autotune(const dsm_sgt_input_t *in, int32_t mode, int32_t precision, int32_t device,
                     const char *cache_path, dsm_sgt_tuning_t *best, double *probe_ms);

int dsm_sgt_get_info(dsm_sgt_t *h, dsm_sgt_info_t *info);
/* Receiver radii (r_n values) and distances in degrees (theta_n values). */
int dsm_sgt_get_grid(dsm_sgt_t *h, double *radii, double *theta_deg);

/* Compute every frequency index i in [i_begin, i_end] except i = 0
 * (the reference never computes i = 0, SGTpsv-MPI.f90:550).  Results stay
 * in device memory until fetched.  Replaces the body of the frequency
 * loop, SGTsh-MPI.f90:389-560 / SGTpsv-MPI.f90:535-1095. */
int dsm_sgt_run(dsm_sgt_t *h, int32_t i_begin, int32_t i_end);

/* Compute the frequencies idx[0..n-1] (any order; duplicates and i = 0
 * are dropped), batched as a range is: ascending, consecutive entries of
 * the set in one batch.  For one MPI rank's frequency set of the
 * reference's zig-zag distribution (SGTpsv-MPI.f90:550, SGTsh-MPI.f90:389)
 * or any other subset; replaces the same loop body as dsm_sgt_run.  With
 * a catalog window the set must fit in the window. */
int dsm_sgt_run_list(dsm_sgt_t *h, const int32_t *idx, int32_t n);

/* Copy one catalog record: kind (DSM_SGT_TSGT/RSGT/SYNN), frequency i and
 * receiver index ir (0-based; ignored for SYNN) into `out`, a Fortran
 * array complex(4) out(num, theta_n), exactly the record the reference
 * writes (SGTsh-MPI.f90:577-611, SGTpsv-MPI.f90:1113-1149). */
int dsm_sgt_fetch(dsm_sgt_t *h, int32_t i, int32_t kind, int32_t ir, void *out);
/* Same record as complex(8) out(num, theta_n): the plain FP64 form's
 * accumulators before the complex64 cast; the mixed forms keep their
 * device catalog in complex64 (the files' precision), widened here. */
int dsm_sgt_fetch_fp64(dsm_sgt_t *h, int32_t i, int32_t kind, int32_t ir, void *out);
/* The l cut-off logged by the reference for frequency i (its llog-1,
 * SGTsh-MPI.f90:559, SGTpsv-MPI.f90:1098). */
int dsm_sgt_lcut(dsm_sgt_t *h, int32_t i, int32_t *llog_minus_1);

/* Diagnostic, no counterpart in the reference: the P-SV band pieces the
 * device assembles for frequency i (the per-zone band blocks and calbc of
 * SGTpsv-MPI.f90:572-604), band[nn][4 band rows][P0, P1, Q0, Q1] complex,
 * interleaved (re, im): 32 doubles per band column.  For checking the
 * device assembly against the host set-up; P-SV handles, between runs. */
int dsm_sgt_psv_band(dsm_sgt_t *h, int32_t i, double *band);
/* Write the catalog files of frequency i under outputDir with the
 * reference's names and byte layout (raw little-endian complex(4),
 * component fastest, no record markers). */
int dsm_sgt_write_catalog(dsm_sgt_t *h, int32_t i, const char *output_dir, const char *model_name);

/* Catalog output of a whole run (the write loop SGTpsv-MPI.f90:1103-1150,
 * SGTsh-MPI.f90:567-612, for every frequency of [i_begin, i_end] except
 * i = 0; each must be computed and resident).  The records are cast to
 * complex64 (plain form; the mixed forms' catalog is complex64) and put in
 * the files' order on the device; the call returns
 * once that is queued, and a copy stream and a writer thread move the
 * records to the host and write the files beside the engine's next
 * dsm_sgt_run.  Files, names and bytes are those of
 * dsm_sgt_write_catalog.  At most one output is in flight per handle: the
 * next output call (or dsm_sgt_write_catalog, dsm_sgt_destroy) first waits
 * for the previous one.  dsm_sgt_wait returns the writer's status. */
int dsm_sgt_write_catalogs_async(dsm_sgt_t *h, int32_t i_begin, int32_t i_end, const char *output_dir,
                                 const char *model_name);
/* The same records copied into host memory `dst` instead of files: per
 * frequency (ascending, i = 0 skipped) one block of complex(4) values,
 * TSGT a(num_tsgt, theta_n) of every receiver, then RSGT a(num_rsgt,
 * theta_n) of every receiver, then SYNN a(num_synn, theta_n); block size
 * theta_n (r_n (num_tsgt + num_rsgt) + num_synn) complex values.  dst
 * from dsm_sgt_host_alloc (pinned) copies at full bus speed; it must stay
 * valid until dsm_sgt_wait.  No counterpart in the reference (it writes
 * each record to disk). */
int dsm_sgt_copy_catalog_async(dsm_sgt_t *h, int32_t i_begin, int32_t i_end, void *dst, size_t dst_bytes);
/* Wait for the handle's asynchronous output; its status. */
int dsm_sgt_wait(dsm_sgt_t *h);
/* Pinned host memory for dsm_sgt_copy_catalog_async (hipHostMalloc). */
int dsm_sgt_host_alloc(size_t bytes, void **ptr);
void dsm_sgt_host_free(void *ptr);
/* Free the device blocks that destroyed engines left cached on HIP device
 * `device` for the process's later engines (an engine holds ~100 GB; the
 * next engine of similar sizes takes them over instead of allocating).  No
 * counterpart in the reference; for callers switching to engines of other
 * sizes (another mode, another grid) or handing the memory to other code. */
int dsm_sgt_release_cache(int32_t device);

int dsm_sgt_get_stats(dsm_sgt_t *h, dsm_sgt_stats_t *stats);
int dsm_sgt_reset_stats(dsm_sgt_t *h);
/* Drop all computed frequencies (device memory is kept for reuse). */
int dsm_sgt_clear(dsm_sgt_t *h);

const char *dsm_sgt_last_error(dsm_sgt_t *h);
void dsm_sgt_destroy(dsm_sgt_t *h);

#ifdef __cplusplus
}
#endif
#endif
