// kernels.h -- host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "psv_band_lane.h"
#include "psv_lane.h"
#include "sh_lane.h"

namespace dsm {

// SH lane bodies and argument structs: sh_lane.h
void launch_sh_solve(const ShSolveArgs &a, int nblocks, hipStream_t s);
void launch_sh_coef(const ShCoefArgs &a, int nblocks, hipStream_t s);
void launch_sh_depth(const ShDepthArgs &a, hipStream_t s);
void launch_sh_cut(const ShCutArgs &a, int nslots, hipStream_t s);

// P-SV lane bodies and argument structs: psv_lane.h
// excitation g of every (frequency, l) into a.gexc, then the band solve;
// grids (l blocks, nslots active frequencies)
void launch_psv_excite(const PsvSolveArgs &a, int nslots, hipStream_t s);
// fast: the fused FP64 operation family of the mixed forms (dev_common.h)
void launch_psv_solve(const PsvSolveArgs &a, int nslots, bool fast, hipStream_t s);
int psv_solve_blocks(int l0, int lend);  // l blocks per frequency of launch_psv_solve
void launch_psv_depth(const PsvDepthArgs &a, hipStream_t s);
void launch_psv_need(const PsvNeedArgs &a, hipStream_t s);
// l cut-off replay on the device (batched forms) and the bf16 row extents
void launch_psv_cut(const PsvCutArgs &a, hipStream_t s);

// Band pieces on the device (psv_band.hip; PsvBandArgs in psv_band_lane.h).
void launch_psv_band(const PsvBandArgs &a, int nslots, hipStream_t s);
void launch_psv_nlr(const int *kstart, const int *rlow, const int *nl, int nslots, int nblk, int r_n, int *nlr,
                    hipStream_t s);
// per-block factor storage of the batched forms (PsvSolveArgs::loff)
void launch_psv_loff(const int *kstart, int nblocks, int n, int klo, int fs, int ckpt, long long *loff,
                     long long *total, hipStream_t s);
// compact band of the fused family (PsvSolveArgs::cband)
void launch_psv_compact(const double2 *band, const int *kind, int nn, const int *slot_list, int nslots,
                        double2 *cband, hipStream_t s);
void launch_psv_coef(const PsvCoefArgs &a, int nblocks, hipStream_t s);
void launch_psv_l0(const PsvL0Args &a, hipStream_t s);

// ---- Legendre / vector-harmonic tables ------------------------------------
// FP64 upward recurrence for P_l^m, m<=3 (trial.f90:85-117), one lane per
// distance; writes the real basis functions of the synthesis for
// l in [0, lmax] and |m| = 1, 2: tab[mu-1][b][l][theta_pad].
void launch_legendre_sh(const double *theta_rad, int theta_n, int theta_pad, int lmax, int lcap,
                        double *tab, hipStream_t s);
// P-SV functions F0..F9 of |m| = 0, 1, 2 (legendre_lane.h): tab[plane][l][theta_pad]
void launch_legendre_psv(const double *theta_rad, int theta_n, int theta_pad, int lmax, int lcap, double *P,
                         double *tab, hipStream_t s);

// ---- zero the l tail of the coefficient rows ------------------------------
void launch_zero_tail(double2 *E, int nslot_rows, int rows_per_slot, const int *slots,
                      const int *nl, int lcap, hipStream_t s);

// ---- synthesis contraction -------------------------------------------------
// acc[(slot,r)][theta] = sum_{j,l} E[slot][q0+j][r][l] * T[b_j][l][theta]
// as a real GEMM on FP64 MFMA (complex A rows split into re/im rows), then
// out[..][comp_t] (=|+=) sign_t * acc for the job's one or two targets.
struct SynthJob {
    int q0, nb, b[4];      // first row, number of theta functions, their table planes
    int comp[2];           // target components (comp[1] < 0: one target)
    double sign[2];        // +-1
    int accumulate;        // add into the targets instead of writing them
    int ncomp;             // record width
    void *out;             // catalog of this kind: [pos][ncomp][r_eff][theta_n], complex FP64 or (c64) complex64
    int c64;
    double msign;          // F(l, pi - theta) = msign (-1)^l F(l, theta) for all the job's functions
};
struct SynthArgs {
    const double2 *E;      // [slot][nq][r_eff][lcap]
    int nq, r_eff;
    const double *tab;     // [plane][lcap][theta_pad]
    int theta_n, theta_pad, lcap;
    int nslot;
    const int *nl;         // [slot] number of l terms
    const int *slot_out;   // [slot] -> catalog frequency position
    const SynthJob *jobs;  // device array
    int njobs;
    int mirror, nh;        // FP64: distance symmetry, nh = (theta_n + 1) / 2 columns computed
    unsigned long long *nonfinite;  // count of non-finite values stored (may be null)
    // FP64 kernel, split K: ksplit > 1 cuts the l series of every theta
    // function into ksplit parts of kchunk chunks of 16 degrees, fixed by
    // the degree (the same parts for every tile, batch size and device),
    // run by separate workgroups; a second launch adds their partial sums
    // (kpart, synth_f64_kpart_size doubles) in part order.  knb: the most
    // theta functions of a job of the launch.
    int ksplit, kchunk, knb;
    double *kpart;
    int kphase;            // set by launch_synth_f64: 0 whole K, 1 partial sums, 2 their sum
};
void launch_synth_f64(const SynthArgs &a, hipStream_t s);
// doubles of SynthArgs::kpart for a launch of a.njobs jobs with a.ksplit parts
size_t synth_f64_kpart_size(const SynthArgs &a);
// workgroups of one launch of launch_synth_f64 without split K
size_t synth_f64_tiles(const SynthArgs &a);
// Mixed form (synth32.hip): FP32 operands on FP32 MFMA, FP64 accumulation
// across 32-wide K chunks; tab32 is the theta table rounded to FP32.
void launch_synth_f32(const SynthArgs &a, const float *tab32, hipStream_t s);
void launch_tab_to_f32(const double *src, float *dst, size_t n, hipStream_t s);

// Mixed forms on BF16 MFMA (synth_b16.hip): operands split into ns = 2 or
// 3 bf16 terms, degrees stored even-l first; with `mirror` the columns
// c < nh = (theta_n+1)/2 give distances c and theta_n-1-c.
struct SynthB16Args {
    const unsigned short *Eb;  // planes q ns + term of rows m = (r nslot + slot) 2 + re|im (b16_at)
    const unsigned short *Tb;  // planes term nplanes + plane of nhp rows (b16_at)
    int nq, r_eff, nslot, lcap, nplanes;
    int theta_n, nh, nhp, mirror;
    const int *nlr;            // [r][slot]: l extent of every (receiver, slot) row
    const int *slot_out;
    const SynthJob *jobs;
    int njobs;
    double *stash;             // [workgroup][8 waves][32][64 lanes]
    const int *tiles;          // optional [grid]: tile (job ny + ty) nx + tx per workgroup, -1 none
    int ntiles;
    unsigned long long *nonfinite;  // count of non-finite values stored (may be null)
};
// zero degrees l in [nl(slot), lend] of every bf16 row
void launch_zero_tail_b16(unsigned short *Eb, int nq, int ns, int nslot, int r_eff, int lcap, const int *nl, int lend,
                          hipStream_t s);
void launch_split_rows(const double2 *E, int nslot, int nq, int r_eff, int lcap, int ns, unsigned short *Eb,
                       const int *nl, hipStream_t s);
// degree 0 of nrows bf16 rows (planes of M rows, b16_at) from one-degree
// rows one[row]
void launch_b16_put_l0(const unsigned short *one, unsigned short *Eb, size_t nrows, size_t M, int lcap,
                       hipStream_t s);
// zero the degrees l in [lfrom, lcap) of nrows bf16 rows (planes of M rows)
void launch_zero_pad_b16(unsigned short *Eb, size_t nrows, size_t M, int lcap, int lfrom, hipStream_t s);
void launch_split_table(const double *tab, int nplanes, int lcap, int theta_pad, int nh, int nhp, int ns,
                        unsigned short *Tb, hipStream_t s);
size_t synth_b16_grid(int nslot, int r_eff, int nhp, int njobs);
// a distance tile with at most this many columns below nh (the last one)
// runs a quarter of the MFMAs of a full tile
constexpr int SYNTH_B16_NARROW = 32;
// tile geometry of launch_synth_b16: nx distance tiles, ny row tiles
void synth_b16_tiles(int nslot, int r_eff, int nhp, int &nx, int &ny);
void launch_synth_b16(const SynthB16Args &a, int ns, hipStream_t s);

// ---- catalog records (catalog.hip) ------------------------------------------
// One kind of the component-major catalog src [pos][ncomp][nrec][theta_n]
// (complex FP64, cast; or complex64) in the files' record order: for the k-th position of
// pos_list, dst[k dst_stride + dst_off + (rec theta_n + t) ncomp + c].
// ncomp <= 20.
struct PackArgs {
    const void *src;       // complex FP64, or complex64 with c64
    int c64;
    int ncomp, nrec, theta_n;
    const int *pos_list;
    float2 *dst;
    size_t dst_stride, dst_off;  // in complex64 values
};
void launch_pack_records(const PackArgs &a, int npos, hipStream_t s);

}  // namespace dsm
