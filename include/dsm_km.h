/*
 * dsm_km.h -- KernelMaker's catalog stage on the GPU, reading the catalogs of
 * dsm_sgt engines (dsm_sgt.h) in device memory instead of the catalog files.
 *
 * The reference's KernelMaker reads, for every radius of its kernel grid,
 * the RSGT and TSGT records of all frequencies from the files SGTpsv and
 * SGTsh wrote (rdsgtomega, src/KernelMaker/fourier.f90:1-123; call sites
 * KernelMaker-1.0.0.f90:719-736), and for every scattering point at that
 * radius interpolates them in distance (clsgt, fourier.f90:128-162) and
 * rotates them into the 3rd-order tensor of the receiver component and the
 * 4th-order tensor contracted with the moment tensor (rsgt2h3freq
 * :278-318, tsgt2h4freq :320-415), as calculateKernel does
 * (calculateK.f90:131-162).  dsm_km_points does that for a set of points
 * at one receiver radius of the engines, with the same arithmetic: the
 * complex(4) values of the records, the reference's single-precision
 * reduction-time phase factors, the SH records added into their P-SV slots.
 *
 * Arrays are plain pointers; complex values interleaved (re, im), Fortran
 * complex(8).  Status codes and dsm_sgt_last_error() as in dsm_sgt.h.
 */
#ifndef DSM_KM_H
#define DSM_KM_H

#include "dsm_sgt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* KernelMaker's parameters of the stage (its input file, others.f90:60-90,
 * and module parameters). */
typedef struct dsm_km_params {
    int32_t fmin, fmax;   /* frequency indices fmin:fmax; each i != 0 computed and resident in the engines */
    int32_t compo;        /* receiver component: 0 Z, 1 R, 2 T */
    int32_t reserved;
    double c_red_reci;    /* reciprocal reduction velocity (s per degree) */
    double ipdistance;    /* width of the sinc interpolation window (degrees) */
    double mt[6];         /* moment tensor rr, tt, pp, rt, rp, tp */
} dsm_km_params_t;

/* npoints scattering points at receiver radius index ir of the P-SV engine
 * `psv` (and, when `sh` is not NULL and holds the same radius, the SH
 * engine's records in their P-SV slots, as KernelMaker adds SGTsh's files
 * for radii above the CMB).  Per point: deltar, deltas its distances
 * (degrees) from the receiver and from the source; trig[12] the cosines and
 * sines of calculateSineCosine (others.f90:768-813): crq srq crq2 srq2 csq ssq
 * csq2 ssq2 cqs sqs cqs2 sqs2.  Out: h3, h4 complex(8) (6, fmin:fmax) per
 * point, consecutive (npoints * 6 * (fmax - fmin + 1) values each).  The
 * engines' catalogs are left unchanged; runs on the P-SV engine's device
 * and stream, synchronously. */
int dsm_km_points(dsm_sgt_t *psv, dsm_sgt_t *sh, int32_t ir, const dsm_km_params_t *params, int32_t npoints,
                  const double *deltar, const double *deltas, const double *trig, void *h3, void *h4);

/* The frequency-to-time transform of KernelMaker (vectorFFT_double,
 * fourier.f90:418-447; tensorFFT_double :449-479 applies it to every
 * component) for nvec spectra at once: ccvec complex(8) (imin:imax) per
 * spectrum, consecutive; the spectrum extended to N = 2 np1 points with
 * c(np1 + i) = conj(c(np1 - i)), i = imin..np1-1, transformed by the
 * inverse DFT, and rvec(i) = Re x(i) exp(omegai t_i) / tlen * 1e3 for
 * t_i = i tlen / N, i = iws..iwe (the reference's de-damping of the complex
 * frequencies omega - i omegai); rvec real(8) (iws:iwe) per spectrum.
 * np1 a power of two, 2..8192; 0 <= imin <= imax < N, 0 <= iws <= iwe < N.
 * Runs on the engine's device and stream, synchronously (h: any created
 * engine; its catalogs are not touched). */
int dsm_km_vectorfft(dsm_sgt_t *h, int32_t nvec, int32_t imin, int32_t imax, int32_t np1, const void *ccvec,
                     double omegai, double tlen, int32_t iws, int32_t iwe, double *rvec);

/* dsm_km_vectorfft on device arrays, without allocating: no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * ccvec and rvec as in dsm_km_vectorfft, but device memory of the engine's
 * GPU (each checked to lie within one device allocation); stream: a HIP
 * stream of that GPU, on which the call returns at once, or NULL for the
 * engine's stream, on which it returns when the transform is done.  The
 * twiddle table lives in static device storage of 8192
 * complex(8) values (128 KiB, held by every GPU the library runs on until
 * the process ends): np1 above 8192 is refused.  Every call rewrites that
 * GPU's table, so calls that overlap on one GPU (other streams, other
 * threads) race; no call may run while another on the same GPU is in
 * flight. */
int dsm_km_vectorfft_noalloc(dsm_sgt_t *h, void *stream, int32_t nvec, int32_t imin, int32_t imax, int32_t np1,
                             const void *ccvec, double omegai, double tlen, int32_t iws, int32_t iwe, double *rvec);

#ifdef __cplusplus
}
#endif
#endif
