/*
 * dsm_oracle.h -- CPU restatement of the DSM strain-Green's-tensor (SGT)
 * engines SGTsh / SGTpsv of IPGP/DSM-Kernel.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the parity checker for the
 * HIP implementation and the timed CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library never links or calls it.
 *
 * Every routine restates the reference routine of the same name, with the
 * reference's argument list (all arguments by pointer, Fortran column-major
 * arrays, 1-based index arithmetic), so that each call site of the
 * restated driver can be switched to the reference routine compiled from
 * its own sources (oracle/_ref) -- see sh_oracle.c, "function table".
 */
#ifndef DSM_ORACLE_H
#define DSM_ORACLE_H

#include <stdint.h>
#include <complex.h>

#define ORC_MAXZONE 64

/* Parsed inputs: the stdin parameter file (src/SGTsh/others.f90:1-60),
 * DSM.config (others.f90:62-92) and the polynomial model file
 * (SGTsh-MPI.f90:83-134, SGTpsv-MPI.f90:100-105).  Polynomials are stored
 * coefficient-fastest, as the reference's rrho(4,nzone). */
typedef struct {
    int32_t nzone;
    int32_t maxlmax;
    int32_t imin, imax;
    int32_t rsgtswitch, tsgtswitch, synnswitch, psgtswitch;
    double vrmin[ORC_MAXZONE], vrmax[ORC_MAXZONE];
    double rho[ORC_MAXZONE][4], vpv[ORC_MAXZONE][4], vph[ORC_MAXZONE][4];
    double vsv[ORC_MAXZONE][4], vsh[ORC_MAXZONE][4], eta[ORC_MAXZONE][4];
    double qmu[ORC_MAXZONE], qkappa[ORC_MAXZONE];
    double re, ratc, ratl, omegai_seed;   /* DSM.config values, verbatim */
    double tlen;
    double rmin_, rmax_, rdelta_;         /* receiver radii (km) */
    double r0min;                         /* source radius (km) */
    double thetamin, thetamax, thetadelta;/* epicentral distances (deg) */
    /* receiver lists in place of the grids (counts > 0; DSM3D_forward's
     * radius and distance files, src/DSM3D_forward/mpi-tipsv.f90:163-210) */
    int32_t r_list_n, theta_list_n;
    const double *r_list;
    const double *theta_list;
} orc_input_t;

/* Sizes of one SH run (after the fluid/inner-core removal and the
 * receiver trim of SGTsh-MPI.f90:66-162). */
typedef struct {
    int32_t r_n, theta_n, nn, nzone;
    double r0, spo;
    int32_t spn, intir0;
} orc_sh_info_t;

/* use_ref: 0 = restated C routines, 1 = the reference's own routines
 * loaded from oracle/_ref/libref_sh.so (see orc_sh_load_ref). */
int orc_sh_info(const orc_input_t *in, orc_sh_info_t *info);
int orc_sh_load_ref(const char *path);
int orc_sh_mix(uint64_t mask);
/* One frequency of SGTsh (SGTsh-MPI.f90:389-613).  Outputs are the FP64
 * accumulators before the complex64 cast, Fortran layout
 * tsgt(10,r_n,theta_n), rsgt(5,r_n,theta_n), synn(5,theta_n).
 * llog receives the reference's log value llog-1 (SGTsh-MPI.f90:559). */
int orc_sh_omega(const orc_input_t *in, int32_t i, int32_t use_ref,
                 double complex *tsgt, double complex *rsgt,
                 double complex *synn, int32_t *llog);

/* Sizes of one P-SV run (SGTpsv-MPI.f90:120-240; no receiver trim). */
typedef struct {
    int32_t r_n, theta_n, nn, nzone;
    double r0, spo;
    int32_t spn, intir0, nslay, nllay;
} orc_psv_info_t;

int orc_psv_info(const orc_input_t *in, orc_psv_info_t *info);
int orc_psv_load_ref(const char *path);
int orc_psv_mix(uint64_t mask);
/* One frequency of SGTpsv (SGTpsv-MPI.f90:542-1084): FP64 accumulators
 * tsgt(20,r_n,theta_n), rsgt(10,r_n,theta_n), synn(10,theta_n); llog is
 * the logged value llog-1 (SGTpsv-MPI.f90:1033). */
int orc_psv_omega(const orc_input_t *in, int32_t i, int32_t use_ref,
                  double complex *tsgt, double complex *rsgt,
                  double complex *synn, int32_t *llog);

/* Test hook of both drivers: while on, every call of calamp (the l cut-off
 * update, SGTpsv-MPI.f90:902-905, SGTsh-MPI.f90:476-478) is recorded as
 * (l, amplitude the call computes) into the caller's arrays, at most cap
 * entries.  orc_calamp_trace(l, amp, cap) turns it on and clears the count;
 * cap = 0 turns it off.  orc_calamp_count() returns the calls recorded
 * (possibly more than cap).  Not thread-safe; one driver call at a time. */
int orc_calamp_trace(int32_t *l, double *amp, int32_t cap);
int orc_calamp_count(void);
/* calcutd of SGTpsv and SGTsh (never called by the reference's drivers),
 * with the reference's argument lists. */
void orc_psv_calcutd(const int *nzone, const int *nnlayer, const int *nnl, const double complex *tmpc,
                     const double *rat, const int *nn, const int *iphase, const double *spo, const int *spn,
                     const double *ra, const int *kkdr, int *kc);
void orc_sh_calcutd(const int *nzone, const int *nnl, const double complex *tmpr, const double *rat, const int *nn,
                    const double *ra, int *kc);
/* Test hook: copy the P-SV solution g0 (nn values, at most cap) of the
 * (l, m, imt) right-hand side of the next orc_psv_omega calls into buf;
 * buf = NULL turns it off. */
int orc_psv_capture(int32_t l, int32_t m, int32_t imt, double complex *buf, int32_t cap);
/* (internal) record one call: l and the n complex values calamp reads */
void orc_trace_calamp(int l, const double complex *g, int n);

#endif
