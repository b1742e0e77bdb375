// km_stage.h -- argument structs and launchers of the KernelMaker catalog
// stage (km_stage.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace dsm {

// rdsgtomega of the RSGT and TSGT records of one receiver radius for nf
// frequencies (fmin + j): om_t[(j T + t) 20 + c], om_r[(j T + t) 10 + c]
struct KmRdArgs {
    const void *cat_t_psv, *cat_r_psv;     // engine catalogs [pos][comp][rec][T]
    int r_psv, ir_psv, c64_psv;            // c64: complex64 catalog, else complex FP64
    const void *cat_t_sh, *cat_r_sh;       // SH engine (null: none)
    int r_sh, ir_sh, c64_sh;
    const int *pos_psv, *pos_sh;           // [j] catalog position, -1: not read (i = 0)
    const float2 *ph;                      // [j][T] the reference's exp(cmplx(0, -2 pi i / tlen redtime))
    int nf, theta_n;
    double2 *om_t, *om_r;
};
void launch_km_rd(const KmRdArgs &a, hipStream_t s);

// clsgt + rsgt2h3freq + tsgt2h4freq per point p and frequency j
struct KmPointArgs {
    const double2 *om_t, *om_r;
    const double *thetaD;                  // catalog distances (deg)
    int theta_n, nf, npoints, compo;
    double thetadelta, ipdistance;
    const double *deltar, *deltas;         // [p] (deg)
    const double *trig;                    // [p][12]: crq srq crq2 srq2 csq ssq csq2 ssq2 cqs sqs cqs2 sqs2
    const float2 *ph_r, *ph_s;             // [p][j] clsgt's exp(cmplx(0, -2 pi j / tlen distance c_red_reci))
    double mt[6];
    double2 *h3, *h4;                      // [p][j][6]
};
void launch_km_points(const KmPointArgs &a, hipStream_t s);

// vectorFFT_double of nvec spectra ccvec[v][imax - imin + 1] into
// rvec[v][iwe - iws + 1]; tw2[k] = e^{2 pi i k / N}, k < N / 2, N = 2 np1;
// np1 a power of two, at most KM_FFT_MAX_NP1 (its N / 2 complex values in LDS)
constexpr int KM_FFT_MAX_NP1 = 8192;
struct KmFftArgs {
    const double2 *ccvec;
    int imin, imax, np1, iws, iwe;
    double omegai, tlen;
    const double2 *tw2;
    double *rvec;
};
void launch_km_fft(const KmFftArgs &a, int nvec, hipStream_t s);
// tw[k] = e^{2 pi i k / N}, k < N / 2 (the table KmFftArgs::tw2 reads)
void launch_km_twiddles(double2 *tw, int N, hipStream_t s);
// Device address of the static twiddle storage (KM_FFT_MAX_NP1 values) of
// `device`, which must be current; looked up once per device.
hipError_t km_static_twiddles(int device, double2 **tw);

}  // namespace dsm
