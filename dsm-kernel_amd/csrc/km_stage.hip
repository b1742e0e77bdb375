// km_stage.hip -- KernelMaker's catalog stage on the device: the catalog of
// one receiver radius read from the engines' device catalogs as KernelMaker
// reads the files (rdsgtomega, KernelMaker/fourier.f90:1-123), then per
// scattering point the sinc interpolation in distance (clsgt, :128-162) and
// the rotation into the receiver component's 3rd-order tensor and the
// moment-tensor-contracted 4th-order tensor (rsgt2h3freq :278-318,
// tsgt2h4freq :320-415), as calculateKernel calls them (calculateK.f90:
// 157-162).
//
// Compiled without FMA contraction, in the reference's operation order: the
// outputs follow its rounding.  The reduction-time phase factors are the
// reference's single-precision exp(cmplx(0.d0, x)); they come in as float
// tables built on the host.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "km_stage.h"

namespace dsm {

namespace {
constexpr double KM_PI = 3.1415926535897932;  // KernelMaker's module parameters

__device__ __forceinline__ double2 km_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// value of a complex(4) record, as the file holds it
__device__ __forceinline__ double2 as_c4(double2 v) { return make_double2((double)(float)v.x, (double)(float)v.y); }
// rdsgtomega's term: record * phase * 1e-21
__device__ __forceinline__ double2 rd_term(double2 v, double2 ph)
{
    const double2 t = km_mul(as_c4(v), ph);
    return make_double2(t.x * 1.0e-21, t.y * 1.0e-21);
}
}  // namespace

// Catalog element idx: complex64 (c64) or complex FP64
__device__ __forceinline__ double2 cat_value(const void *cat, size_t idx, int c64)
{
    if (c64) {
        const float2 f = ((const float2 *)cat)[idx];
        return make_double2((double)f.x, (double)f.y);
    }
    return ((const double2 *)cat)[idx];
}

// One lane per (frequency j, distance t): TSGT into om_t[(j T + t) 20 + c],
// RSGT into om_r[(j T + t) 10 + c]; SH records added into their P-SV slots
// (TSGT_SH 1 -> 7, 2..10 -> 12..20; RSGT_SH 1..5 -> 6..10).
__global__ __launch_bounds__(256) void km_rd_kernel(KmRdArgs a)
{
    const int T = a.theta_n;
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long)a.nf * T) return;
    const int j = (int)(idx / T), t = (int)(idx - (long)j * T);
    double2 *ot = a.om_t + idx * 20, *orr = a.om_r + idx * 10;
    const int pos = a.pos_psv[j];
    if (pos < 0) {  // i = 0: never read
        for (int c = 0; c < 20; c++) ot[c] = make_double2(0.0, 0.0);
        for (int c = 0; c < 10; c++) orr[c] = make_double2(0.0, 0.0);
        return;
    }
    const double2 ph = make_double2((double)a.ph[idx].x, (double)a.ph[idx].y);
    double2 vt[20], vr[10];
    {
        const size_t cs = (size_t)a.r_psv * T;
        const size_t ct = ((size_t)pos * 20 * a.r_psv + a.ir_psv) * T + t;
        const size_t cr = ((size_t)pos * 10 * a.r_psv + a.ir_psv) * T + t;
        for (int c = 0; c < 20; c++) {
            const double2 v = rd_term(cat_value(a.cat_t_psv, ct + c * cs, a.c64_psv), ph);
            vt[c] = make_double2(0.0 + v.x, 0.0 + v.y);
        }
        for (int c = 0; c < 10; c++) {
            const double2 v = rd_term(cat_value(a.cat_r_psv, cr + c * cs, a.c64_psv), ph);
            vr[c] = make_double2(0.0 + v.x, 0.0 + v.y);
        }
    }
    if (a.cat_t_sh && a.pos_sh[j] >= 0) {
        const int ps = a.pos_sh[j];
        const size_t cs = (size_t)a.r_sh * T;
        const size_t ct = ((size_t)ps * 10 * a.r_sh + a.ir_sh) * T + t;
        const size_t cr = ((size_t)ps * 5 * a.r_sh + a.ir_sh) * T + t;
        for (int c = 0; c < 10; c++) {
            const double2 v = rd_term(cat_value(a.cat_t_sh, ct + c * cs, a.c64_sh), ph);
            const int slot = c == 0 ? 6 : 10 + c;
            vt[slot] = make_double2(vt[slot].x + v.x, vt[slot].y + v.y);
        }
        for (int c = 0; c < 5; c++) {
            const double2 v = rd_term(cat_value(a.cat_r_sh, cr + c * cs, a.c64_sh), ph);
            vr[5 + c] = make_double2(vr[5 + c].x + v.x, vr[5 + c].y + v.y);
        }
    }
    for (int c = 0; c < 20; c++) ot[c] = vt[c];
    for (int c = 0; c < 10; c++) orr[c] = vr[c];
}

namespace {
// clsgt: NC components at `dist` from om[(j T + t) NC + c], into F
template <int NC>
__device__ __forceinline__ void km_interp(const KmPointArgs &a, const double2 *om, double dist, int j, double2 ph,
                                          double2 *F)
{
    const int T = a.theta_n;
    const int idelta = 1 + (int)((dist - a.thetaD[0]) / a.thetadelta);
    const int nh = (int)(a.ipdistance / a.thetadelta / 2.0);
    int is = idelta - nh, ie = idelta + nh;  // findIdeltaRegion, fourier.f90:166-190
    if (is < 1) {
        ie = ie + (1 - is);
        is = 1;
    }
    if (ie > T) {
        is = is - (ie - T);
        ie = T;
    }
    if (is < 1) is = 1;
    if (ie > T) ie = T;
    for (int c = 0; c < NC; c++) F[c] = make_double2(0.0, 0.0);
    double norm = 0.0;
    for (int kk = is; kk <= ie; kk++) {
        const double x = KM_PI * ((dist - a.thetaD[is - 1]) / a.thetadelta - (double)(kk - is));
        const double2 *src = om + ((size_t)j * T + (kk - 1)) * NC;
        if (x == 0.0) {
            for (int c = 0; c < NC; c++) {
                const double2 v = src[c];
                F[c] = make_double2(F[c].x + v.x, F[c].y + v.y);
            }
            norm = norm + 1.0;
        } else {
            const double w = sin(x) / x;
            for (int c = 0; c < NC; c++) {
                const double2 v = src[c];
                F[c] = make_double2(F[c].x + w * v.x, F[c].y + w * v.y);
            }
            norm = norm + w;
        }
    }
    for (int c = 0; c < NC; c++) F[c] = km_mul(make_double2(F[c].x / norm, F[c].y / norm), ph);
}

__device__ __forceinline__ double2 sc(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 ad(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 sb(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 ng(double2 a) { return make_double2(-a.x, -a.y); }
}  // namespace

// One lane per (point p, frequency j): h3[(p nf + j) 6 + k], h4 likewise.
__global__ __launch_bounds__(128) void km_point_kernel(KmPointArgs a)
{
    const int j = blockIdx.y;
    const int p = blockIdx.x * 128 + threadIdx.x;
    if (p >= a.npoints || j >= a.nf) return;
    const double *tg = a.trig + (size_t)p * 12;
    double2 *h3 = a.h3 + ((size_t)p * a.nf + j) * 6;
    double2 *h4 = a.h4 + ((size_t)p * a.nf + j) * 6;
    {
        double2 r[10];
        const float2 phr = a.ph_r[(size_t)p * a.nf + j];
        km_interp<10>(a, a.om_r, a.deltar[p], j, make_double2((double)phr.x, (double)phr.y), r);
        const double crq = tg[0], srq = tg[1], crq2 = tg[2], srq2 = tg[3];
        if (a.compo == 0) {  // Z
            h3[0] = r[0];
            h3[1] = sb(sc(r[2], crq2), r[3]);
            h3[2] = sb(sc(ng(r[2]), crq2), r[3]);
            h3[3] = sc(sc(r[1], 2.0), crq);
            h3[4] = sc(sc(r[1], 2.0), srq);
            h3[5] = sc(sc(r[2], 2.0), srq2);
        } else if (a.compo == 1) {  // R
            h3[0] = r[4];
            h3[1] = sb(sc(ng(r[6]), crq2), r[5]);
            h3[2] = sb(sc(r[6], crq2), r[5]);
            h3[3] = sc(sc(r[8], 4.0), crq);
            h3[4] = sc(sc(r[8], 4.0), srq);
            h3[5] = sc(sc(r[6], -2.0), srq2);
        } else {  // T: the reference negates its h3(3), zero at KernelMaker's points
            h3[0] = make_double2(0.0, 0.0);
            h3[1] = sc(r[7], srq2);
            h3[2] = make_double2(0.0, 0.0);
            h3[3] = sc(sc(r[9], 2.0), srq);
            h3[4] = sc(sc(r[9], -2.0), crq);
            h3[5] = sc(sc(r[7], -2.0), crq2);
        }
    }
    double2 t[21];  // t[k]: tsgtF(k)
    {
        const float2 phs = a.ph_s[(size_t)p * a.nf + j];
        km_interp<20>(a, a.om_t, a.deltas[p], j, make_double2((double)phs.x, (double)phs.y), t + 1);
    }
    const double csq = tg[4], ssq = tg[5], csq2 = tg[6], ssq2 = tg[7];
    const double cqs = tg[8], sqs = tg[9], cqs2 = tg[10], sqs2 = tg[11];
    const double *mt = a.mt;
    const double2 h01 = sc(t[7], cqs), h02 = sc(t[7], sqs), h03 = sc(t[8], csq), h04 = sc(t[8], ssq);
    const double2 h05 = sc(t[9], cqs2), h06 = sc(t[10], csq2), h07 = sc(t[11], csq2), h08 = sc(t[11], ssq2);
    const double2 h09 = sc(t[12], cqs2), h10 = sc(t[12], sqs2);
    const double2 h11 = sc(t[13], csq * cqs - ssq * sqs), h12 = sc(t[13], ssq * cqs + csq * sqs);
    const double2 h13 = sc(t[14], csq * cqs + ssq * sqs), h14 = sc(t[14], ssq * cqs - csq * sqs);
    const double2 h15 = sc(t[15], csq2 * cqs - ssq2 * sqs), h16 = sc(t[15], ssq2 * cqs + csq2 * sqs);
    const double2 h17 = sc(t[17], csq * cqs2 - ssq * sqs2), h18 = sc(t[17], ssq * cqs2 + csq * sqs2);
    const double2 h19 = sc(t[19], csq2 * cqs2 - ssq2 * sqs2), h20 = sc(t[19], ssq2 * cqs2 + csq2 * sqs2);
    const double2 h21 = sc(t[20], csq2 * cqs2 + ssq2 * sqs2), h22 = sc(t[20], ssq2 * cqs2 - csq2 * sqs2);
    const double2 h23 = sc(t[16], csq2 * cqs + ssq2 * sqs), h24 = sc(t[16], ssq2 * cqs - csq2 * sqs);
    const double2 h25 = sc(t[18], csq * cqs2 + ssq * sqs2), h26 = sc(t[18], ssq * cqs2 - csq * sqs2);
    // rr
    h4[0] = ad(sb(ad(ad(sc(t[1], mt[0]), sc(sb(h05, t[2]), mt[1])), sc(sb(ng(h05), t[2]), mt[2])),
                  sc(t[5], cqs * mt[3] + sqs * mt[4])),
               sc(sc(t[9], sqs2), mt[5]));
    // tt
    h4[1] = ad(ad(ad(ad(ad(sc(sb(h06, t[3]), mt[0]), sc(ad(sb(sb(ad(h21, h19), h07), h09), t[4]), mt[1])),
                        sc(ad(ad(sb(sb(ng(h21), h19), h07), h09), t[4]), mt[2])),
                     sc(ad(sb(h23, h15), h01), mt[3])),
                  sc(ad(sb(ng(h24), h16), h02), mt[4])),
               sc(sb(ad(ng(h22), h20), h10), mt[5]));
    // pp
    h4[2] = ad(ad(ad(ad(ad(sc(sb(ng(h06), t[3]), mt[0]), sc(ad(sb(ad(sb(ng(h21), h19), h07), h09), t[4]), mt[1])),
                        sc(ad(ad(ad(ad(h21, h19), h07), h09), t[4]), mt[2])),
                     sc(ad(ad(ng(h23), h15), h01), mt[3])),
                  sc(ad(ad(h24, h16), h02), mt[4])),
               sc(sb(sb(h22, h20), h10), mt[5]));
    // rt
    h4[3] = ad(ad(ad(ad(ad(sc(sc(ng(t[6]), csq), mt[0]), sc(ad(sb(h25, h17), h03), mt[1])),
                        sc(ad(ad(ng(h25), h17), h03), mt[2])),
                     sc(ad(ng(h13), h11), mt[3])),
                  sc(ad(h14, h12), mt[4])),
               sc(sb(ng(h26), h18), mt[5]));
    // rp
    h4[4] = ad(ad(ad(ad(ad(sc(ng(sc(t[6], ssq)), mt[0]), sc(ad(sb(h26, h18), h04), mt[1])),
                        sc(ad(ad(ng(h26), h18), h04), mt[2])),
                     sc(ad(ng(h14), h12), mt[3])),
                  sc(sb(ng(h13), h11), mt[4])),
               sc(ad(h25, h17), mt[5]));
    // tp
    h4[5] = ad(ad(ad(ad(ad(sc(sc(t[10], ssq2), mt[0]), sc(sb(ad(h22, h20), h08), mt[1])),
                        sc(sb(sb(ng(h22), h20), h08), mt[2])),
                     sc(sb(h24, h16), mt[3])),
                  sc(ad(h23, h15), mt[4])),
               sc(sb(h21, h19), mt[5]));
}

// vectorFFT_double (fourier.f90:418-447) of one spectrum per workgroup.
// The reference builds c(0 : N-1), N = 2 np1, from ccvec(imin:imax) and
// c(np1 + i) = conj(c(np1 - i)) for i = imin..np1-1, transforms it with
// cdft (the unnormalised inverse DFT) and keeps Re x(n) exp(omegai n /
// samplingHz) / tlen * 1e3 on the window.  Re x is the inverse DFT of the
// Hermitian part H(k) = (c(k) + conj(c(N-k))) / 2, whose transform is real:
// its even and odd samples are the real and imaginary parts of an N/2-point
// inverse DFT of Z(k) = E(k) + i O(k), E(k) = H(k) + H(k + N/2),
// O(k) = (H(k) - H(k + N/2)) e^{2 pi i k / N}.  Z is transformed in LDS by an
// in-place radix-2 decimation-in-frequency FFT (output in bit-reversed
// order).  LDS: N/2 complex values.
namespace {
__device__ __forceinline__ double2 km_spec(const KmFftArgs &a, const double2 *cc, int k)
{
    const int np1 = a.np1;
    if (k >= np1 + a.imin && k - np1 <= np1 - 1) {  // the mirrored half
        const int src = 2 * np1 - k;
        const double2 v = (src >= a.imin && src <= a.imax) ? cc[src - a.imin] : make_double2(0.0, 0.0);
        return make_double2(v.x, -v.y);
    }
    return (k >= a.imin && k <= a.imax) ? cc[k - a.imin] : make_double2(0.0, 0.0);
}
__device__ __forceinline__ double2 km_herm(const KmFftArgs &a, const double2 *cc, int k)
{
    const int N = 2 * a.np1;
    const double2 u = km_spec(a, cc, k), v = km_spec(a, cc, (N - k) % N);
    return make_double2(0.5 * (u.x + v.x), 0.5 * (u.y - v.y));
}
}  // namespace

__global__ __launch_bounds__(256) void km_fft_kernel(KmFftArgs a)
{
    extern __shared__ double2 z[];
    const int M = a.np1;  // N / 2
    const int v = blockIdx.x;
    const double2 *cc = a.ccvec + (size_t)v * (a.imax - a.imin + 1);
    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        const double2 h0 = km_herm(a, cc, k), h1 = km_herm(a, cc, k + M);
        const double2 e = make_double2(h0.x + h1.x, h0.y + h1.y);
        const double2 w = a.tw2[k];  // e^{2 pi i k / N}
        const double2 o = km_mul(make_double2(h0.x - h1.x, h0.y - h1.y), w);
        z[k] = make_double2(e.x - o.y, e.y + o.x);  // E + i O
    }
    __syncthreads();
    // radix-2 DIF, inverse sign: twiddle e^{2 pi i j / M} = tw2[2 j (M / 2h)]
    for (int hh = M / 2; hh >= 1; hh >>= 1) {
        const int stride = M / (2 * hh);
        for (int b = threadIdx.x; b < M / 2; b += blockDim.x) {
            const int blk = b / hh, jj = b - blk * hh;
            const int i0 = blk * 2 * hh + jj, i1 = i0 + hh;
            const double2 x0 = z[i0], x1 = z[i1];
            z[i0] = make_double2(x0.x + x1.x, x0.y + x1.y);
            z[i1] = km_mul(make_double2(x0.x - x1.x, x0.y - x1.y), a.tw2[2 * jj * stride]);
        }
        __syncthreads();
    }
    const int lg = 31 - __clz(M);
    const double sampling = (double)(2 * a.np1) / a.tlen;
    double *out = a.rvec + (size_t)v * (a.iwe - a.iws + 1);
    for (int n = a.iws + threadIdx.x; n <= a.iwe; n += blockDim.x) {
        const int m = n >> 1;
        const int r = lg ? (int)(__brev((unsigned)m) >> (32 - lg)) : 0;
        const double2 zm = z[r];
        const double x = (n & 1) ? zm.y : zm.x;
        out[n - a.iws] = x * exp(a.omegai * (double)n / sampling) / a.tlen * 1.0e3;
    }
}

void launch_km_fft(const KmFftArgs &a, int nvec, hipStream_t s)
{
    if (nvec <= 0) return;
    const int lds = (int)sizeof(double2) * a.np1;  // up to 128 KiB of the CU's 160
    (void)hipFuncSetAttribute((const void *)km_fft_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    hipLaunchKernelGGL(km_fft_kernel, dim3(nvec), dim3(256), lds, s, a);
}

// Twiddle table of the transform: tw[k] = e^{2 pi i k / N}, k < N / 2.
__global__ __launch_bounds__(256) void km_twiddle_kernel(double2 *tw, int N)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N / 2) return;
    double sn, cs;
    sincos(2.0 * 3.141592653589793 * (double)k / (double)N, &sn, &cs);
    tw[k] = make_double2(cs, sn);
}

void launch_km_twiddles(double2 *tw, int N, hipStream_t s)
{
    if (N < 2) return;
    hipLaunchKernelGGL(km_twiddle_kernel, dim3((N / 2 + 255) / 256), dim3(256), 0, s, tw, N);
}

// Twiddle storage of the allocation-free transform: KM_FFT_MAX_NP1 values
// (128 KiB) per device, held until the process ends.  Host code reaches it
// only through the address km_static_twiddles returns for the device.
static __device__ __align__(256) double2 scratch[KM_FFT_MAX_NP1];

hipError_t km_static_twiddles(int device, double2 **tw)
{
    static std::mutex mu;
    static std::vector<double2 *> addr;  // per device
    *tw = nullptr;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = (double2 *)p;
    }
    *tw = addr[device];
    return hipSuccess;
}

void launch_km_rd(const KmRdArgs &a, hipStream_t s)
{
    const long n = (long)a.nf * a.theta_n;
    if (n <= 0) return;
    hipLaunchKernelGGL(km_rd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
}

void launch_km_points(const KmPointArgs &a, hipStream_t s)
{
    if (a.npoints <= 0 || a.nf <= 0) return;
    hipLaunchKernelGGL(km_point_kernel, dim3((a.npoints + 127) / 128, a.nf), dim3(128), 0, s, a);
}

}  // namespace dsm
