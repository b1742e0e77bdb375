"""CPU: the oracle of KernelMaker's catalog stage and frequency-to-time
transform (oracle/km_oracle.c) against the reference's own routines
(src/KernelMaker/fourier.f90 with module.f90, built by flang from their
sources into oracle/_ref/libref_km.so behind oracle/km_harness.f90) on the
same inputs, and against what the mathematics fixes:

* rdsgtomega: the records are written as catalog files in the reference's
  names and layout and read by the reference's own rdsgtomega (P-SV and SH
  kinds, the SH -> P-SV slot map, SYNN's halved slots, the reduction-time
  phase and 1e-21);
* clsgt + rsgt2h3freq + tsgt2h4freq at random scattering points, all three
  receiver components;
* vectorFFT_double (cdft, bitrv2) on random spectra;
* closed forms: cdft is the unnormalised inverse DFT (numpy); the sinc
  interpolation of a distance-independent catalog returns it exactly (the
  weights are normalised); at a catalog distance it returns that distance's
  value; a single-frequency spectrum transforms to a damped-out cosine.

The reference forms its phase factors in single precision
(exp(cmplx(0.d0, x)) with a default-kind CMPLX); the oracle takes them from
the C library's cosf / sinf of the float argument, as flang's runtime does:
rdsgtomega's catalogs agree bit for bit, the points to ~1e-15 (the
sin(x) / x weights and the rotations round alike up to operation order).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol

REF = os.path.join(ol.ORACLE_DIR, "_ref", "libref_km.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF), reason="reference routines not built (no reference sources)")

P = C.c_void_p


def _orc():
    L = ol.lib()
    L.orc_km_rdsgt.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_double, C.c_double, P, P]
    L.orc_km_clsgt.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_double, C.c_double,
                               C.c_double, C.c_double, P, P]
    L.orc_km_rsgt2h3.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P]
    L.orc_km_tsgt2h4.argtypes = [C.c_int, C.c_int, P, P, P, P]
    L.orc_km_vectorfft.argtypes = [C.c_int, C.c_int, C.c_int, P, P, C.c_double, C.c_double, C.c_int, C.c_int, P, P]
    L.orc_km_cdft.argtypes = [C.c_int, C.c_double, C.c_double, P, P]
    return L


def _ref():
    R = C.CDLL(REF)
    R.km_ref_init.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                              C.c_double, P, C.c_int, C.c_char_p, C.c_char_p]
    R.km_ref_read.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
    R.km_ref_omega.argtypes = [C.c_int, P]
    R.km_ref_point.argtypes = [C.c_double, C.c_double, P, P, P]
    R.km_ref_vectorfft.argtypes = [C.c_int, C.c_int, C.c_int, P, P, C.c_double, C.c_double, C.c_int, C.c_int]
    return R


def ptr(a):
    return a.ctypes.data_as(P)


# catalog: distances 0..30 deg every 0.5 deg, frequencies 0..24 (0 is never
# computed), source 6000 km, scattering radius 6200 km
TH = dict(theta_n=61, thetamin=0.0, thetadelta=0.5)
FMIN, FMAX, TLEN, CRED, IPD = 0, 24, 1638.4, 0.08, 4.0
RS, RX = 6000.0, 6200.0
KINDS = ((200, 20, 20), (100, 10, 20), (20, 10, 10), (10, 5, 10), (2, 10, 10), (1, 5, 10))  # (ipsvorsh, num, num_psv)


def _i7(v):
    return f"{v:7d}".replace(" ", "0")


def _name(ipsvorsh, i):
    suf = "PSV" if ipsvorsh in (2, 20, 200) else "SH"
    if ipsvorsh in (200, 100):
        return f"TSGT/m4.{_i7(int(RS * 1e3))}.{_i7(int(RX * 1e3))}.{_i7(i)}.TSGT_{suf}"
    kind = "RSGT" if ipsvorsh in (20, 10) else "SYNN"
    return f"RSGT/m4.{_i7(int(RX * 1e3))}.{_i7(i)}.{kind}_{suf}"


def _records(rng):
    """Random complex64 records per kind and frequency: [f][theta][num]."""
    nf, T = FMAX - FMIN + 1, TH["theta_n"]
    return {k: (rng.standard_normal((nf, T, n)) + 1j * rng.standard_normal((nf, T, n))).astype(np.complex64)
            for k, n, _ in KINDS}


def _thetad():
    return np.array([TH["thetamin"] + i * TH["thetadelta"] for i in range(TH["theta_n"])])


def _oracle_omega(L, recs):
    nf, T = FMAX - FMIN + 1, TH["theta_n"]
    th = _thetad()
    out = {20: np.zeros((T, nf, 20), np.complex128), 10: np.zeros((T, nf, 10), np.complex128),
           2: np.zeros((T, nf, 10), np.complex128)}
    for k, n, npsv in KINDS:
        dst = out[{200: 20, 100: 20, 20: 10, 10: 10, 2: 2, 1: 2}[k]]
        rec = np.ascontiguousarray(recs[k])
        L.orc_km_rdsgt(k, n, npsv, FMIN, FMAX, T, ptr(th), TLEN, CRED, ptr(rec), ptr(dst))
    return out[20], out[10], out[2]  # tsgt(20,f,theta), rsgt(10,...), synn(10,...) as [theta][f][c]


def _ref_read(R, tmp, recs, compo=0, mt=np.array([1.0, -0.4, 0.3, 0.2, -0.7, 0.5])):
    for k, n, _ in KINDS:
        for f in range(FMIN, FMAX + 1):
            path = tmp / _name(k, f)
            path.parent.mkdir(parents=True, exist_ok=True)
            recs[k][f - FMIN].tofile(path)
    R.km_ref_init(TH["theta_n"], TH["thetamin"], TH["thetadelta"], FMIN, FMAX, TLEN, CRED, IPD, ptr(mt), compo,
                  str(tmp).encode(), b"m4")
    for k, n, npsv in KINDS:
        R.km_ref_read(RX if k not in (200, 100) else RS, RX if k in (200, 100) else 0.0, n, npsv, k)


@needs_ref
def test_rdsgtomega_matches_reference(tmp_path):
    L, R = _orc(), _ref()
    recs = _records(np.random.default_rng(7))
    _ref_read(R, tmp_path, recs)
    mine = _oracle_omega(L, recs)
    nf, T = FMAX - FMIN + 1, TH["theta_n"]
    for kind, (n, want) in enumerate(((20, mine[0]), (10, mine[1]), (10, mine[2]))):
        got = np.zeros(n * nf * T, np.complex128)
        R.km_ref_omega(kind, ptr(got))
        got = got.reshape(T, nf, n)
        assert np.all(got[:, 0, :] == 0)  # i = 0 is never read
        assert np.array_equal(got, want), kind
        # slots the reference leaves empty stay empty
        assert np.array_equal(got == 0, want == 0)


def _trig(rng):
    a = rng.uniform(0, 2 * np.pi, 3)  # phirq, phisq, phiqs
    return np.array([np.cos(a[0]), np.sin(a[0]), np.cos(2 * a[0]), np.sin(2 * a[0]),
                     np.cos(a[1]), np.sin(a[1]), np.cos(2 * a[1]), np.sin(2 * a[1]),
                     np.cos(a[2]), np.sin(a[2]), np.cos(2 * a[2]), np.sin(2 * a[2])])


def oracle_point(L, omega, dr, ds, trig, compo, mt):
    """clsgt + rsgt2h3freq + tsgt2h4freq of the oracle: h3, h4 as [f][6]."""
    tsg, rsg = omega
    nf, T = FMAX - FMIN + 1, TH["theta_n"]
    th = _thetad()
    rF = np.zeros((nf, 10), np.complex128)
    tF = np.zeros((nf, 20), np.complex128)
    L.orc_km_clsgt(dr, 10, FMIN, FMAX, T, ptr(th), TH["thetadelta"], IPD, TLEN, CRED, ptr(rsg), ptr(rF))
    L.orc_km_clsgt(ds, 20, FMIN, FMAX, T, ptr(th), TH["thetadelta"], IPD, TLEN, CRED, ptr(tsg), ptr(tF))
    h3 = np.zeros((nf, 6), np.complex128)
    h4 = np.zeros((nf, 6), np.complex128)
    t1 = np.ascontiguousarray(trig[:4])
    t2 = np.ascontiguousarray(np.concatenate([trig[8:10], trig[4:6], trig[10:12], trig[6:8]]))
    L.orc_km_rsgt2h3(compo, FMIN, FMAX, ptr(rF), ptr(t1), ptr(h3))
    L.orc_km_tsgt2h4(FMIN, FMAX, ptr(tF), ptr(t2), ptr(np.ascontiguousarray(mt)), ptr(h4))
    return h3, h4


@needs_ref
@pytest.mark.parametrize("compo", [0, 1, 2])
def test_points_match_reference(tmp_path, compo):
    L, R = _orc(), _ref()
    rng = np.random.default_rng(11 + compo)
    recs = _records(rng)
    mt = rng.standard_normal(6)
    _ref_read(R, tmp_path, recs, compo, mt)
    tsg, rsg, _ = _oracle_omega(L, recs)
    nf = FMAX - FMIN + 1
    for _ in range(6):
        dr, ds = rng.uniform(0.0, 30.0, 2)
        trig = _trig(rng)
        h3r = np.zeros((nf, 6), np.complex128)
        h4r = np.zeros((nf, 6), np.complex128)
        R.km_ref_point(dr, ds, ptr(trig), ptr(h3r), ptr(h4r))
        h3, h4 = oracle_point(L, (tsg, rsg), dr, ds, trig, compo, mt)
        assert np.abs(h3 - h3r).max() <= 1e-13 * np.abs(h3r).max()
        assert np.abs(h4 - h4r).max() <= 1e-13 * np.abs(h4r).max()


def test_sinc_interpolation_closed_forms():
    """No reduction phase (c_red_reci = 0 in the oracle call): a catalog
    constant in distance is returned exactly; at a catalog distance the
    grid value is returned (the other weights sin(pi n) / (pi n) vanish to
    round-off)."""
    L = _orc()
    nf, T = FMAX - FMIN + 1, TH["theta_n"]
    th = _thetad()
    rng = np.random.default_rng(3)
    const = rng.standard_normal((nf, 10)) + 1j * rng.standard_normal((nf, 10))
    omega = np.ascontiguousarray(np.broadcast_to(const, (T, nf, 10)))
    for d in (0.0, 3.3, 17.25, 29.9):
        out = np.zeros((nf, 10), np.complex128)
        L.orc_km_clsgt(d, 10, FMIN, FMAX, T, ptr(th), TH["thetadelta"], IPD, TLEN, 0.0, ptr(omega), ptr(out))
        assert np.abs(out - const).max() <= 1e-13 * np.abs(const).max()
    omega = rng.standard_normal((T, nf, 10)) + 1j * rng.standard_normal((T, nf, 10))
    for k in (0, 12, 40, 60):
        out = np.zeros((nf, 10), np.complex128)
        L.orc_km_clsgt(th[k], 10, FMIN, FMAX, T, ptr(th), TH["thetadelta"], IPD, TLEN, 0.0, ptr(omega), ptr(out))
        assert np.abs(out - omega[k]).max() <= 1e-14 * np.abs(omega).max() * 20


def test_cdft_is_the_inverse_dft():
    L = _orc()
    rng = np.random.default_rng(5)
    for n in (4, 16, 256, 2048):
        c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        a = np.zeros(2 * n)
        cc = c.copy()
        L.orc_km_cdft(2 * n, np.cos(np.pi / n), np.sin(np.pi / n), ptr(cc), ptr(a))
        assert np.abs(cc - np.fft.ifft(c) * n).max() <= 1e-12 * np.abs(c).sum() / np.sqrt(n)


def test_vectorfft_closed_form():
    """One line k0 of the spectrum: c(k0) = a gives, with its Hermitian
    image, x_n = 2 Re(a e^{2 pi i k0 n / N}) exp(omegai t_n) / tlen 1e3."""
    L = _orc()
    np1, tlen, omegai = 256, 1638.4, 0.003
    N = 2 * np1
    k0, a = 17, 0.7 - 0.2j
    spec = np.zeros(40, np.complex128)
    spec[k0] = a
    iws, iwe = 0, N - 1
    out = np.zeros(iwe - iws + 1)
    cv = np.zeros(N, np.complex128)
    work = np.zeros(2 * N)
    L.orc_km_vectorfft(0, 39, np1, ptr(spec), ptr(out), omegai, tlen, iws, iwe, ptr(cv), ptr(work))
    n = np.arange(N)
    t = n * tlen / N
    want = 2 * np.real(a * np.exp(2j * np.pi * k0 * n / N)) * np.exp(omegai * t) / tlen * 1e3
    assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max()


@needs_ref
def test_vectorfft_matches_reference():
    L, R = _orc(), _ref()
    rng = np.random.default_rng(9)
    for np1, imin, imax, iws, iwe in ((256, 0, 120, 0, 511), (1024, 1, 700, 100, 900)):
        spec = rng.standard_normal(imax - imin + 1) + 1j * rng.standard_normal(imax - imin + 1)
        out = np.zeros(iwe - iws + 1)
        ref = np.zeros(iwe - iws + 1)
        cv = np.zeros(2 * np1, np.complex128)
        work = np.zeros(4 * np1)
        L.orc_km_vectorfft(imin, imax, np1, ptr(spec), ptr(out), 0.002, 1638.4, iws, iwe, ptr(cv), ptr(work))
        R.km_ref_vectorfft(imin, imax, np1, ptr(spec), ptr(ref), 0.002, 1638.4, iws, iwe)
        assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()
