"""GPU: KernelMaker's catalog stage (dsm_km_points) and frequency-to-time
transform (dsm_km_vectorfft) against the oracle.

A P-SV and an SH engine compute a small catalog (3 radii above the CMB,
distances 0..30 degrees every 0.5 degree, frequencies 1..24); the oracle
(oracle/km_oracle.c, itself bit-identical to the reference's rdsgtomega
and within 1e-13 of its clsgt / rsgt2h3freq / tsgt2h4freq, see
test_oracle_km.py) reads the records the engines return through
dsm_sgt_fetch -- the complex(4) values the catalog files hold -- and the
device stage reads the engines' device catalogs.  Bar: the largest
difference below 1e-12 of the largest value, for every receiver component,
with and without the SH engine.  The transform (oracle: bit-identical to
the reference's vectorFFT_double) at np1 = 256, 4096 and 8192 (the 5 s and
2.5 s catalogs' FFT lengths), with and without i = 0, full and partial
windows: within 1e-12 of the largest sample (a different FFT factorisation
rounds differently).
"""
import numpy as np
import pytest

import dsm_inputs as di
from test_oracle_km import _orc, ptr

pytestmark = pytest.mark.gpu

GRID = dict(tlen=1638.4, imax=24, rmin_=5421., rmax_=6371., rdelta_=475., thetamin=0., thetamax=30.,
            thetadelta=0.5, maxlmax=1200, r0=6000.0)
FMIN, FMAX, CRED, IPD = 0, 24, 0.08, 4.0


@pytest.fixture(scope="module")
def engines():
    import dsm_sgt
    inp = di.synthetic_input(**GRID)
    psv = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=dsm_sgt.PSV)
    sh = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=dsm_sgt.SH)
    psv.run(0, GRID["imax"])
    sh.run(0, GRID["imax"])
    yield psv, sh
    psv.close()
    sh.close()


def _omega(L, psv, sh, ir):
    """The oracle's rdsgtomega of the engines' records at receiver ir."""
    import dsm_sgt
    nf, T = FMAX - FMIN + 1, psv.info.theta_n
    th = np.ascontiguousarray(psv.theta)
    om_t = np.zeros((T, nf, 20), np.complex128)
    om_r = np.zeros((T, nf, 10), np.complex128)
    kinds = [(200, dsm_sgt.TSGT, 20, psv, om_t), (20, dsm_sgt.RSGT, 10, psv, om_r)]
    if sh is not None:
        kinds += [(100, dsm_sgt.TSGT, 10, sh, om_t), (10, dsm_sgt.RSGT, 5, sh, om_r)]
    for ipsvorsh, kind, num, eng, dst in kinds:
        rec = np.zeros((nf, T, num), np.complex64)
        for i in range(max(FMIN, 1), FMAX + 1):
            rec[i - FMIN] = eng.fetch(i, kind, ir)
        L.orc_km_rdsgt(ipsvorsh, num, dst.shape[2], FMIN, FMAX, T, ptr(th), GRID["tlen"], CRED, ptr(rec), ptr(dst))
    return om_t, om_r


def _oracle_points(L, om, theta, compo, mt, dr, ds, trig):
    om_t, om_r = om
    nf, T = FMAX - FMIN + 1, theta.size
    h3 = np.zeros((dr.size, nf, 6), np.complex128)
    h4 = np.zeros_like(h3)
    for p in range(dr.size):
        rF = np.zeros((nf, 10), np.complex128)
        tF = np.zeros((nf, 20), np.complex128)
        L.orc_km_clsgt(dr[p], 10, FMIN, FMAX, T, ptr(theta), GRID["thetadelta"], IPD, GRID["tlen"], CRED, ptr(om_r),
                       ptr(rF))
        L.orc_km_clsgt(ds[p], 20, FMIN, FMAX, T, ptr(theta), GRID["thetadelta"], IPD, GRID["tlen"], CRED, ptr(om_t),
                       ptr(tF))
        t1 = np.ascontiguousarray(trig[p, :4])
        t2 = np.ascontiguousarray(np.concatenate([trig[p, 8:10], trig[p, 4:6], trig[p, 10:12], trig[p, 6:8]]))
        a3 = np.zeros((nf, 6), np.complex128)
        a4 = np.zeros((nf, 6), np.complex128)
        L.orc_km_rsgt2h3(compo, FMIN, FMAX, ptr(rF), ptr(t1), ptr(a3))
        L.orc_km_tsgt2h4(FMIN, FMAX, ptr(tF), ptr(t2), ptr(np.ascontiguousarray(mt)), ptr(a4))
        h3[p], h4[p] = a3, a4
    return h3, h4


def _points(rng, n):
    dr, ds = rng.uniform(0.0, 30.0, n), rng.uniform(0.0, 30.0, n)
    a = rng.uniform(0.0, 2 * np.pi, (n, 3))
    trig = np.stack([np.cos(a[:, 0]), np.sin(a[:, 0]), np.cos(2 * a[:, 0]), np.sin(2 * a[:, 0]),
                     np.cos(a[:, 1]), np.sin(a[:, 1]), np.cos(2 * a[:, 1]), np.sin(2 * a[:, 1]),
                     np.cos(a[:, 2]), np.sin(a[:, 2]), np.cos(2 * a[:, 2]), np.sin(2 * a[:, 2])], axis=1)
    # a point exactly on a catalog distance, and the window at both ends
    dr[:3] = (10.0, 0.0, 30.0)
    ds[:3] = (0.25, 30.0, 12.5)
    return dr, ds, trig


@pytest.mark.parametrize("compo", [0, 1, 2])
@pytest.mark.parametrize("with_sh", [True, False])
def test_km_points_vs_oracle(engines, compo, with_sh):
    import dsm_sgt
    psv, sh = engines
    L = _orc()
    rng = np.random.default_rng(100 + compo + 3 * with_sh)
    mt = rng.standard_normal(6)
    dr, ds, trig = _points(rng, 40)
    for ir in (0, psv.info.r_n - 1):
        om = _omega(L, psv, sh if with_sh else None, ir)
        w3, w4 = _oracle_points(L, om, np.ascontiguousarray(psv.theta), compo, mt, dr, ds, trig)
        g3, g4 = dsm_sgt.km_points(psv, sh if with_sh else None, ir, FMIN, FMAX, compo, CRED, IPD, mt, dr, ds, trig)
        e3 = np.abs(g3 - w3).max() / np.abs(w3).max()
        e4 = np.abs(g4 - w4).max() / np.abs(w4).max()
        print(f"compo {compo} sh {with_sh} ir {ir}: h3 {e3:.1e} h4 {e4:.1e}")
        assert e3 <= 1e-12 and e4 <= 1e-12
        assert np.all(g3[:, 0] == 0) and np.all(g4[:, 0] == 0)  # i = 0 is never read


def test_km_points_refuses_missing_frequencies(engines):
    import dsm_sgt
    psv, sh = engines
    dr, ds, trig = _points(np.random.default_rng(1), 4)
    with pytest.raises(dsm_sgt.DsmSgtError):
        dsm_sgt.km_points(psv, sh, 0, 0, GRID["imax"] + 1, 0, CRED, IPD, np.ones(6), dr, ds, trig)


@pytest.mark.parametrize("np1,imin,imax,iws,iwe", [(256, 0, 200, 0, 511), (4096, 1, 656, 100, 8000),
                                                   (8192, 0, 1311, 0, 16383), (8192, 5, 9000, 777, 4242)])
def test_vectorfft_vs_oracle(engines, np1, imin, imax, iws, iwe):
    import dsm_sgt
    psv, _ = engines
    L = _orc()
    rng = np.random.default_rng(np1 + imin)
    nvec = 5
    spec = rng.standard_normal((nvec, imax - imin + 1)) + 1j * rng.standard_normal((nvec, imax - imin + 1))
    spec *= np.exp(-np.arange(imax - imin + 1) / 400.0)  # a decaying spectrum, as the catalogs'
    omegai, tlen = np.log(50.0) / 3276.8, 3276.8
    got = dsm_sgt.km_vectorfft(psv, spec, imin, np1, omegai, tlen, iws, iwe)
    cv = np.zeros(2 * np1, np.complex128)
    work = np.zeros(4 * np1)
    for v in range(nvec):
        want = np.zeros(iwe - iws + 1)
        sv = np.ascontiguousarray(spec[v])
        L.orc_km_vectorfft(imin, imax, np1, ptr(sv), ptr(want), omegai, tlen, iws, iwe, ptr(cv), ptr(work))
        err = np.abs(got[v] - want).max() / np.abs(want).max()
        assert err <= 1e-12, (v, err)
    print(f"np1 {np1}: max relative difference {err:.1e}")
