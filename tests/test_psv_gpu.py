"""GPU parity of the P-SV engine (HIP path through the C-ABI) against the
CPU oracle, on the FP64 accumulators before the complex64 cast.

Bar: relative L2 <= 1e-9 per catalog component (components that are
round-off zeros of the reference are compared against the largest
component norm; see parity.py), and the same l cut-off as the reference's
calamp sequence.
"""
import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol
from parity import component_errors

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _engine(inp):
    import dsm_sgt
    return dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=dsm_sgt.PSV)


def _check(eng, inp, freqs, tol=TOL, fast=False):
    worst = 0.0
    for i in freqs:
        t, r, s, llog = ol.psv_omega(inp, i, fast=fast)
        gt, gr, gs = eng.fetch_all_fp64(i)
        assert eng.lcut(i) == llog, (i, eng.lcut(i), llog)
        for name, x, y in (("tsgt", gt, t), ("rsgt", gr, r), ("synn", gs, s)):
            e = component_errors(x, y)
            print(f"i={i} {name} max err {e.max():.2e}", flush=True)
            worst = max(worst, e.max())
            assert e.max() <= tol, (i, name, e)
    return worst


# receivers in the inner core, the fluid outer core, the mantle and the
# crust; surface source; poles included
SMALL = dict(tlen=1638.4, imax=16, rmin_=1000., rmax_=6371., rdelta_=250., thetamin=0.,
             thetamax=180., thetadelta=22.5, maxlmax=1200)


def test_psv_small_all_frequencies():
    inp = di.synthetic_input(**SMALL)
    eng = _engine(inp)
    info = ol.psv_info(inp)
    assert (eng.info.r_n, eng.info.theta_n, eng.info.nn) == (info.r_n, info.theta_n, info.nn)
    assert (eng.info.num_tsgt, eng.info.num_rsgt, eng.info.num_synn) == (20, 10, 10)
    eng.run(0, SMALL["imax"])
    _check(eng, inp, range(1, SMALL["imax"] + 1))


def test_psv_deep_source_cutoff():
    """A 400 km deep source: the l series is cut early, differently per
    frequency; the GPU replays the reference's calamp sequence."""
    p = dict(tlen=1638.4, imax=64, rmin_=3000., rmax_=6371., rdelta_=337., thetamin=0.,
             thetamax=180., thetadelta=30., maxlmax=1200, r0=5971.0)
    inp = di.synthetic_input(**p)
    eng = _engine(inp)
    eng.run(1, 64)
    cuts = [eng.lcut(i) for i in range(1, 65)]
    assert max(cuts) < p["maxlmax"] - 1
    _check(eng, inp, (1, 3, 40, 64))


def test_psv_5s_resolution():
    """The 5 s P-SV configuration's radial grid (3727 band rows, lmax 8000)
    on six receivers (one in the fluid core) and seven distances, at the
    lowest and highest frequencies."""
    p = dict(tlen=3276.8, imax=655, rmin_=3421., rmax_=6371., rdelta_=590., thetamin=0.,
             thetamax=180., thetadelta=30., maxlmax=8000)
    inp = di.synthetic_input(**p)
    eng = _engine(inp)
    assert eng.info.nn == 3727 and eng.info.r_n == 6
    eng.run(1, 1)
    eng.run(655, 655)
    _check(eng, inp, (1, 655), fast=True)


def test_psv_catalog_files(tmp_path):
    """Files, names and byte layout of the P-SV catalog (SGTpsv-MPI.f90:
    1103-1149): complex64 records of 20 / 10 / 10 components per distance."""
    inp = di.synthetic_input(**dict(SMALL, imax=3))
    eng = _engine(inp)
    eng.run(1, 3)
    eng.write_catalog(2, str(tmp_path), "m4")
    rr = int(eng.radii[-1] * 1e3)
    f = tmp_path / "TSGT" / f"m4.{int(6371.0e3):07d}.{rr:07d}.{2:07d}.TSGT_PSV"
    data = np.fromfile(f, np.complex64).reshape(eng.info.theta_n, 20)
    assert np.array_equal(data, eng.fetch(2, 0, len(eng.radii) - 1))
    f = tmp_path / "RSGT" / f"m4.{rr:07d}.{2:07d}.RSGT_PSV"
    assert np.fromfile(f, np.complex64).size == eng.info.theta_n * 10
    f = tmp_path / "RSGT" / f"m4.{int(6371.0e3):07d}.{2:07d}.SYNN_PSV"
    assert np.fromfile(f, np.complex64).size == eng.info.theta_n * 10
