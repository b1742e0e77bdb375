"""GPU parity of the SH engine (HIP path through the C-ABI) against the CPU
oracle, on the FP64 accumulators before the complex64 cast.

Bar: relative L2 <= 1e-9 per catalog component (components that are
round-off zeros of the reference, e.g. the strain trace of toroidal motion,
are compared against the largest component norm; see parity.py), and the
same l cut-off as the reference's calamp sequence for every frequency.
"""
import os
import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol
from parity import component_errors

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _engine(inp):
    import dsm_sgt
    return dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=dsm_sgt.SH)


def _check(eng, inp, freqs, tol=TOL):
    worst = 0.0
    for i in freqs:
        t, r, s, llog = ol.sh_omega(inp, i)
        gt, gr, gs = eng.fetch_all_fp64(i)
        assert eng.lcut(i) == llog, (i, eng.lcut(i), llog)
        for name, x, y in (("tsgt", gt, t), ("rsgt", gr, r), ("synn", gs, s)):
            e = component_errors(x, y)
            print(f"i={i} {name} max err {e.max():.2e}", flush=True)
            worst = max(worst, e.max())
            assert e.max() <= tol, (i, name, e)
    return worst


# coarse grid (58 band rows), every frequency of a 32-step run, poles included
SMALL = dict(tlen=3276.8, imax=32, rmin_=4421., rmax_=6371., rdelta_=250., thetamin=0.,
             thetamax=180., thetadelta=15., maxlmax=600)


def test_sh_small_all_frequencies():
    inp = di.synthetic_input(**SMALL)
    eng = _engine(inp)
    info = ol.sh_info(inp)
    assert (eng.info.r_n, eng.info.theta_n, eng.info.nn) == (info.r_n, info.theta_n, info.nn)
    eng.run(0, SMALL["imax"])
    _check(eng, inp, range(1, SMALL["imax"] + 1))


def test_sh_c2_resolution():
    """The 10 s SH configuration's radial grid (531 rows, 40 receivers,
    lmax 4000) on a coarse distance grid, lowest and highest frequencies."""
    p = dict(tlen=3276.8, imax=328, rmin_=4421., rmax_=6371., rdelta_=50., thetamin=0.,
             thetamax=180., thetadelta=20., maxlmax=4000, imin=0)
    inp = di.synthetic_input(**p)
    eng = _engine(inp)
    assert eng.info.nn == 531 and eng.info.r_n == 40
    eng.run(1, 2)
    eng.run(327, 328)
    _check(eng, inp, (1, 2, 327, 328))


def test_sh_deep_source_cutoff():
    """A 400 km deep source: the surface amplitude decays with l, so the
    reference's l cut-off ends the series early (and differently per
    frequency); the GPU must replay it exactly."""
    p = dict(tlen=1638.4, imax=40, rmin_=5671., rmax_=6371., rdelta_=100., thetamin=0.,
             thetamax=180., thetadelta=10., maxlmax=3000, r0=5971.0)
    inp = di.synthetic_input(**p)
    eng = _engine(inp)
    eng.run(1, 40)
    cuts = [eng.lcut(i) for i in range(1, 41)]
    assert max(cuts) < p["maxlmax"] - 1  # the cut-off fired everywhere
    _check(eng, inp, (1, 13, 40))


def test_sh_catalog_files(tmp_path):
    """Catalog files: the reference's names and byte layout (raw complex64,
    component fastest), equal to the complex64 cast of the FP64 record."""
    inp = di.synthetic_input(**SMALL)
    eng = _engine(inp)
    eng.run(3, 3)
    eng.write_catalog(3, str(tmp_path), "model4")
    nt = eng.info.theta_n
    intir0 = eng.info.intir0
    for ir, r in enumerate(eng.radii):
        f = tmp_path / "TSGT" / f"model4.{intir0:07d}.{int(r * 1e3):07d}.{3:07d}.TSGT_SH"
        data = np.fromfile(f, dtype=np.complex64)
        assert data.size == 10 * nt
        ref = eng.fetch_fp64(3, 0, ir).astype(np.complex64).ravel()
        assert np.array_equal(data, ref)
        g = tmp_path / "RSGT" / f"model4.{int(r * 1e3):07d}.{3:07d}.RSGT_SH"
        assert os.path.getsize(g) == 8 * 5 * nt
    s = tmp_path / "RSGT" / f"model4.{intir0:07d}.{3:07d}.SYNN_SH"
    assert os.path.getsize(s) == 8 * 5 * nt


def test_sh_receivers_at_and_below_the_cmb():
    """Receivers from 3030 km every 450 km: 3030 km lies in the core and
    3480 km exactly on the CMB; SGTsh drops both (SGTsh-MPI.f90:152-162:
    r <= rmin of the remaining solid model), keeps 3930..6180 km, and the
    catalog matches the oracle's on the trimmed set."""
    import dsm_sgt
    p = dict(tlen=3276.8, imax=8, rmin_=3030., rmax_=6371., rdelta_=450., thetamin=0., thetamax=180.,
             thetadelta=30., maxlmax=600)
    inp = di.synthetic_input(**p)
    eng = _engine(inp)
    info = ol.sh_info(inp)
    assert eng.info.r_n == info.r_n == 6
    assert np.allclose(eng.radii, 3930.0 + 450.0 * np.arange(6))
    eng.run(1, p["imax"])
    _check(eng, inp, (1, 4, 8))
