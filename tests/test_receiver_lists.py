"""Receiver geometry given as lists, as DSM3D_forward reads it
(src/DSM3D_forward/mpi-tipsv.f90:163-210: a file of receiver radii and a
file of distances) instead of SGTpsv/SGTsh's uniform grids; the frequency
loop is the engines' (dsm_sgt_input_t.r_list / theta_list).

CPU: the oracle's list path against its grid path.  Lists holding the grid's
own values give the grid run bit for bit; lists of every other radius and
every third distance give exactly those rows and columns of the grid run
(receivers and distances are independent of one another: the solves do not
depend on them, each receiver's stencil depends on its own radius only);
SH's trim of radii at or below the CMB applies to a list.  Lists that are not
strictly ascending, or hold a non-finite value, are refused at create.

GPU: both engines on seeded random lists (radii in every region, the SH
list running into the core; distances including both poles) against the
oracle on the same lists: FP64 at 1e-9 with the same l cut-off, BF16X6 at
the small grids' bound (a receiver next to the surface source).
"""
import ctypes as C
import math

import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol
from parity import component_errors

BASE = dict(tlen=1638.4, imax=12, rmin_=3600., rmax_=6350., rdelta_=250., thetamin=0., thetamax=180.,
            thetadelta=15., maxlmax=800, r0=6200.0)


def _with_lists(inp, radii=None, thetas=None):
    """A copy of inp with the given lists (the arrays are kept on the copy)."""
    out = type(inp).from_buffer_copy(inp)
    out._keep = []
    if radii is not None:
        a = np.ascontiguousarray(radii, np.float64)
        out._keep.append(a)
        out.r_list_n, out.r_list = a.size, a.ctypes.data
    if thetas is not None:
        a = np.ascontiguousarray(thetas, np.float64)
        out._keep.append(a)
        out.theta_list_n, out.theta_list = a.size, a.ctypes.data
    return out


def _grid(p):
    r = p["rmin_"] + p["rdelta_"] * np.arange(int((p["rmax_"] - p["rmin_"]) / p["rdelta_"]) + 1)
    t = p["thetamin"] + p["thetadelta"] * np.arange(int((p["thetamax"] - p["thetamin"]) / p["thetadelta"]) + 1)
    return r, t


@pytest.mark.parametrize("mode", ["psv", "sh"])
def test_oracle_lists_equal_grid(mode):
    fn = ol.psv_omega if mode == "psv" else ol.sh_omega
    grid = di.synthetic_input(**BASE)
    r, t = _grid(BASE)
    same = _with_lists(grid, r, t)
    sub = _with_lists(grid, r[::2], t[::3])
    info = ol.psv_info if mode == "psv" else ol.sh_info
    keep = r.size - info(grid).r_n  # SH: radii at or below the CMB dropped from the front
    for i in (1, 7, 12):
        a, b, c = fn(grid, i, fast=True), fn(same, i, fast=True), fn(sub, i, fast=True)
        assert a[3] == b[3] == c[3]
        for k in range(3):
            assert np.array_equal(a[k], b[k]), (i, k)
        # the subset: distances every third, receivers every other (after the trim)
        rows = [j - keep for j in range(0, r.size, 2) if j >= keep]
        assert np.array_equal(c[0], a[0][::3][:, rows]), i
        assert np.array_equal(c[1], a[1][::3][:, rows]), i
        assert np.array_equal(c[2], a[2][::3]), i


@pytest.mark.parametrize("bad", ["descending", "nan", "negative_theta"])
def test_create_refuses_bad_lists(bad):
    import dsm_sgt
    inp = dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(**BASE))
    if bad == "descending":
        inp = _with_lists(inp, radii=[6000.0, 5000.0])
    elif bad == "nan":
        inp = _with_lists(inp, radii=[5000.0, math.nan])
    else:
        inp = _with_lists(inp, thetas=[-1.0, 10.0])
    with pytest.raises(dsm_sgt.DsmSgtError) as e:
        dsm_sgt.Engine(inp)
    assert e.value.code == 1 and "list" in str(e.value)


def _random_lists(seed, mode):
    rng = np.random.default_rng(seed)
    lo = 1500.0 if mode == "psv" else 3000.0  # P-SV: receivers in the inner core, the fluid, the mantle
    radii = np.unique(np.round(rng.uniform(lo, 6371.0, 9), 3))
    thetas = np.unique(np.concatenate([[0.0, 180.0], np.round(rng.uniform(0.0, 180.0, 9), 4)]))
    return radii, thetas


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["PSV", "SH"])
@pytest.mark.parametrize("prec,tol", [("FP64", 1e-9), ("BF16X6", None)])
def test_engine_lists_vs_oracle(mode, prec, tol):
    import dsm_sgt
    from test_mixed_gpu import SMALL_TOL
    radii, thetas = _random_lists(11 if mode == "PSV" else 12, mode.lower())
    inp = _with_lists(di.synthetic_input(**BASE), radii, thetas)
    fn = ol.psv_omega if mode == "PSV" else ol.sh_omega
    bound = tol if tol is not None else SMALL_TOL[mode.lower(), prec]
    eng = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=getattr(dsm_sgt, mode),
                         precision=getattr(dsm_sgt, prec))
    info = ol.psv_info(inp) if mode == "PSV" else ol.sh_info(inp)
    assert eng.info.r_n == info.r_n and eng.info.theta_n == thetas.size
    assert np.array_equal(eng.theta, thetas)
    assert np.array_equal(eng.radii, radii[radii.size - info.r_n:])
    eng.run(1, BASE["imax"])
    worst = 0.0
    for i in (1, 5, 12):
        ref = fn(inp, i, fast=True)
        got = eng.fetch_all_fp64(i)
        assert eng.lcut(i) == ref[3]
        for k, (x, y) in enumerate(zip(got, ref[:3])):
            e = component_errors(x, y)
            worst = max(worst, e.max())
            assert e.max() <= bound, (i, k, e)
    print(f"{mode} {prec} on lists ({info.r_n} radii, {thetas.size} distances): max rel L2 {worst:.2e}")
    eng.close()
