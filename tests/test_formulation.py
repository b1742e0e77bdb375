"""CPU: the product's SH formulation (the HIP kernels' per-lane bodies,
the set-up code and the contraction table, built for the host as
libdsmsgt_hostcheck.so) against the oracle.  This pins the recast of the
reference's per-distance loops into coefficient rows x theta functions
without a GPU; the GPU tests then only have to pin the device execution.
"""
import ctypes as C
import os

import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol
from parity import component_errors

LIB = os.path.join(di.ROOT, "dsm-kernel_amd", "libdsmsgt_hostcheck.so")


def _hc():
    if not os.path.exists(LIB):
        pytest.skip("host check library not built")
    L = C.CDLL(LIB)
    L.dsm_hostcheck_sh.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def _run(L, inp, i):
    t, r, s, llog = ol.sh_omega(inp, i)
    gt, gr, gs = np.zeros_like(t), np.zeros_like(r), np.zeros_like(s)
    ll = C.c_int32()
    assert L.dsm_hostcheck_sh(C.byref(inp), i, gt.ctypes.data, gr.ctypes.data, gs.ctypes.data, C.byref(ll)) == 0
    assert ll.value == llog
    return max(component_errors(gt, t).max(), component_errors(gr, r).max(), component_errors(gs, s).max())


def test_formulation_small_grid():
    L = _hc()
    inp = di.synthetic_input(tlen=3276.8, imax=32, rmin_=4421., rmax_=6371., rdelta_=250., thetamin=0.,
                             thetamax=180., thetadelta=15., maxlmax=600)
    for i in (1, 2, 17, 32):
        assert _run(L, inp, i) <= 1e-12


def test_formulation_deep_source():
    L = _hc()
    inp = di.synthetic_input(tlen=1638.4, imax=40, rmin_=5671., rmax_=6371., rdelta_=100., thetamin=0.,
                             thetamax=180., thetadelta=20., maxlmax=3000, r0=5971.0)
    for i in (3, 40):
        assert _run(L, inp, i) <= 1e-12


# ---- P-SV ------------------------------------------------------------------
def _hc_psv():
    L = _hc()
    L.dsm_hostcheck_psv.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def _run_psv(L, inp, i):
    t, r, s, llog = ol.psv_omega(inp, i)
    gt, gr, gs = np.zeros_like(t), np.zeros_like(r), np.zeros_like(s)
    ll = C.c_int32()
    assert L.dsm_hostcheck_psv(C.byref(inp), i, gt.ctypes.data, gr.ctypes.data, gs.ctypes.data, C.byref(ll)) == 0
    assert ll.value == llog
    return max(component_errors(gt, t).max(), component_errors(gr, r).max(), component_errors(gs, s).max())


def test_psv_formulation_all_regions():
    """Six distinct solves per (omega,l), 63 coefficient rows per receiver
    and the two-pass contraction reproduce the reference's 15 solves and
    per-distance strain loops: receivers in the inner core, the fluid core
    (calupfluid), the mantle and the crust; the lowest frequencies are the
    worst conditioned (1e-11 there, 1e-14 elsewhere)."""
    L = _hc_psv()
    inp = di.synthetic_input(tlen=1638.4, imax=64, rmin_=1000., rmax_=6371., rdelta_=250., thetamin=0.,
                             thetamax=180., thetadelta=22.5, maxlmax=1200)
    for i in (1, 3, 40):
        assert _run_psv(L, inp, i) <= 1e-10


def test_psv_formulation_deep_source():
    L = _hc_psv()
    inp = di.synthetic_input(tlen=1638.4, imax=64, rmin_=3000., rmax_=6371., rdelta_=337., thetamin=0.,
                             thetamax=180., thetadelta=22.5, maxlmax=1200, r0=5971.0)
    for i in (3, 40):
        assert _run_psv(L, inp, i) <= 1e-12
