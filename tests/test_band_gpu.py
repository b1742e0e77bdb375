"""GPU: the per-frequency P-SV band pieces the device assembles
(psv_band.hip, through dsm_sgt_psv_band) equal, value for value, those of
the host set-up (psv_omega_setup + psv_pack_band, through the test-only
host build libdsmsgt_hostcheck.so) -- solid and fluid zones, zone
boundaries (calbc), the frequency 0 and the highest frequency.
"""
import ctypes as C
import os

import numpy as np
import pytest

import dsm_inputs as di

pytestmark = pytest.mark.gpu

HOSTCHECK = os.path.join(di.ROOT, "dsm-kernel_amd", "libdsmsgt_hostcheck.so")

GRIDS = {
    # the 5 s catalog: inner core, fluid outer core, mantle, crust
    "5s": (dict(tlen=3276.8, imax=656, rmin_=3716., rmax_=6371., rdelta_=45., thetamin=0., thetamax=180.,
                thetadelta=0.1, maxlmax=8000), (0, 1, 2, 300, 655, 656)),
    # coarse grid, receivers down to 1000 km radius
    "small": (dict(tlen=1638.4, imax=16, rmin_=1000., rmax_=6371., rdelta_=250., thetamin=0., thetamax=180.,
                   thetadelta=22.5, maxlmax=1200), (0, 1, 7, 16)),
}


def _host_band(inp, i, nn):
    L = C.CDLL(HOSTCHECK)
    L.dsm_hostcheck_psv_band.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    out = np.zeros((nn, 4, 4), np.complex128)
    assert L.dsm_hostcheck_psv_band(C.byref(inp), i, out.ctypes.data) == 0
    return out


@pytest.mark.parametrize("grid", sorted(GRIDS))
def test_device_band_equals_host(grid):
    import dsm_sgt
    p, freqs = GRIDS[grid]
    inp = di.synthetic_input(**p)
    eng = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6)
    nn = eng.info.nn
    for i in freqs:
        dev = eng.psv_band(i)
        host = _host_band(inp, i, nn)
        assert np.all(np.isfinite(dev)), i
        bad = np.argwhere(dev != host)
        assert bad.size == 0, (i, bad[:5], dev[tuple(bad[0])], host[tuple(bad[0])])
        print(f"{grid} i={i}: {nn} columns equal, max |entry| {np.abs(host).max():.3e}", flush=True)
    eng.close()
