"""CPU: the device band assembly's code (psv_band_column, psv_band_lane.h,
run on the host by the test-only build libdsmsgt_hostcheck.so) equals the
host set-up's band pieces (psv_omega_setup + psv_pack_band) value for
value -- every column, solid and fluid zones, zone boundaries (calbc) --
over the frequencies of the 5 s catalog grid and a coarse grid.  The GPU
test test_band_gpu.py checks the kernel itself against the same host band.
"""
import ctypes as C
import os

import numpy as np
import pytest

import dsm_inputs as di

HOSTCHECK = os.path.join(di.ROOT, "dsm-kernel_amd", "libdsmsgt_hostcheck.so")

GRIDS = {
    "5s": (dict(tlen=3276.8, imax=656, rmin_=3716., rmax_=6371., rdelta_=45., thetamin=0., thetamax=180.,
                thetadelta=0.1, maxlmax=8000), list(range(0, 657, 41)) + [655, 656]),
    "small": (dict(tlen=1638.4, imax=16, rmin_=1000., rmax_=6371., rdelta_=250., thetamin=0., thetamax=180.,
                   thetadelta=22.5, maxlmax=1200), list(range(0, 17))),
}


def _lib():
    if not os.path.exists(HOSTCHECK):
        pytest.skip("libdsmsgt_hostcheck.so not built (make -C dsm-kernel_amd)")
    L = C.CDLL(HOSTCHECK)
    for f in ("dsm_hostcheck_psv_band", "dsm_hostcheck_psv_band_lane"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    return L


@pytest.mark.parametrize("grid", sorted(GRIDS))
def test_band_lane_equals_host_setup(grid):
    L = _lib()
    p, freqs = GRIDS[grid]
    inp = di.synthetic_input(**p)
    nn = None
    for i in freqs:
        host = np.zeros((8192, 4, 4), np.complex128)
        lane = np.zeros((8192, 4, 4), np.complex128)
        assert L.dsm_hostcheck_psv_band(C.byref(inp), i, host.ctypes.data) == 0
        assert L.dsm_hostcheck_psv_band_lane(C.byref(inp), i, lane.ctypes.data) == 0
        if nn is None:
            nz = np.nonzero(np.abs(host).reshape(8192, -1).max(axis=1))[0]
            nn = int(nz.max()) + 1
            assert nn < 8192
        assert np.all(np.isfinite(lane[:nn]))
        bad = np.argwhere(lane != host)
        assert bad.size == 0, (i, bad[:5])
