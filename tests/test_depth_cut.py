"""CPU: the reference's depth truncation (calcutd with ratc) against the
depth cut of the batched forms.

calcutd (src/SGTpsv/others.f90:1421-1499, src/SGTsh/others.f90:928-982) is
defined but never called by SGTpsv / SGTsh, and ratc is read from
DSM.config and then unused.  The oracle restates both versions with their
quirks (P-SV: the last node above the threshold, the amplitudes of the
rows at odd offsets from kkdr; SH: the real part, the first node above the
threshold) and they are pinned against the reference's compiled routines.

The batched forms cut each (omega, l) system from an a-priori decay
estimate (psv_depth_kernel, restated on the host by the product's test
build).  Following the original-DSM semantics of calcutd (SH's: the first
node from the bottom whose |U| exceeds ratc times the largest), the full
solution of real systems on the 5 s grid (3733 rows) must keep every row
calcutd keeps: the cut row is at or below calcutd's, at ratc = 1e-10.
"""
import ctypes as C
import os

import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol

HC = os.path.join(di.ROOT, "dsm-kernel_amd", "libdsmsgt_hostcheck.so")
P = C.c_void_p


def _i(v):
    return C.byref(C.c_int(v))


def _d(v):
    return C.byref(C.c_double(v))


def _psv_calcutd(fn, nzone, nnl, iphase, kkdr, ra, tmpc, rat, nn):
    kc = C.c_int(-1)
    nnl = np.ascontiguousarray(nnl, np.int32)
    iph = np.ascontiguousarray(iphase, np.int32)
    kk = np.ascontiguousarray(kkdr, np.int32)
    ra = np.ascontiguousarray(ra, np.float64)
    fn(_i(nzone), _i(int(nnl.sum())), nnl.ctypes.data_as(P), tmpc.ctypes.data_as(P), _d(rat), _i(nn),
       iph.ctypes.data_as(P), _d(0.5), _i(1), ra.ctypes.data_as(P), kk.ctypes.data_as(P), C.byref(kc))
    return kc.value


def _sh_calcutd(fn, nzone, nnl, tmpr, rat, nn, ra):
    kc = C.c_int(-1)
    nnl = np.ascontiguousarray(nnl, np.int32)
    ra = np.ascontiguousarray(ra, np.float64)
    fn(_i(nzone), nnl.ctypes.data_as(P), tmpr.ctypes.data_as(P), _d(rat), _i(nn), ra.ctypes.data_as(P), C.byref(kc))
    return kc.value


@pytest.mark.skipif(not (ol.have_ref("psv") and ol.have_ref("sh")), reason="reference routines not built")
def test_calcutd_restatements_match_reference():
    L = ol.lib()
    RP = C.CDLL(ol.ref_path("psv"), mode=os.RTLD_LAZY)  # (symbols of the uncalled I/O routines stay unresolved)
    RS = C.CDLL(ol.ref_path("sh"), mode=os.RTLD_LAZY)
    rng = np.random.default_rng(4)
    # zone layouts: (nnl, iphase); kkdr as calspdr builds it (2 rows per
    # solid node, 1 per fluid node, interface nodes of unlike zones shared)
    for nnl, iph in (([10], [1]), ([6, 5, 7], [1, 2, 1]), ([3, 8, 4, 6], [1, 2, 1, 1])):
        nz = len(nnl)
        kkdr = [1]
        for z in range(1, nz):
            rows = 2 * nnl[z - 1] if iph[z - 1] == 1 else nnl[z - 1]
            kkdr.append(kkdr[-1] + rows + (1 if iph[z - 1] == iph[z] and iph[z] == 1 else 0))
        nn = kkdr[-1] + (2 * nnl[-1] + 1 if iph[-1] == 1 else nnl[-1])
        nodes = sum(nnl) + nz + 1
        ra = np.sort(rng.uniform(1000.0, 6371.0, nodes))
        for rat in (1e-1, 1e-3, 1e-10):
            # P-SV: amplitudes falling upward, so that the last node above the
            # threshold lies below the top one (at the top the reference reads
            # an unassigned zone index)
            decay = np.exp(np.linspace(0, -40, nn))
            tmpc = (decay * (rng.standard_normal(nn) + 1j * rng.standard_normal(nn))).astype(np.complex128)
            a = _psv_calcutd(L.orc_psv_calcutd, nz, nnl, iph, kkdr, ra, tmpc, rat, nn)
            b = _psv_calcutd(RP.calcutd_, nz, nnl, iph, kkdr, ra, tmpc, rat, nn)
            assert a == b and a > 0, (nnl, rat, a, b)
            # SH: one value per node, rising upward
            tmpr = (np.exp(np.linspace(-40, 0, nodes)) * (1.0 + rng.standard_normal(nodes) + 1j)).astype(np.complex128)
            a = _sh_calcutd(L.orc_sh_calcutd, nz, nnl, tmpr, rat, nodes, ra)
            b = _sh_calcutd(RS.calcutd_, nz, nnl, tmpr, rat, nodes, ra)
            assert a == b and a > 0, (nnl, rat, a, b)


# (source radius, frequency, degree): a 400 km and a 100 km deep source, the
# 5 s grid, degrees below the reference's cut-off
CASES = [(5971.0, 100, 300), (5971.0, 600, 1500), (6271.0, 100, 300), (6271.0, 300, 1000), (6271.0, 600, 1500)]
RATC = 1e-10


@pytest.mark.skipif(not os.path.exists(HC), reason="build the test library first")
def test_depth_cut_keeps_the_rows_calcutd_keeps():
    H = C.CDLL(HC)
    H.dsm_hostcheck_psv_depth.argtypes = [P, C.c_int32, C.c_int32, C.c_double, C.c_double] + [P] * 5
    L = ol.lib(fast=True)
    L.orc_psv_capture.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, C.c_int32]
    rows = []
    for r0, i, l in CASES:
        inp = di.synthetic_input(tlen=3276.8, imax=656, rmin_=6371., rmax_=6371., rdelta_=10., thetamin=10.,
                                 thetamax=10., thetadelta=1., maxlmax=l, r0=r0)
        out = [C.c_int32() for _ in range(4)]
        nn = ol.psv_info(inp).nn
        urow = np.zeros(nn, np.int32)
        lb = max(1, 256 * (l // 256))  # the depth cut is taken per block of 256 degrees, at its lowest
        rc = H.dsm_hostcheck_psv_depth(C.byref(inp), i, lb, 40.0, 0.35, *[C.byref(o) for o in out],
                                       urow.ctypes.data_as(P))
        assert rc == 0
        first, row0, jtmp, nn2 = [o.value for o in out]
        assert nn2 == nn
        g = np.full(nn, np.nan, np.complex128)
        L.orc_psv_capture(l, 0, 1, g.ctypes.data_as(P), nn)
        try:
            ol.psv_omega(inp, i, fast=True)
        finally:
            L.orc_psv_capture(0, 0, 0, None, 0)
        assert not np.isnan(g.real).any(), "degree past the cut-off"
        u = np.where(urow == 1, np.abs(g), 0.0)
        jstar = int(np.argmax(u > RATC * u.max())) + 1  # first U row above the threshold
        rows.append((r0, i, l, nn - first + 1, nn - jstar + 1))
        assert first <= jstar, (r0, i, l, first, jstar)
    for r0, i, l, ours, cut in rows:
        print(f"source {6371 - r0:.0f} km, i={i}, l={l}: rows solved {ours}, calcutd(ratc=1e-10) keeps {cut}")
