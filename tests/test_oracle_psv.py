"""CPU: the P-SV oracle against the reference's own routines and against
properties the mathematics fixes.

The oracle (oracle/psv_oracle.c + psv_driver.inc) restates the SGTpsv
driver loop and every routine it calls, with a routine table that can be
switched to the reference's routines compiled from src/SGTpsv/*.f90
(oracle/_ref).  The committed fixture tests/golden/psv_small_ref.npz holds
the reference-routine results, so the check also runs where the reference
sources are absent.
"""
import os

import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol
from parity import component_errors

GOLD = os.path.join(os.path.dirname(__file__), "golden", "psv_small_ref.npz")
CASES = {
    "surf": (dict(tlen=1638.4, imax=64, rmin_=1000., rmax_=6371., rdelta_=250., thetamin=0.,
                  thetamax=180., thetadelta=45., maxlmax=1200, r0=6371.), (2, 40)),
    "deep": (dict(tlen=1638.4, imax=64, rmin_=3000., rmax_=6371., rdelta_=337., thetamin=0.,
                  thetamax=180., thetadelta=45., maxlmax=1200, r0=5971.), (3, 40)),
}
# flang's complex division (scaled textbook formula) against gcc's Smith
# method in the oracle: the measured spread is 5e-10 at frequency 2 on this
# coarse grid and <1e-12 elsewhere (same mechanism as the SH fixture).
TOL_REF = 5e-9


def test_psv_oracle_matches_reference_golden():
    g = np.load(GOLD)
    for name, (kw, freqs) in CASES.items():
        inp = di.synthetic_input(**kw)
        for i in freqs:
            t, r, s, llog = ol.psv_omega(inp, i)
            assert llog == int(g[f"{name}_llog_{i}"])
            for comp, x in (("tsgt", t), ("rsgt", r), ("synn", s)):
                e = component_errors(x, g[f"{name}_{comp}_{i}"])
                assert e.max() <= TOL_REF, (name, i, comp, e)


@pytest.mark.skipif(not ol.have_ref("psv"), reason="reference routines not built (oracle/_ref)")
def test_psv_oracle_matches_reference_routines_fluid_receivers():
    """Receivers across the inner core, the fluid outer core (calupfluid
    branch) and the lowermost mantle, source at 371 km depth."""
    inp = di.synthetic_input(tlen=1638.4, imax=64, rmin_=1100., rmax_=3600., rdelta_=100.,
                             thetamin=0., thetamax=180., thetadelta=30., maxlmax=1200, r0=6000.)
    a = ol.psv_omega(inp, 40)
    b = ol.psv_omega(inp, 40, use_ref=True)
    assert a[3] == b[3]
    for k in range(3):
        assert component_errors(a[k], b[k]).max() <= 1e-10


def test_psv_fluid_strain_is_isotropic():
    """In the fluid outer core the reference's strain is the isotropic
    omega*Q/(3 lambda) (calupfluid): off-diagonal and deviatoric slots
    vanish, t3 = -t1 for the rr source."""
    kw, _ = CASES["surf"]
    inp = di.synthetic_input(**kw)
    t, r, s, _ = ol.psv_omega(inp, 40)
    radii = kw["rmin_"] + kw["rdelta_"] * np.arange(t.shape[1])
    fl = (radii > 1221.0) & (radii <= 3480.0)
    assert fl.sum() >= 8
    tf = t[:, fl, :]
    scale = np.abs(tf[..., 0]).max()
    assert scale > 0
    for slot in (5, 9, 12, 13, 14, 15, 16, 17, 18, 19):
        assert np.abs(tf[..., slot]).max() <= 1e-14 * scale, slot
    assert np.allclose(tf[..., 2], -tf[..., 0], rtol=1e-14, atol=0)


def test_psv_poles_keep_only_axisymmetric_terms():
    """On the source axis only the m=0 harmonics are non-zero
    (src/SGTpsv/trial.f90:112-131), so the sources that excite |m|=1,2
    (imt 4,5,6) and every slot built from u_rt, u_rp, u_tp or u_tt-u_pp
    vanish at 0 and 180 degrees."""
    kw, _ = CASES["surf"]
    inp = di.synthetic_input(**kw)
    t, r, s, _ = ol.psv_omega(inp, 40)
    top = np.abs(t).max()
    for it in (0, -1):
        for slot in (4, 5, 6, 7, 9, 10, 12, 13, 14, 15, 16, 17, 18, 19):
            assert np.abs(t[it, :, slot]).max() <= 1e-13 * top, (it, slot)
    assert np.abs(t[0, :, 0]).max() > 1e-3 * top
    assert np.abs(t[1, :, 12]).max() > 1e-6 * top


def test_psv_lcut_after_lsuf():
    """The l series of a deep source stops once the receiver-side amplitude
    has stayed below ratl for several l past l_suf (calamp, l >= lsuf)."""
    inp = di.synthetic_input(tlen=1638.4, imax=64, rmin_=5471., rmax_=6371., rdelta_=300.,
                             thetamin=0., thetamax=180., thetadelta=90., maxlmax=3000, r0=5971.)
    for i in (10, 40):
        cut = ol.psv_omega(inp, i)[3]
        omega = 2 * np.pi * i / 1638.4
        lsuf = int(omega * 6371.0 / 3.9 - 0.5) + 1
        assert lsuf + 3 <= cut < 2999, (i, cut, lsuf)
