"""Writes tests/golden/psv_small_ref.npz: SGTpsv accumulators computed with
the reference's own numeric routines (src/SGTpsv/{calmat,dcsymbdl,dcsymbdl3,
glu2,rk3,trial,others}.f90, compiled from their sources by oracle/Makefile
into oracle/_ref/libref_psv.so) called in the reference driver's loop order
(oracle/psv_driver.inc, use_ref=1).  Run in the container that has the
reference sources; the fixture travels, the reference does not.

Cases: a surface source seen from receivers in the inner core, the fluid
outer core, the mantle and the crust; and a source at 400 km depth.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import dsm_inputs as di  # noqa: E402
import oracle_lib as ol  # noqa: E402

CASES = {
    "surf": (dict(tlen=1638.4, imax=64, rmin_=1000., rmax_=6371., rdelta_=250., thetamin=0.,
                  thetamax=180., thetadelta=45., maxlmax=1200, r0=6371.), (2, 40)),
    "deep": (dict(tlen=1638.4, imax=64, rmin_=3000., rmax_=6371., rdelta_=337., thetamin=0.,
                  thetamax=180., thetadelta=45., maxlmax=1200, r0=5971.), (3, 40)),
}


def main():
    if not ol.have_ref("psv"):
        raise SystemExit("oracle/_ref/libref_psv.so missing: run make -C oracle ref")
    out = {}
    for name, (kw, freqs) in CASES.items():
        inp = di.synthetic_input(**kw)
        for i in freqs:
            t, r, s, llog = ol.psv_omega(inp, i, use_ref=True)
            out[f"{name}_tsgt_{i}"], out[f"{name}_rsgt_{i}"], out[f"{name}_synn_{i}"] = t, r, s
            out[f"{name}_llog_{i}"] = np.array(llog)
    np.savez_compressed(os.path.join(HERE, "psv_small_ref.npz"), **out)


if __name__ == "__main__":
    main()
