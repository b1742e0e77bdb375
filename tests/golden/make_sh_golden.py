"""Writes tests/golden/sh_small_ref.npz: SGTsh accumulators computed with
the reference's own numeric routines (src/SGTsh/{calmat,dclib,trial,others,
glu2}.f90, compiled from their sources by oracle/Makefile into
oracle/_ref/libref_sh.so) called in the reference driver's loop order
(oracle/sh_oracle.c, use_ref=1).  Run in the container that has the
reference sources; the fixture travels, the reference does not.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import dsm_inputs as di  # noqa: E402
import oracle_lib as ol  # noqa: E402

SMALL = dict(tlen=3276.8, imax=32, rmin_=4421., rmax_=6371., rdelta_=250., thetamin=0.,
             thetamax=180., thetadelta=15., maxlmax=600)
FREQS = (1, 4, 32)


def main():
    if not ol.have_ref("sh"):
        raise SystemExit("oracle/_ref/libref_sh.so missing: run make -C oracle ref")
    inp = di.synthetic_input(**SMALL)
    out = {}
    for i in FREQS:
        t, r, s, llog = ol.sh_omega(inp, i, use_ref=True)
        out[f"tsgt_{i}"], out[f"rsgt_{i}"], out[f"synn_{i}"] = t, r, s
        out[f"llog_{i}"] = np.array(llog)
    np.savez_compressed(os.path.join(HERE, "sh_small_ref.npz"), **out)


if __name__ == "__main__":
    main()
