"""GPU parity on the multi-depth 2.5 s configuration's grid: P-SV and SH,
tlen 3276.8 s, re 0.1 (about 7450 P-SV / 2110 SH band rows), lmax 16000,
60 receivers 3716..6371 km every 45 km, with the distance grid reduced to
0..180 degrees every 15 degrees (13 distances, both poles) so that the
oracle finishes in seconds.  Two of the 16 seeded source depths
(numpy default_rng(5).uniform(0, 700, 16)): the deepest (699.4 km, whose
series the reference cuts near l = 150) and the shallowest (31.7 km,
cut near l = 3400), at the lowest frequencies.

The plain FP64 form against the oracle at 1e-9 relative L2 per component
with the same l cut-off; the batched mixed form (BF16x6 synthesis, depth
cut, l bound) at the 1e-6 bar (no receiver at a source radius here).
"""
import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol
from parity import component_errors

pytestmark = pytest.mark.gpu

C5 = dict(tlen=3276.8, imax=1311, rmin_=3716., rmax_=6371., rdelta_=45., thetamin=0., thetamax=180.,
          thetadelta=15., maxlmax=16000, re=1e-1)
DEPTHS = np.random.default_rng(5).uniform(0.0, 700.0, 16)
CASES = ((float(DEPTHS.max()), (1, 2)), (float(DEPTHS.min()), (1,)))
_ORACLE = {}


def _oracle(mode, depth, i):
    key = (mode, depth, i)
    if key not in _ORACLE:
        inp = di.synthetic_input(**C5, r0=6371.0 - depth)
        _ORACLE[key] = (ol.psv_omega if mode == "PSV" else ol.sh_omega)(inp, i, fast=True)
    return _ORACLE[key]


@pytest.mark.parametrize("form,tol", [("FP64", 1e-9), ("BF16X6", 1e-6)])
@pytest.mark.parametrize("mode", ["PSV", "SH"])
def test_c5_grid_vs_oracle(mode, form, tol):
    import dsm_sgt
    worst = 0.0
    for depth, freqs in CASES:
        inp = di.synthetic_input(**C5, r0=6371.0 - depth)
        eng = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=getattr(dsm_sgt, mode),
                             precision=getattr(dsm_sgt, form), tuning={"nslot": 2})
        try:
            assert eng.info.theta_n == 13 and eng.info.r_n == 60
            eng.run(freqs[0], freqs[-1])
            for i in freqs:
                t, r, s, llog = _oracle(mode, depth, i)
                assert eng.lcut(i) == llog, (mode, form, depth, i, eng.lcut(i), llog)
                got = eng.fetch_all_fp64(i)
                for name, x, y in zip(("tsgt", "rsgt", "synn"), got, (t, r, s)):
                    e = component_errors(x, y).max()
                    worst = max(worst, e)
                    assert e <= tol, (mode, form, depth, i, name, e)
            print(f"C5 grid {mode} {form} depth {depth:.1f} km: {eng.info.nn} rows, l cut-off "
                  f"{[eng.lcut(i) for i in freqs]}, worst component error so far {worst:.2e}", flush=True)
        finally:
            eng.close()
