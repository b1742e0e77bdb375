"""CPU: the C-ABI library loads, exports every entry point include/*.h
declares, and fails loudly (no CPU fallback) where no GPU is present; the
product's input readers agree with the test-side readers."""
import ctypes as C
import os
import re

import pytest

import dsm_inputs as di

ROOT = di.ROOT
LIB = os.path.join(ROOT, "dsm-kernel_amd", "libdsmsgt.so")


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(dsm_(?:sgt|km)_\w+)\s*\(", txt, re.M))
    return names


def test_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first"
    L = C.CDLL(LIB)
    decl = _declared()
    assert "dsm_sgt_create" in decl and "dsm_sgt_run" in decl
    for n in sorted(decl):
        assert hasattr(L, n), n
    import dsm_sgt
    assert set(dsm_sgt.EXPORTS) == decl


def test_no_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import dsm_sgt
    inp = di.synthetic_input(tlen=3276.8, imax=8, rmin_=4421., rmax_=6371., rdelta_=250., thetamin=0.,
                             thetamax=180., thetadelta=15., maxlmax=300)
    with pytest.raises(dsm_sgt.DsmSgtError) as e:
        dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp))
    assert e.value.code == 3  # EHIP


@pytest.mark.parametrize("field", ["rho", "vsv", "qmu", "thetadelta", "omegai_seed"])
def test_create_refuses_non_finite_input(field):
    """dsm_sgt_create checks the inputs before any device call: a NaN in the
    model or the grid is EINVAL naming the bad value (GPU or not)."""
    import math
    import dsm_sgt
    inp = dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(
        tlen=3276.8, imax=8, rmin_=4421., rmax_=6371., rdelta_=250., thetamin=0., thetamax=180., thetadelta=15.,
        maxlmax=300))
    if field in ("rho", "vsv"):
        getattr(inp, field)[1][2] = math.nan
    elif field == "qmu":
        inp.qmu[0] = math.inf
    else:
        setattr(inp, field, math.nan)
    with pytest.raises(dsm_sgt.DsmSgtError) as e:
        dsm_sgt.Engine(inp)
    assert e.value.code == 1, e.value  # EINVAL
    assert "finite" in str(e.value) or "zone" in str(e.value)


PARAM = """# parameter file in the reference's grammar
../configs/DSM.sample.config
out/
model4.psv.model
model4
# tlen
3276.8d0
4421.d0 6371.d0 50.d0
6371.d0
0.d0 180.d0 1.d-1
0 328
1 1 1
end
"""
CONF = """# DSM configuration
1.d-1
1.d-10
1.d-5
2.d-2
4000
end
"""


def test_input_readers_agree():
    import dsm_sgt.inputs as pi
    model = di.model4_text()
    a = pi.make_input(pi.read_model(model), pi.read_dsm_config(CONF), pi.read_params(PARAM))
    b = di.make_input(model, di.read_dsm_config(CONF), di.read_params(PARAM))
    assert bytes(a) == bytes(b)
    assert a.nzone == 4 and a.imax == 328 and abs(a.thetadelta - 0.1) < 1e-15
    assert a.vsv[1][0] == 0.0 and a.vsv[2][0] == 10.6998604327 and a.qkappa[3] == 57823.0
    assert (a.rsgtswitch, a.tsgtswitch, a.synnswitch, a.psgtswitch) == (1, 1, 1, 0)


def test_struct_layouts_match_the_header(tmp_path):
    """The Python bindings' structures have the C-ABI's sizes and field
    offsets (dsm_sgt_tuning_t, dsm_sgt_stats_t, dsm_sgt_info_t,
    dsm_sgt_input_t), checked against the header compiled by gcc."""
    import shutil
    import subprocess
    import dsm_sgt
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    src = tmp_path / "layout.c"
    src.write_text("""#include <stddef.h>
#include <stdio.h>
#include "dsm_sgt.h"
int main(void) {
  printf("%zu %zu %zu %zu\\n", sizeof(dsm_sgt_tuning_t), sizeof(dsm_sgt_stats_t), sizeof(dsm_sgt_info_t),
         sizeof(dsm_sgt_input_t));
  printf("%zu %zu %zu %zu\\n", offsetof(dsm_sgt_tuning_t, catalog_window), offsetof(dsm_sgt_stats_t, solve_alg_bytes),
         offsetof(dsm_sgt_info_t, r0), offsetof(dsm_sgt_stats_t, nonfinite));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes, offs = [list(map(int, ln.split())) for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                        check=True).stdout.split("\n")[:2]]
    assert sizes == [C.sizeof(dsm_sgt.Tuning), C.sizeof(dsm_sgt.Stats), C.sizeof(dsm_sgt.Info),
                     C.sizeof(dsm_sgt.DsmSgtInput)]
    assert offs == [dsm_sgt.Tuning.catalog_window.offset, dsm_sgt.Stats.solve_alg_bytes.offset, dsm_sgt.Info.r0.offset,
                    dsm_sgt.Stats.nonfinite.offset]
