"""GPU: the Fortran drop-in programs (dsm-kernel_amd/host/sgtpsv_gpu,
sgtsh_gpu) read the reference's three input files from stdin exactly as
the reference's drivers do and write the reference's catalog and log files
(SGTpsv-MPI.f90:483-531, 1095-1160; SGTsh-MPI.f90:355-380, 560-613).  The
records are compared with the oracle's FP64 accumulators cast to complex64
as the reference casts them."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol
from parity import component_errors

pytestmark = pytest.mark.gpu

HOST = os.path.join(di.ROOT, "dsm-kernel_amd", "host")

CONF = """# DSM configuration
1.d-1
1.d-10
1.d-5
2.d-2
400
end
"""


def _param(r_range, imax):
    return f"""# parameter file in the reference's grammar
DSM.config
out
model4.psv.model
m4
1638.4d0
{r_range}
6371.d0
0.d0 180.d0 30.d0
0 {imax}
1 1 1
end
"""


def _run(tmp_path, prog, r_range, imax):
    shutil.copy(os.path.join(di.CONFIGS, "model4.psv.model"), tmp_path / "model4.psv.model")
    (tmp_path / "DSM.config").write_text(CONF)
    param = _param(r_range, imax)
    (tmp_path / "database.inf").write_text(param)
    with open(tmp_path / "database.inf") as f:
        res = subprocess.run([os.path.join(HOST, prog)], stdin=f, cwd=tmp_path, capture_output=True,
                             text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    conf = di.read_dsm_config(CONF)
    return di.make_input(di.model4_text(), conf, di.read_params(param))


@pytest.mark.parametrize("prog,mode", [("sgtpsv_gpu", "PSV"), ("sgtsh_gpu", "SH")])
def test_dropin_program_writes_the_catalog(tmp_path, prog, mode):
    imax = 4
    inp = _run(tmp_path, prog, "5421.d0 6371.d0 475.d0", imax)
    out = tmp_path / "out"
    ntsgt, nrsgt, nsynn = (20, 10, 10) if mode == "PSV" else (10, 5, 5)
    oracle = ol.psv_omega if mode == "PSV" else ol.sh_omega
    info = ol.psv_info(inp) if mode == "PSV" else ol.sh_info(inp)
    radii = [inp.rmin_ + k * inp.rdelta_ for k in range(int((inp.rmax_ - inp.rmin_) / inp.rdelta_) + 1)]
    radii = radii[len(radii) - info.r_n:]
    T = info.theta_n
    r0 = f"{int(6371.0e3):07d}"
    for i in range(1, imax + 1):
        t, r, s, llog = oracle(inp, i)
        gt = np.stack([np.fromfile(out / "TSGT" / f"m4.{r0}.{int(rr * 1e3):07d}.{i:07d}.TSGT_{mode}",
                                   np.complex64).reshape(T, ntsgt) for rr in radii], axis=1)
        gr = np.stack([np.fromfile(out / "RSGT" / f"m4.{int(rr * 1e3):07d}.{i:07d}.RSGT_{mode}",
                                   np.complex64).reshape(T, nrsgt) for rr in radii], axis=1)
        gs = np.fromfile(out / "RSGT" / f"m4.{r0}.{i:07d}.SYNN_{mode}", np.complex64).reshape(T, nsynn)
        for x, y in ((gt, t), (gr, r), (gs, s)):
            assert component_errors(x.astype(np.complex128), y.astype(np.complex64).astype(np.complex128),
                                    floor=1e-5).max() <= 1e-6
    suffix = "" if mode == "PSV" else "SH"
    logs = os.listdir(out / "log")
    assert f"calLog{suffix}.m4.0000000.{imax:07d}" in logs
    lst = (out / "log" / f"list{suffix}.m4.0000000.{imax:07d}.0000000").read_text().split("\n")
    rows = [ln.split() for ln in lst if ln.strip()]
    assert [int(rw[0]) for rw in rows] == list(range(1, imax + 1))
    assert int(rows[-1][2]) == oracle(inp, imax)[3]


def test_dropin_program_mixed_form_in_windows(tmp_path, monkeypatch):
    """The drop-in program with the batched mixed form
    (DSM_SGT_PRECISION=bf16x6) and a catalog window of 3 frequencies
    (DSM_SGT_WINDOW=3, so the 8 frequencies are written window by window):
    the same files, records within the mixed form's small-grid bound of
    the oracle, the same l cut-offs in the list file."""
    monkeypatch.setenv("DSM_SGT_PRECISION", "bf16x6")
    monkeypatch.setenv("DSM_SGT_WINDOW", "3")
    imax = 8
    inp = _run(tmp_path, "sgtpsv_gpu", "5421.d0 6371.d0 475.d0", imax)
    out = tmp_path / "out"
    info = ol.psv_info(inp)
    radii = [inp.rmin_ + k * inp.rdelta_ for k in range(int((inp.rmax_ - inp.rmin_) / inp.rdelta_) + 1)]
    T = info.theta_n
    r0 = f"{int(6371.0e3):07d}"
    lst = (out / "log" / f"list.m4.0000000.{imax:07d}.0000000").read_text().split("\n")
    rows = [ln.split() for ln in lst if ln.strip()]
    assert [int(rw[0]) for rw in rows] == list(range(1, imax + 1))
    worst = 0.0
    for i in (1, 3, 4, 8):
        t, r, s, llog = ol.psv_omega(inp, i)
        assert int(rows[i - 1][2]) == llog
        gt = np.stack([np.fromfile(out / "TSGT" / f"m4.{r0}.{int(rr * 1e3):07d}.{i:07d}.TSGT_PSV",
                                   np.complex64).reshape(T, 20) for rr in radii], axis=1)
        gs = np.fromfile(out / "RSGT" / f"m4.{r0}.{i:07d}.SYNN_PSV", np.complex64).reshape(T, 10)
        for x, y in ((gt, t), (gs, s)):
            worst = max(worst, component_errors(x.astype(np.complex128), y.astype(np.complex64).astype(np.complex128),
                                                floor=1e-5).max())
    assert worst <= 8e-6, worst
    print(f"drop-in bf16x6 in windows of 3: worst component error {worst:.2e}")
