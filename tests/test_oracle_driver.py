"""CPU: the oracle's driver loop pinned by what the physics and the
reference's text fix, independently of the oracle's own code.

The restated routines are pinned one by one against the reference's
compiled routines (test_oracle_psv.py, test_oracle_sh.py); these checks
cover the loop around them (SGTpsv-MPI.f90:535-1100, SGTsh-MPI.f90:389-560):
which right-hand sides are solved, the moment-tensor elements (setmt), the
back-propagated surface forces, and the l cut-off's calamp calls.

1. Source-receiver reciprocity, SYNN against RSGT.  SYNN is the surface
   displacement u_n at distance theta from a unit moment-tensor element
   M_pq at radius r0; RSGT is the strain e_pq at radius r0, distance theta,
   from a unit force in direction n at the surface station.  Reciprocity
   of the elastic Green's tensor gives u_n = sum_pq M_pq e^(n)_pq(r0) for the
   same pair of points; rotating the pair onto the catalog's geometry
   (source or station at the pole, the other point at azimuth 0) changes
   the signs of the theta and phi indices and multiplies the fields of the
   m = 1 forces by -1.  A record component of SYNN and the RSGT component
   of the same index are the same combination of these quantities, so
   SYNN_c = k_c RSGT_c(r0), with |k_c| = 2 where the moment element is
   off-diagonal (M_pq = M_qp = 1) and 1 otherwise.  The signs of k_c are
   fixed by the directions of the back-propagated forces: vertical up, and
   for P-SV the horizontal forces (-x, +y) in the catalog frame of the
   station (x toward the point at azimuth 0), for SH (-x, -y).  (The two
   drivers differ in the transverse force's sign: SGTpsv solves with
   +conjg(rdvec)/lsq, SGTsh with -conjg(rvec)/lsq; SGTpsv-MPI.f90:1028,
   SGTsh-MPI.f90:520.  KernelMaker halves SYNN's off-diagonal slots,
   fourier.f90:79-87, which leaves +-1.)
2. Reciprocity of the strain Green's tensor, TSGT against TSGT with source
   and receiver radii swapped: the 20 TSGT components at azimuth 0 are an
   invertible linear map (udertotsgt, others.f90:231-285) of the 20 strain
   values e_ij[pq] that do not vanish there; reciprocity gives
   e_ij[pq](a <- b) / c_pq = s_ij s_pq e_pq[ij](b <- a) / c_ij with
   s = -1 for the rt and rp elements, c = 2 for off-diagonal elements.
3. The calamp sequence (SGTpsv-MPI.f90:592-609, 902-905; SGTsh-MPI.f90:
   431-436, 476-478): one call per (l, m) at imt = 1 (P-SV) / imt = 3 (SH),
   for every m the loop visits, in m order, the zero-RHS m's reading zero;
   replaying the reference's ismall/maxamp rule on the recorded amplitudes
   gives the logged cut-off, and no call follows it.

Where the reference's own quirks break reciprocity, the affected
quantities are left out, with the reason: bvecdt(1,-m) is never assigned
(trial.f90:197-219), so the e_rtheta strain of m != 0 fields lacks the
m < 0 term (RSGT component 9, and e_rt of the tt, pp, rt sources); at the
poles locallyCartesianDerivatives keeps only the diagonal with u_r / r
(others.f90:317-322), so the poles are left out.

Each check has teeth: test_mutations_are_caught builds the oracle from a
scratch copy of its sources with one plausible mistake at a time (a
back-propagated force's sign, the calamp call site, a setmt entry, the
right-hand-side selection) and requires the matching check to fail.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import dsm_inputs as di
import oracle_lib as ol

BASE = dict(tlen=1638.4, imax=16, thetamin=0., thetamax=180., thetadelta=22.5, maxlmax=1200)
SH_BASE = dict(tlen=3276.8, imax=32, thetamin=0., thetamax=180., thetadelta=15., maxlmax=600)
# source / receiver radii inside their layers (calspo leaves them unchanged)
RA, RB = 6000.0, 5500.0
FREQS = (3, 8)
TOL = 3e-2  # the discretization: measured <= 1.5e-2 at these frequencies

# SYNN_c = K_c RSGT_c(r0); None: left out (the bvecdt(1,-m) quirk)
K_PSV = (1, 2, 1, 1, 1, 1, 1, -2, None, 2)
K_SH = (0, 1, 2, 2, -2)  # SH slots 6..10; slot 6 vanishes (toroidal strain is traceless)


def _inp(src, rec, base=BASE):
    return di.synthetic_input(rmin_=rec, rmax_=rec, rdelta_=10., r0=src, **base)


def synn_rsgt_error(fn, k, base, L=None):
    """Worst relative L2 (over the distances off the poles) of
    SYNN_c - k_c RSGT_c at the source radius, over components and FREQS."""
    worst = 0.0
    inp = _inp(RA, RA, base)
    for i in FREQS:
        t, r, s, _ = fn(inp, i, L=L)
        rr = r[1:-1, 0, :]
        ss = s[1:-1, :]
        for c, kc in enumerate(k):
            if kc is None:
                continue
            ref = np.linalg.norm(ss[:, c]) if kc else np.linalg.norm(ss)
            if ref == 0.0:
                worst = max(worst, 1.0)  # every used component carries signal here
                continue
            worst = max(worst, np.linalg.norm(ss[:, c] - kc * rr[:, c]) / ref)
    return worst


# ---- TSGT: the strain values behind the 20 components ------------------------
_SRC = ("rr", "tt", "pp", "rt", "rp", "tp")
_UNK = [(st, src) for src in ("rr", "tt", "pp", "rt") for st in ("rr", "tt", "pp", "rt")] + \
       [(st, src) for src in ("rp", "tp") for st in ("rp", "tp")]
_IDX = {u: k for k, u in enumerate(_UNK)}


def _tsgt_map():
    """TSGT component k (1-based) = sum M[k-1, j] e[_UNK[j]]: udertotsgt
    (others.f90:231-285) on the strain values that do not vanish at azimuth
    0 (uder(i,j) + uder(j,i) = 2 e_ij)."""
    M = np.zeros((20, 20))

    def add(k, st, src, c):
        M[k - 1, _IDX[st, src]] += c
    add(1, "rr", "rr", 1); add(3, "tt", "rr", -.5); add(3, "pp", "rr", -.5)
    add(6, "rt", "rr", -1); add(10, "tt", "rr", .5); add(10, "pp", "rr", -.5)
    for src, g in (("tt", 1), ("pp", -1)):
        add(2, "rr", src, -.5); add(4, "tt", src, .25); add(4, "pp", src, .25)
        add(8, "rt", src, .5 * g); add(9, "rr", src, .5 * g)
        add(11, "tt", src, -.25); add(11, "pp", src, .25)
        add(12, "tt", src, -.25 * g); add(12, "pp", src, -.25 * g)
        add(17, "rt", src, -.25); add(18, "rt", src, .25)
        for k in (19, 20):
            add(k, "tt", src, .125 * g); add(k, "pp", src, -.125 * g)
    add(5, "rr", "rt", 1); add(7, "tt", "rt", -.5); add(7, "pp", "rt", -.5)
    add(13, "rt", "rt", -.5); add(14, "rt", "rt", .5)
    add(15, "tt", "rt", .25); add(15, "pp", "rt", -.25); add(16, "tt", "rt", -.25); add(16, "pp", "rt", .25)
    add(13, "rp", "rp", .5); add(14, "rp", "rp", .5); add(15, "tp", "rp", -.5); add(16, "tp", "rp", -.5)
    add(17, "rp", "tp", .5); add(18, "rp", "tp", .5); add(19, "tp", "tp", -.5); add(20, "tp", "tp", .5)
    return M


_S = {"rr": 1, "tt": 1, "pp": 1, "rt": -1, "rp": -1, "tp": 1}
_C = {"rr": 1, "tt": 1, "pp": 1, "rt": 2, "rp": 2, "tp": 2}
# e_rt of the m != 0 sources lacks its m < 0 term (bvecdt(1,-m) quirk)
_QUIRK = {("rt", "tt"), ("rt", "pp"), ("rt", "rt")}


def tsgt_reciprocity_error(L=None):
    Minv = np.linalg.inv(_tsgt_map())
    worst = 0.0
    for i in FREQS:
        e_ab = ol.psv_omega(_inp(RB, RA), i, L=L)[0][1:-1, 0, :] @ Minv.T  # at RA from a source at RB
        e_ba = ol.psv_omega(_inp(RA, RB), i, L=L)[0][1:-1, 0, :] @ Minv.T
        for st, src in _UNK:
            if (st, src) in _QUIRK or (src, st) in _QUIRK:
                continue
            x = e_ab[:, _IDX[st, src]] / _C[src]
            y = _S[st] * _S[src] * e_ba[:, _IDX[src, st]] / _C[st]
            nx = np.linalg.norm(x)
            worst = max(worst, np.linalg.norm(x - y) / nx if nx > 0.0 else 1.0)
    return worst


# ---- calamp ------------------------------------------------------------------
def _lsuf(inp, i):
    """callsuf (others.f90:1269-1287): the top zone's vsv coefficients
    summed without powers of r (the surface value), l_suf = int(omega R / vs
    - 0.5) + 1."""
    nz = inp.nzone
    tvs = sum(inp.vsv[nz - 1][k] for k in range(4))
    omega = 2.0 * np.pi * i / inp.tlen
    return int(omega * inp.vrmax[nz - 1] / tvs - 0.5) + 1


def calamp_violations(mode, inp, i, L=None):
    """Differences between the recorded calamp calls of one frequency and
    the reference's loop: calls per l, zero-RHS m's, and the cut-off the
    reference's rule gives on the recorded amplitudes."""
    fn = ol.psv_omega if mode == "psv" else ol.sh_omega
    ls, amps, res = ol.calamp_trace(fn, inp, i, L=L)
    llog_oracle = res[3]
    bad = []
    if not len(ls):
        return ["no calamp call"]
    lmax = int(ls.max())
    ms_of = (lambda l: [m for m in range(-2, 3) if abs(m) <= l]) if mode == "psv" else \
        (lambda l: [m for m in range(-2, 3) if m != 0 and abs(m) <= l])
    zero_m = (lambda m: m != 0) if mode == "psv" else (lambda m: abs(m) == 1)
    pos = 0
    for l in range(1, lmax + 1):
        ms = ms_of(l)
        got = ls[pos:pos + len(ms)]
        if len(got) != len(ms) or (got != l).any():
            bad.append(f"l={l}: {int((ls == l).sum())} calls, the loop makes {len(ms)}")
            break
        for m, a in zip(ms, amps[pos:pos + len(ms)]):
            if zero_m(m) and a != 0.0:
                bad.append(f"l={l} m={m}: amplitude {a} where the right-hand side is zero")
            if not zero_m(m) and a == 0.0:
                bad.append(f"l={l} m={m}: zero amplitude")
        pos += len(ms)
    if pos != len(ls):
        bad.append(f"{len(ls) - pos} calls outside the (l, m) loop order")
    # replay: SGTpsv-MPI.f90:592-609 / SGTsh-MPI.f90:431-436 with calamp
    lsuf, ismall, maxamp, llog, k = _lsuf(inp, i), 0, -1.0, inp.maxlmax, 0
    for l in range(0, inp.maxlmax + 1):
        if ismall > 20:
            llog = min(llog, l)
            break
        while k < len(ls) and ls[k] == l:
            amp = amps[k]
            maxamp = max(maxamp, amp)
            ratio = amp / maxamp if amp != 0.0 and maxamp != 0.0 else 0.0
            past = l >= lsuf if mode == "psv" else l > lsuf
            ismall = ismall + 1 if (ratio < inp.ratl and past) else 0
            k += 1
    if llog - 1 != llog_oracle:
        bad.append(f"replayed cut-off {llog - 1}, logged {llog_oracle}")
    if lmax >= llog:
        bad.append(f"calls at l = {lmax} after the cut-off at l = {llog}")
    return bad


# ---- tests -------------------------------------------------------------------
def test_psv_synn_rsgt_reciprocity():
    e = synn_rsgt_error(ol.psv_omega, K_PSV, BASE)
    print(f"P-SV SYNN vs RSGT at r0: worst relative L2 {e:.2e}")
    assert e <= TOL


def test_sh_synn_rsgt_reciprocity():
    e = synn_rsgt_error(ol.sh_omega, K_SH, SH_BASE)
    print(f"SH SYNN vs RSGT at r0: worst relative L2 {e:.2e}")
    assert e <= TOL


def test_psv_tsgt_reciprocity():
    e = tsgt_reciprocity_error()
    print(f"P-SV TSGT(a <- b) vs TSGT(b <- a): worst relative L2 {e:.2e}")
    assert e <= TOL


DEEP = dict(BASE, rmin_=3000., rmax_=6371., rdelta_=337., thetadelta=30.)


@pytest.mark.parametrize("mode,i", [("psv", 3), ("psv", 40), ("sh", 10)])
def test_calamp_sequence(mode, i):
    base = DEEP if mode == "psv" else dict(SH_BASE, rmin_=4421., rmax_=6371., rdelta_=250.)
    inp = di.synthetic_input(r0=5971.0, **base)
    bad = calamp_violations(mode, inp, i)
    assert not bad, bad


# (file, text, replacement, check): one plausible mistake in a scratch copy
MUTATIONS = [
    ("psv_driver.inc", "g0[nn - 2] = -conj(rdvec[(icomp - 1) + 3 * (m + 2)]);\n                    F->dcsbdlv0",
     "g0[nn - 2] = conj(rdvec[(icomp - 1) + 3 * (m + 2)]);\n                    F->dcsbdlv0", "psv_synn"),
    ("psv_driver.inc", "g0[nn - 2] = -conj(rdvec[(icomp - 1) + 3 * (m + 2)]);\n                    F->rea2_back",
     "g0[nn - 2] = conj(rdvec[(icomp - 1) + 3 * (m + 2)]);\n                    F->rea2_back", "psv_synn"),
    ("psv_driver.inc", "g0[nn - 1] = conj(rdvec[(icomp - 1) + 3 * (m + 2)]) / CMPLX(lsq, 0.0);",
     "g0[nn - 1] = -conj(rdvec[(icomp - 1) + 3 * (m + 2)]) / CMPLX(lsq, 0.0);", "psv_synn"),
    ("psv_driver.inc", "if (imt == 1) {\n                        orc_trace_calamp",
     "if (imt == 1 && m == 0) {\n                        orc_trace_calamp", "psv_calamp"),
    ("psv_driver.inc", "if (imt == 2) mt[4] = 1.0;", "if (imt == 2) mt[8] = 1.0;", "psv_synn"),
    ("psv_driver.inc", "return imt == 2 || imt == 3 || imt == 6;", "return imt == 2 || imt == 3;", "psv_tsgt"),
    ("psv_driver.inc", "if (imt == 4) mt[3] = 1.0;", "if (imt == 4) mt[6] = 1.0;", "psv_tsgt"),
    ("sh_oracle.c", "g0[nn - 1] = -conj(rvec[(icomp - 1) + 3 * (m + 2)]) / CMPLX(lsq, 0.0);",
     "g0[nn - 1] = conj(rvec[(icomp - 1) + 3 * (m + 2)]) / CMPLX(lsq, 0.0);", "sh_synn"),
    ("sh_oracle.c", "if (imt == 3) {\n                    orc_trace_calamp",
     "if (imt == 3 && m < 0) {\n                    orc_trace_calamp", "sh_calamp"),
]


def _check(name, L):
    """True if the check passes on oracle build L."""
    if name == "psv_synn":
        return synn_rsgt_error(ol.psv_omega, K_PSV, BASE, L) <= TOL
    if name == "sh_synn":
        return synn_rsgt_error(ol.sh_omega, K_SH, SH_BASE, L) <= TOL
    if name == "psv_tsgt":
        return tsgt_reciprocity_error(L) <= TOL
    mode = "psv" if name == "psv_calamp" else "sh"
    base = DEEP if mode == "psv" else dict(SH_BASE, rmin_=4421., rmax_=6371., rdelta_=250.)
    return not calamp_violations(mode, di.synthetic_input(r0=5971.0, **base), 10 if mode == "sh" else 3, L)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("k", range(len(MUTATIONS)))
def test_mutations_are_caught(tmp_path, k):
    fname, old, new, check = MUTATIONS[k]
    src = ol.ORACLE_DIR
    for f in ("sh_oracle.c", "psv_oracle.c", "psv_driver.inc", "dsm_oracle.h"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    text = (tmp_path / fname).read_text()
    assert text.count(old) == 1, (fname, old)
    (tmp_path / fname).write_text(text.replace(old, new))
    so = tmp_path / f"libmut{k}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=gnu11", "-shared", str(tmp_path / "sh_oracle.c"),
                    str(tmp_path / "psv_oracle.c"), "-o", str(so), "-ldl", "-lm"], check=True)
    L = ol.load(str(so))
    assert not _check(check, L), f"mutation {k} ({fname}) not caught by {check}"
