"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the
reference routines compiled from their own sources (oracle/_ref).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
import ctypes as C
import os

import numpy as np

from dsm_inputs import DsmInput, ROOT

ORACLE_DIR = os.path.join(ROOT, "oracle")


class ShInfo(C.Structure):
    _fields_ = [("r_n", C.c_int32), ("theta_n", C.c_int32), ("nn", C.c_int32),
                ("nzone", C.c_int32), ("r0", C.c_double), ("spo", C.c_double),
                ("spn", C.c_int32), ("intir0", C.c_int32)]


class PsvInfo(C.Structure):
    _fields_ = [("r_n", C.c_int32), ("theta_n", C.c_int32), ("nn", C.c_int32),
                ("nzone", C.c_int32), ("r0", C.c_double), ("spo", C.c_double),
                ("spn", C.c_int32), ("intir0", C.c_int32), ("nslay", C.c_int32),
                ("nllay", C.c_int32)]


_LIBS = {}


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _LIBS:
        path = os.path.join(ORACLE_DIR, name)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run __graft_entry__.build())")
        _LIBS[name] = load(path)
    return _LIBS[name]


def load(path):
    """An oracle build at `path` (e.g. a scratch build of modified sources)."""
    L = C.CDLL(path)
    if True:
        L.orc_sh_info.argtypes = [C.POINTER(DsmInput), C.POINTER(ShInfo)]
        L.orc_sh_load_ref.argtypes = [C.c_char_p]
        L.orc_sh_omega.argtypes = [C.POINTER(DsmInput), C.c_int32, C.c_int32,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]
        L.orc_psv_info.argtypes = [C.POINTER(DsmInput), C.POINTER(PsvInfo)]
        L.orc_psv_load_ref.argtypes = [C.c_char_p]
        L.orc_psv_mix.argtypes = [C.c_uint64]
        L.orc_psv_omega.argtypes = [C.POINTER(DsmInput), C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]
        L.orc_calamp_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        L.orc_calamp_count.argtypes = []
    return L


def ref_path(kind):
    return os.path.join(ORACLE_DIR, "_ref", f"libref_{kind}.so")


def have_ref(kind="sh"):
    return os.path.exists(ref_path(kind))


def sh_info(inp):
    info = ShInfo()
    rc = lib().orc_sh_info(C.byref(inp), C.byref(info))
    if rc:
        raise RuntimeError(f"orc_sh_info rc={rc}")
    return info


def sh_omega(inp, i, use_ref=False, fast=False, L=None):
    """One SGTsh frequency: FP64 accumulators (tsgt[theta][r][10],
    rsgt[theta][r][5], synn[theta][5]) and the logged l cut-off.  L: an
    oracle build from load() instead of the built one."""
    L = L or lib(fast)
    if use_ref:
        rc = L.orc_sh_load_ref(ref_path("sh").encode())
        if rc:
            raise RuntimeError(f"cannot load reference routines rc={rc}")
    info = sh_info(inp)
    rn, tn = info.r_n, info.theta_n
    tsgt = np.zeros((tn, rn, 10), np.complex128)
    rsgt = np.zeros((tn, rn, 5), np.complex128)
    synn = np.zeros((tn, 5), np.complex128)
    llog = C.c_int32()
    rc = L.orc_sh_omega(C.byref(inp), i, int(use_ref), tsgt.ctypes.data, rsgt.ctypes.data,
                        synn.ctypes.data, C.byref(llog))
    if rc:
        raise RuntimeError(f"orc_sh_omega rc={rc}")
    return tsgt, rsgt, synn, llog.value


def psv_info(inp):
    info = PsvInfo()
    rc = lib().orc_psv_info(C.byref(inp), C.byref(info))
    if rc:
        raise RuntimeError(f"orc_psv_info rc={rc}")
    return info


def psv_omega(inp, i, use_ref=False, fast=False, mix=None, L=None):
    """One SGTpsv frequency: FP64 accumulators (tsgt[theta][r][20],
    rsgt[theta][r][10], synn[theta][10]) and the logged l cut-off.
    use_ref: the reference's routines for every call; mix: a bit mask
    selecting the reference routine per entry of the oracle's table; L: an
    oracle build from load() instead of the built one."""
    L = L or lib(fast)
    if use_ref or mix is not None:
        rc = L.orc_psv_load_ref(ref_path("psv").encode())
        if rc:
            raise RuntimeError(f"cannot load reference routines rc={rc}")
    mode = 1 if use_ref else 0
    if mix is not None:
        rc = L.orc_psv_mix(mix)
        if rc:
            raise RuntimeError(f"orc_psv_mix rc={rc}")
        mode = 2
    info = psv_info(inp)
    rn, tn = info.r_n, info.theta_n
    tsgt = np.zeros((tn, rn, 20), np.complex128)
    rsgt = np.zeros((tn, rn, 10), np.complex128)
    synn = np.zeros((tn, 10), np.complex128)
    llog = C.c_int32()
    rc = L.orc_psv_omega(C.byref(inp), i, mode, tsgt.ctypes.data, rsgt.ctypes.data,
                         synn.ctypes.data, C.byref(llog))
    if rc:
        raise RuntimeError(f"orc_psv_omega rc={rc}")
    return tsgt, rsgt, synn, llog.value


def calamp_trace(fn, inp, i, L=None, cap=1 << 20):
    """Run one frequency (fn = psv_omega or sh_omega) with the calamp trace
    on: (l of every call, amplitude every call computed, result of fn)."""
    L = L or lib()
    ls = np.zeros(cap, np.int32)
    amps = np.zeros(cap, np.float64)
    L.orc_calamp_trace(ls.ctypes.data, amps.ctypes.data, cap)
    try:
        res = fn(inp, i, L=L)
        n = L.orc_calamp_count()
    finally:
        L.orc_calamp_trace(None, None, 0)
    if n > cap:
        raise RuntimeError("calamp trace overflow")
    return ls[:n].copy(), amps[:n].copy(), res
