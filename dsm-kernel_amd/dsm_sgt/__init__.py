"""Python host side of the MI355X SGT engine (ctypes over libdsmsgt.so).

Mirrors the reference's driver interface for this path: the same three
input files (stdin parameter file, DSM.config, polynomial model) and the
same catalog records (TSGT / RSGT / SYNN, complex64, component fastest).
The compute runs in the HIP library; there is no CPU fallback: creating an
Engine without the built library or without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

from .inputs import DsmSgtInput, read_inputs, read_model, read_dsm_config, read_params, make_input  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libdsmsgt.so")

SH, PSV = 1, 2
FP64, FP32, BF16X3, BF16X6 = 0, 1, 2, 3
TSGT, RSGT, SYNN = 0, 1, 2

ERRORS = {1: "EINVAL", 2: "ESTATE", 3: "EHIP", 4: "ENOMEM", 5: "EIO", 6: "EUNSUPPORTED", 7: "ENUMERIC"}


class Info(C.Structure):
    _fields_ = [("mode", C.c_int32), ("r_n", C.c_int32), ("theta_n", C.c_int32), ("nn", C.c_int32),
                ("num_tsgt", C.c_int32), ("num_rsgt", C.c_int32), ("num_synn", C.c_int32),
                ("intir0", C.c_int32), ("r0", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("solve_ms", C.c_double), ("legendre_ms", C.c_double), ("synth_ms", C.c_double),
                ("other_ms", C.c_double), ("solve_launches", C.c_int64), ("synth_launches", C.c_int64),
                ("systems", C.c_int64), ("solve_bytes", C.c_double), ("synth_flops", C.c_double),
                ("synth_bytes", C.c_double), ("synth_mfma_flops", C.c_double),
                ("prep_wait_ms", C.c_double), ("device_wait_ms", C.c_double), ("solve_retries", C.c_int64),
                ("solve_alg_bytes", C.c_double), ("small_pivots", C.c_int64), ("nonfinite", C.c_int64)]


class Tuning(C.Structure):
    """Launch plan (dsm_sgt_tuning_t): 0 in a field takes the default."""
    _fields_ = [("nslot", C.c_int32), ("lchunk", C.c_int32), ("xcd_group", C.c_int32), ("graph", C.c_int32),
                ("catalog_window", C.c_int32), ("overlap", C.c_int32), ("reserved", C.c_int32 * 2)]

    def as_dict(self):
        return {"nslot": self.nslot, "lchunk": self.lchunk, "xcd_group": self.xcd_group, "graph": self.graph,
                "catalog_window": self.catalog_window, "overlap": self.overlap}


class KmParams(C.Structure):
    """dsm_km_params_t (include/dsm_km.h)."""
    _fields_ = [("fmin", C.c_int32), ("fmax", C.c_int32), ("compo", C.c_int32), ("reserved", C.c_int32),
                ("c_red_reci", C.c_double), ("ipdistance", C.c_double), ("mt", C.c_double * 6)]


class DsmSgtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None

# every entry point declared in include/dsm_sgt.h
EXPORTS = ("dsm_sgt_create", "dsm_sgt_get_info", "dsm_sgt_get_grid", "dsm_sgt_run", "dsm_sgt_fetch",
           "dsm_sgt_fetch_fp64", "dsm_sgt_lcut", "dsm_sgt_write_catalog", "dsm_sgt_get_stats",
           "dsm_sgt_reset_stats", "dsm_sgt_clear", "dsm_sgt_last_error", "dsm_sgt_destroy", "dsm_sgt_psv_band",
           "dsm_sgt_create_tuned", "dsm_sgt_get_tuning", "dsm_sgt_autotune", "dsm_sgt_run_list",
           "dsm_sgt_write_catalogs_async", "dsm_sgt_copy_catalog_async", "dsm_sgt_wait", "dsm_sgt_host_alloc",
           "dsm_sgt_host_free", "dsm_sgt_release_cache", "dsm_km_points", "dsm_km_vectorfft",
           "dsm_km_vectorfft_noalloc")


def lib():
    """Load libdsmsgt.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.dsm_sgt_create.argtypes = [C.POINTER(DsmSgtInput), C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_void_p)]
        L.dsm_sgt_create_tuned.argtypes = [C.POINTER(DsmSgtInput), C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(Tuning), C.POINTER(C.c_void_p)]
        L.dsm_sgt_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.dsm_sgt_autotune.argtypes = [C.POINTER(DsmSgtInput), C.c_int32, C.c_int32, C.c_int32, C.c_char_p,
                                       C.POINTER(Tuning), C.POINTER(C.c_double)]
        L.dsm_sgt_get_info.argtypes = [C.c_void_p, C.POINTER(Info)]
        L.dsm_sgt_get_grid.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.dsm_sgt_run.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.dsm_sgt_run_list.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        L.dsm_sgt_write_catalogs_async.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_char_p, C.c_char_p]
        L.dsm_sgt_copy_catalog_async.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t]
        L.dsm_sgt_wait.argtypes = [C.c_void_p]
        L.dsm_sgt_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
        L.dsm_sgt_host_free.argtypes = [C.c_void_p]
        L.dsm_sgt_host_free.restype = None
        L.dsm_sgt_release_cache.argtypes = [C.c_int32]
        L.dsm_km_points.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(KmParams), C.c_int32, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.dsm_km_vectorfft.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                       C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p]
        L.dsm_km_vectorfft_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                               C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p]
        L.dsm_sgt_fetch.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.dsm_sgt_fetch_fp64.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.dsm_sgt_lcut.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        L.dsm_sgt_psv_band.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.dsm_sgt_write_catalog.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_char_p]
        L.dsm_sgt_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.dsm_sgt_reset_stats.argtypes = [C.c_void_p]
        L.dsm_sgt_clear.argtypes = [C.c_void_p]
        L.dsm_sgt_last_error.argtypes = [C.c_void_p]
        L.dsm_sgt_last_error.restype = C.c_char_p
        L.dsm_sgt_destroy.argtypes = [C.c_void_p]
        L.dsm_sgt_destroy.restype = None
        _lib = L
    return _lib


class Engine:
    """One engine (handle) on one GPU: `Engine(inp, mode=SH)` runs the
    omega-independent set-up; `run(i0, i1)` computes frequencies; `fetch*`
    return catalog records as numpy arrays shaped (theta_n, num)."""

    def __init__(self, inp, mode=SH, precision=FP64, device=0, tuning=None):
        L = lib()
        self._h = C.c_void_p()
        if tuning is not None and not isinstance(tuning, Tuning):
            tuning = Tuning(**tuning)
        rc = L.dsm_sgt_create_tuned(C.byref(inp), mode, precision, device,
                                    C.byref(tuning) if tuning is not None else None, C.byref(self._h))
        if rc:
            msg = L.dsm_sgt_last_error(self._h).decode() if self._h else ""
            if self._h:
                L.dsm_sgt_destroy(self._h)
            self._h = None
            raise DsmSgtError(rc, msg)
        self.info = Info()
        self._check(L.dsm_sgt_get_info(self._h, C.byref(self.info)))
        self.radii = np.zeros(self.info.r_n)
        self.theta = np.zeros(self.info.theta_n)
        self._check(L.dsm_sgt_get_grid(self._h, self.radii.ctypes.data, self.theta.ctypes.data))
        self.mode = mode

    def _check(self, rc):
        if rc:
            raise DsmSgtError(rc, lib().dsm_sgt_last_error(self._h).decode())

    def num(self, kind):
        return (self.info.num_tsgt, self.info.num_rsgt, self.info.num_synn)[kind]

    def run(self, i0, i1):
        self._check(lib().dsm_sgt_run(self._h, i0, i1))

    def run_list(self, freqs):
        """Compute an arbitrary set of frequencies (dsm_sgt_run_list)."""
        idx = np.ascontiguousarray(np.asarray(list(freqs), dtype=np.int32))
        self._check(lib().dsm_sgt_run_list(self._h, idx.ctypes.data, idx.size))

    def record_block(self):
        """complex64 values of one frequency's records (TSGT, RSGT, SYNN)."""
        i = self.info
        return i.theta_n * (i.r_n * (i.num_tsgt + i.num_rsgt) + i.num_synn)

    def write_catalogs_async(self, i0, i1, output_dir, model_name):
        """Files of every frequency in [i0, i1], written beside the next run;
        wait() returns when they are on disk."""
        self._check(lib().dsm_sgt_write_catalogs_async(self._h, i0, i1, output_dir.encode(), model_name.encode()))

    def copy_catalog_async(self, i0, i1, dst):
        """Records of [i0, i1] into `dst` (a HostBuffer or a writable
        complex64 numpy array of len(freqs) * record_block() values)."""
        ptr, nbytes = (dst.ptr, dst.nbytes) if hasattr(dst, "ptr") else (dst.ctypes.data, dst.nbytes)
        self._keep = dst
        self._check(lib().dsm_sgt_copy_catalog_async(self._h, i0, i1, ptr, nbytes))

    def wait(self):
        self._check(lib().dsm_sgt_wait(self._h))

    def fetch(self, i, kind, ir=0):
        out = np.zeros((self.info.theta_n, self.num(kind)), np.complex64)
        self._check(lib().dsm_sgt_fetch(self._h, i, kind, ir, out.ctypes.data))
        return out

    def fetch_fp64(self, i, kind, ir=0):
        out = np.zeros((self.info.theta_n, self.num(kind)), np.complex128)
        self._check(lib().dsm_sgt_fetch_fp64(self._h, i, kind, ir, out.ctypes.data))
        return out

    def fetch_all_fp64(self, i, kinds=(TSGT, RSGT, SYNN)):
        """(tsgt[theta][r][num_tsgt], rsgt[theta][r][num_rsgt], synn[theta][num_synn]);
        a kind not in `kinds` (e.g. switched off) is None."""
        R = self.info.r_n
        t = np.stack([self.fetch_fp64(i, TSGT, ir) for ir in range(R)], axis=1) if TSGT in kinds else None
        r = np.stack([self.fetch_fp64(i, RSGT, ir) for ir in range(R)], axis=1) if RSGT in kinds else None
        s = self.fetch_fp64(i, SYNN) if SYNN in kinds else None
        return t, r, s

    def psv_band(self, i):
        """Band pieces the device assembles for frequency i (P-SV):
        complex128 [nn][4 band rows][P0, P1, Q0, Q1]."""
        out = np.zeros((self.info.nn, 4, 4), np.complex128)
        self._check(lib().dsm_sgt_psv_band(self._h, i, out.ctypes.data))
        return out

    def tuning(self):
        t = Tuning()
        self._check(lib().dsm_sgt_get_tuning(self._h, C.byref(t)))
        return t

    def lcut(self, i):
        v = C.c_int32()
        self._check(lib().dsm_sgt_lcut(self._h, i, C.byref(v)))
        return v.value

    def write_catalog(self, i, output_dir, model_name):
        self._check(lib().dsm_sgt_write_catalog(self._h, i, output_dir.encode(), model_name.encode()))

    def stats(self):
        s = Stats()
        self._check(lib().dsm_sgt_get_stats(self._h, C.byref(s)))
        return s

    def reset_stats(self):
        self._check(lib().dsm_sgt_reset_stats(self._h))

    def clear(self):
        self._check(lib().dsm_sgt_clear(self._h))

    def close(self):
        if self._h:
            lib().dsm_sgt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def km_points(psv, sh, ir, fmin, fmax, compo, c_red_reci, ipdistance, mt, deltar, deltas, trig):
    """KernelMaker's catalog stage on the GPU (dsm_km_points): h3, h4 as
    complex128 arrays (npoints, fmax - fmin + 1, 6)."""
    prm = KmParams(fmin=fmin, fmax=fmax, compo=compo, c_red_reci=c_red_reci, ipdistance=ipdistance)
    for k in range(6):
        prm.mt[k] = float(mt[k])
    dr = np.ascontiguousarray(deltar, dtype=np.float64)
    ds = np.ascontiguousarray(deltas, dtype=np.float64)
    tg = np.ascontiguousarray(trig, dtype=np.float64).reshape(-1, 12)
    n = dr.size
    h3 = np.zeros((n, fmax - fmin + 1, 6), np.complex128)
    h4 = np.zeros_like(h3)
    rc = lib().dsm_km_points(psv._h, sh._h if sh is not None else None, ir, C.byref(prm), n, dr.ctypes.data,
                             ds.ctypes.data, tg.ctypes.data, h3.ctypes.data, h4.ctypes.data)
    if rc:
        raise DsmSgtError(rc, lib().dsm_sgt_last_error(psv._h).decode())
    return h3, h4


def km_vectorfft(eng, spectra, imin, np1, omegai, tlen, iws, iwe):
    """KernelMaker's frequency-to-time transform on the GPU
    (dsm_km_vectorfft): spectra complex (nvec, imax - imin + 1) -> real
    (nvec, iwe - iws + 1)."""
    sp = np.ascontiguousarray(spectra, dtype=np.complex128)
    if sp.ndim == 1:
        sp = sp[None, :]
    nvec, nk = sp.shape
    out = np.zeros((nvec, iwe - iws + 1))
    rc = lib().dsm_km_vectorfft(eng._h, nvec, imin, imin + nk - 1, np1, sp.ctypes.data, omegai, tlen, iws, iwe,
                                out.ctypes.data)
    if rc:
        raise DsmSgtError(rc, lib().dsm_sgt_last_error(eng._h).decode())
    return out


class HostBuffer:
    """Pinned host memory (dsm_sgt_host_alloc); `array` views it as
    complex64."""

    def __init__(self, nbytes):
        self.ptr_c = C.c_void_p()
        rc = lib().dsm_sgt_host_alloc(nbytes, C.byref(self.ptr_c))
        if rc:
            raise DsmSgtError(rc, f"cannot allocate {nbytes} bytes of pinned host memory")
        self.nbytes = nbytes
        self.ptr = self.ptr_c.value

    def view(self, offset, nbytes):
        """A part of the buffer, for copy_catalog_async."""
        if offset < 0 or offset + nbytes > self.nbytes:
            raise ValueError("view outside the buffer")
        v = type("HostView", (), {})()
        v.ptr, v.nbytes, v.parent = self.ptr + offset, nbytes, self
        return v

    @property
    def array(self):
        buf = (C.c_char * self.nbytes).from_address(self.ptr)
        return np.frombuffer(buf, dtype=np.complex64)

    def close(self):
        if self.ptr_c:
            lib().dsm_sgt_host_free(self.ptr_c)
            self.ptr_c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def release_cache(device=0):
    """Free the device blocks destroyed engines left cached for later
    engines (dsm_sgt_release_cache)."""
    rc = lib().dsm_sgt_release_cache(device)
    if rc:
        raise DsmSgtError(rc, "release_cache failed")


def autotune(inp, mode=PSV, precision=BF16X6, device=0, cache_path=None):
    """Run the library's autotuner (dsm_sgt_autotune): returns (Tuning,
    probe milliseconds; 0.0 when the plan came from the cache file)."""
    t, ms = Tuning(), C.c_double()
    rc = lib().dsm_sgt_autotune(C.byref(inp), mode, precision, device,
                                cache_path.encode() if cache_path else None, C.byref(t), C.byref(ms))
    if rc:
        raise DsmSgtError(rc, "autotune failed")
    return t, ms.value
