"""GPU: the allocation-free transform dsm_km_vectorfft_noalloc, called
through the C-ABI, against dsm_km_vectorfft on the same spectra.

Both entry points fill the twiddle table with the same kernel and run the
same transform kernel, so the windows must agree sample for sample.  Two
calls in a row share the device's static twiddle storage: the first at the
largest size (np1 = 8192, the 2.5 s catalog's FFT length) on a stream of
the caller's, the second smaller (np1 = 256) on the engine's stream, so
that a table left behind by the first call would show in the second.
Sizes beyond the static storage and arrays that are not device memory of
the engine's GPU, or are too short, are refused before anything is
launched.
"""
import ctypes as C

import numpy as np
import pytest

import dsm_inputs as di

pytestmark = pytest.mark.gpu

OMEGAI, TLEN = np.log(50.0) / 3276.8, 3276.8
# (np1, imin, imax, iws, iwe, nvec, caller's stream)
CALLS = [(8192, 0, 1311, 0, 16383, 5, True), (256, 3, 200, 17, 300, 2, False)]


def _entry(L):
    f = L.dsm_km_vectorfft_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_double,
                  C.c_double, C.c_int32, C.c_int32, C.c_void_p]
    return f


class Hip:
    """The caller's side: device arrays and a stream from the HIP runtime
    the library is linked against (already loaded in this process)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.rt.hipMemGetAddressRange.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.c_void_p]
        self.bufs = []

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def put(self, a):
        """A device copy of the numpy array a."""
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), a.nbytes))
        self.bufs.append(p.value)
        self.ok(self.rt.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1))  # host to device
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        self.ok(self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2))  # device to host
        return out

    def end(self, ptr):
        """The end address of the device allocation holding ptr."""
        base, size = C.c_void_p(), C.c_size_t()
        self.ok(self.rt.hipMemGetAddressRange(C.byref(base), C.byref(size), ptr))
        return base.value + size.value

    def close(self):
        for p in self.bufs:
            self.ok(self.rt.hipFree(p))
        self.bufs = []


@pytest.fixture(scope="module")
def engine():
    import dsm_sgt
    inp = di.synthetic_input(tlen=1638.4, imax=8, rmin_=6371., rmax_=6371., rdelta_=1., thetamin=0., thetamax=10.,
                             thetadelta=1., maxlmax=200, r0=6000.0)
    eng = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=dsm_sgt.SH)
    yield eng
    eng.close()


def _spectra(np1, imin, imax, nvec):
    rng = np.random.default_rng(np1 + imin)
    nk = imax - imin + 1
    sp = rng.standard_normal((nvec, nk)) + 1j * rng.standard_normal((nvec, nk))
    return np.ascontiguousarray(sp * np.exp(-np.arange(nk) / 400.0))


def test_noalloc_matches_vectorfft_on_two_calls(engine):
    import dsm_sgt
    L = dsm_sgt.lib()
    f = _entry(L)
    hip = Hip()
    stream = C.c_void_p()
    hip.ok(hip.rt.hipStreamCreate(C.byref(stream)))
    try:
        for np1, imin, imax, iws, iwe, nvec, own_stream in CALLS:
            spec = _spectra(np1, imin, imax, nvec)
            want = dsm_sgt.km_vectorfft(engine, spec, imin, np1, OMEGAI, TLEN, iws, iwe)
            blank = np.full((nvec, iwe - iws + 1), np.nan)
            d_in, d_out = hip.put(spec), hip.put(blank)
            print(f"noalloc transform: np1 {np1} nvec {nvec} spectrum {imin}..{imax} window {iws}..{iwe} "
                  f"{'caller stream' if own_stream else 'engine stream'}", flush=True)
            rc = f(engine._h, stream.value if own_stream else None, nvec, imin, imax, np1, d_in, OMEGAI, TLEN, iws,
                   iwe, d_out)
            assert rc == 0, L.dsm_sgt_last_error(engine._h).decode()
            if own_stream:
                hip.ok(hip.rt.hipStreamSynchronize(stream.value))
            # (on the engine's stream the call returns when the transform is done)
            got = hip.get(d_out, blank)
            assert np.array_equal(got, want), np.abs(got - want).max()
    finally:
        hip.close()
        hip.ok(hip.rt.hipStreamDestroy(stream.value))


def test_noalloc_refuses_before_launch(engine):
    import dsm_sgt
    L = dsm_sgt.lib()
    f = _entry(L)
    hip = Hip()
    try:
        d_in, d_out = hip.put(np.zeros(2 * 4096)), hip.put(np.zeros(4096))
        host = np.zeros(2 * 4096)
        e_in, e_out = hip.end(d_in), hip.end(d_out)
        cases = [  # (np1, imax, iwe, input, output, words in the message)
            (16384, 100, 100, d_in, d_out, "static"),
            (256, 100, 100, host.ctypes.data, d_out, "device memory"),
            (256, 100, 100, d_in, host.ctypes.data, "device memory"),
            (256, 99, 100, e_in - 16 * 50, d_out, "extends past"),  # 100 values, 50 inside the allocation
            (256, 100, 199, d_in, e_out - 8 * 100, "extends past"),  # 200 samples, 100 inside
        ]
        for np1, imax, iwe, pin, pout, words in cases:
            print(f"noalloc refusal: np1 {np1} imax {imax} iwe {iwe} ({words})", flush=True)
            rc = f(engine._h, None, 1, 0, imax, np1, pin, OMEGAI, TLEN, 0, iwe, pout)
            assert rc == 1, (np1, rc)
            assert words in L.dsm_sgt_last_error(engine._h).decode()
    finally:
        hip.close()
