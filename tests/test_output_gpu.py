"""GPU: the list-form run entry and the catalog output path, through the
C-ABI.

* dsm_sgt_run_list on one MPI rank's frequency set of the reference's
  zig-zag distribution (SGTpsv-MPI.f90:550) gives, bit for bit, the
  records and l cut-offs of dsm_sgt_run over the whole range.
* dsm_sgt_write_catalogs_async writes the same files, byte for byte, as
  the records dsm_sgt_fetch returns (the reference's complex64 cast of the
  FP64 accumulators, component fastest), also when the next catalog window
  is computed while the files are written.
* dsm_sgt_copy_catalog_async puts the same records in host memory, one
  block per frequency.
"""
import os

import numpy as np
import pytest

import dsm_inputs as di

SMALL = dict(tlen=1638.4, imax=16, rmin_=1000., rmax_=6371., rdelta_=250., thetamin=0.,
             thetamax=180., thetadelta=22.5, maxlmax=1200)
SH_SMALL = dict(tlen=3276.8, imax=16, rmin_=4421., rmax_=6371., rdelta_=250., thetamin=0.,
                thetamax=180., thetadelta=15., maxlmax=600)


def _engine(p, mode, prec, **tuning):
    import dsm_sgt
    inp = di.synthetic_input(**p)
    return dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=mode, precision=prec,
                          tuning=tuning or None)


def zigzag(imax, rank, nproc):
    """Frequencies of MPI rank `rank` of `nproc` (SGTpsv-MPI.f90:550)."""
    return [i for i in range(0, imax + 1)
            if i != 0 and ((imax - rank - i) % (2 * nproc) == 0 or (imax + rank + 1 - i) % (2 * nproc) == 0)]


def test_zigzag_sets_partition_the_range():
    """The rank sets are disjoint and cover 1..imax (host logic only)."""
    imax, nproc = 16, 3
    sets = [zigzag(imax, r, nproc) for r in range(nproc)]
    flat = sorted(i for s in sets for i in s)
    assert flat == list(range(1, imax + 1))


@pytest.mark.gpu
@pytest.mark.parametrize("mode_name,prec_name", [("PSV", "FP64"), ("PSV", "BF16X6"), ("SH", "FP64"),
                                                 ("SH", "BF16X6")])
def test_run_list_zigzag_equals_range(mode_name, prec_name):
    import dsm_sgt
    mode, prec = getattr(dsm_sgt, mode_name), getattr(dsm_sgt, prec_name)
    p = SMALL if mode_name == "PSV" else SH_SMALL
    full = _engine(p, mode, prec)
    full.run(1, p["imax"])
    for rank in range(3):
        freqs = zigzag(p["imax"], rank, 3)
        part = _engine(p, mode, prec)
        part.run_list(freqs[::-1])  # any order
        for i in freqs:
            assert part.lcut(i) == full.lcut(i)
            for a, b in zip(part.fetch_all_fp64(i), full.fetch_all_fp64(i)):
                assert np.array_equal(a, b), (mode_name, prec_name, rank, i)
        part.close()
    full.close()


def _files(out, eng, i, model="m4"):
    """The records of frequency i read back from the files."""
    suf = "PSV" if eng.mode == 2 else "SH"
    T = eng.info.theta_n
    r0 = f"{eng.info.intir0:07d}"
    t = [np.fromfile(out / "TSGT" / f"{model}.{r0}.{int(rr * 1e3):07d}.{i:07d}.TSGT_{suf}", np.complex64)
         .reshape(T, eng.info.num_tsgt) for rr in eng.radii]
    r = [np.fromfile(out / "RSGT" / f"{model}.{int(rr * 1e3):07d}.{i:07d}.RSGT_{suf}", np.complex64)
         .reshape(T, eng.info.num_rsgt) for rr in eng.radii]
    s = np.fromfile(out / "RSGT" / f"{model}.{r0}.{i:07d}.SYNN_{suf}", np.complex64).reshape(T, eng.info.num_synn)
    return t, r, s


def _records(eng, i):
    import dsm_sgt
    R = eng.info.r_n
    return ([eng.fetch(i, dsm_sgt.TSGT, ir) for ir in range(R)], [eng.fetch(i, dsm_sgt.RSGT, ir) for ir in range(R)],
            eng.fetch(i, dsm_sgt.SYNN))


def _same(files, recs):
    for a, b in zip(files[0] + files[1] + [files[2]], recs[0] + recs[1] + [recs[2]]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("mode_name", ["PSV", "SH"])
def test_async_files_in_windows(tmp_path, mode_name):
    """Windows of 4 frequencies: the files of window w are written while
    window w + 1 is computed."""
    import dsm_sgt
    mode = getattr(dsm_sgt, mode_name)
    p = dict(SMALL if mode_name == "PSV" else SH_SMALL, imax=12)
    eng = _engine(p, mode, dsm_sgt.BF16X6, catalog_window=4)
    want = {}
    for i0 in (1, 5, 9):
        eng.run(i0, i0 + 3)
        for i in range(i0, i0 + 4):
            want[i] = _records(eng, i)
        eng.write_catalogs_async(i0, i0 + 3, str(tmp_path), "m4")
    eng.wait()
    for i in range(1, 13):
        _same(_files(tmp_path, eng, i), want[i])
    eng.close()


@pytest.mark.gpu
def test_write_catalog_one_frequency_matches_fetch(tmp_path):
    import dsm_sgt
    eng = _engine(dict(SMALL, imax=3), dsm_sgt.PSV, dsm_sgt.FP64)
    eng.run(1, 3)
    eng.write_catalog(2, str(tmp_path), "m4")
    _same(_files(tmp_path, eng, 2), _records(eng, 2))
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode_name", ["PSV", "SH"])
def test_copy_catalog_to_host_memory(mode_name):
    import dsm_sgt
    mode = getattr(dsm_sgt, mode_name)
    p = dict(SMALL if mode_name == "PSV" else SH_SMALL, imax=5)
    eng = _engine(p, mode, dsm_sgt.FP64)
    eng.run(0, 5)
    blk = eng.record_block()
    buf = dsm_sgt.HostBuffer(8 * blk * 5)
    eng.copy_catalog_async(0, 5, buf)
    eng.wait()
    a = buf.array
    R, T = eng.info.r_n, eng.info.theta_n
    nt, nr, ns = eng.info.num_tsgt, eng.info.num_rsgt, eng.info.num_synn
    for k, i in enumerate(range(1, 6)):
        b = a[k * blk:(k + 1) * blk]
        t, r, s = _records(eng, i)
        for ir in range(R):
            assert np.array_equal(b[ir * T * nt:(ir + 1) * T * nt].reshape(T, nt), t[ir])
            off = R * T * nt + ir * T * nr
            assert np.array_equal(b[off:off + T * nr].reshape(T, nr), r[ir])
        assert np.array_equal(b[R * T * (nt + nr):].reshape(T, ns), s)
    # a destination too small is refused
    small = np.zeros(blk, np.complex64)
    with pytest.raises(dsm_sgt.DsmSgtError):
        eng.copy_catalog_async(1, 2, small)
    buf.close()
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode_name", ["PSV", "SH"])
@pytest.mark.parametrize("prec_name", ["FP64", "BF16X6"])
def test_switches_skip_rsgt_and_synn(tmp_path, mode_name, prec_name):
    """rsgtswitch = synnswitch = 0: as the reference, the engine computes
    neither kind (no catalog, no jobs); TSGT equals a full run's bit for bit,
    fetching a skipped kind is ESTATE, and only TSGT files are written."""
    import dsm_sgt
    p = dict(tlen=1638.4, imax=6, rmin_=5371., rmax_=6371., rdelta_=500., thetamin=0., thetamax=180.,
             thetadelta=10., maxlmax=500, r0=6100.0)
    mode, prec = getattr(dsm_sgt, mode_name), getattr(dsm_sgt, prec_name)
    full = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(**p)), mode=mode, precision=prec)
    part_in = dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(**p, switches=(0, 1, 0, 0)))
    part = dsm_sgt.Engine(part_in, mode=mode, precision=prec)
    full.run(1, p["imax"])
    part.run(1, p["imax"])
    for i in range(1, p["imax"] + 1):
        a = full.fetch_all_fp64(i)[0]
        b = part.fetch_all_fp64(i, kinds=(dsm_sgt.TSGT,))[0]
        assert np.array_equal(a, b), i
        assert part.lcut(i) == full.lcut(i)
    for kind in (dsm_sgt.RSGT, dsm_sgt.SYNN):
        with pytest.raises(dsm_sgt.DsmSgtError) as e:
            part.fetch(1, kind, 0)
        assert e.value.code == 2  # ESTATE
    part.write_catalogs_async(1, p["imax"], str(tmp_path), "m")
    part.wait()
    names = [f for _, _, fs in os.walk(tmp_path) for f in fs]
    assert names and all("TSGT" in f or f.startswith(("calLog", "list")) for f in names), names[:5]
    full.close()
    part.close()
