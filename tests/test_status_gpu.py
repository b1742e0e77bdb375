"""GPU: the numerical status of a run.  The synthesis counts the
non-finite values it stores into the catalogs (a run that produced any
returns DSM_SGT_ENUMERIC and keeps the catalog); the SH band
factorization counts pivots with |d| <= 1e-14, the test of the
reference's dclisb0 (src/SGTsh/dclib.f90:37-86), which reports such a
pivot and continues (P-SV's dcsymbdl0 receives eps = 0 from its driver
and tests nothing).  Valid inputs give neither, in every form; inputs with
non-finite values are refused at create (test_abi.py)."""
import pytest

import dsm_inputs as di

pytestmark = pytest.mark.gpu

SMALL = dict(tlen=1638.4, imax=12, rmin_=4421., rmax_=6371., rdelta_=250., thetamin=0., thetamax=180.,
             thetadelta=15., maxlmax=800)


@pytest.mark.parametrize("mode", ["SH", "PSV"])
@pytest.mark.parametrize("prec", ["FP64", "FP32", "BF16X6"])
def test_valid_runs_report_clean_status(mode, prec):
    import dsm_sgt
    inp = di.synthetic_input(**SMALL)
    eng = dsm_sgt.Engine(dsm_sgt.DsmSgtInput.from_buffer_copy(inp), mode=getattr(dsm_sgt, mode),
                         precision=getattr(dsm_sgt, prec))
    eng.run(0, SMALL["imax"])  # raises on any non-zero status
    st = eng.stats()
    assert st.nonfinite == 0 and st.small_pivots == 0, (st.nonfinite, st.small_pivots)
    eng.close()


def test_catalog_larger_than_memory_asks_for_a_window():
    """A complex64 catalog of ~400 GB (1311 frequencies x 355 radii x 3601
    distances) without a catalog window is refused at create, before the
    large allocations, with ENOMEM naming the window that would fit."""
    import dsm_sgt
    p = dict(tlen=3276.8, imax=1311, rmin_=3716., rmax_=6371., rdelta_=7.5, thetamin=0., thetamax=180.,
             thetadelta=0.05, maxlmax=600)
    inp = dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(**p))
    with pytest.raises(dsm_sgt.DsmSgtError) as e:
        dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6)
    assert e.value.code == 4 and "catalog_window <=" in str(e.value), e.value
