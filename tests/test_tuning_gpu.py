"""GPU: launch plans (dsm_sgt_tuning_t) change how the batched pipeline is
launched, never what it computes.

* graph mode (each batch's launches after the factor-storage read-back run
  as two HIP graphs) gives the catalog of the eager launches bit for bit,
  run after run, with eager work in between;
* the synthesis tile grouping (per-XCD chunks of 2 or 8, or no tile list)
  and the batch size give the same catalog bit for bit: each tile sums the
  same products in the same order, and rows past a frequency's cut-off are
  exact zeros;
* the autotuner returns a plan among its candidates, writes it to the cache
  file and reads it back without probing.

All through the C-ABI (dsm_sgt_create_tuned, dsm_sgt_get_tuning,
dsm_sgt_autotune).
"""
import numpy as np
import pytest

import dsm_inputs as di

pytestmark = pytest.mark.gpu

# 40 frequencies (a short first batch of 16, then batches of the plan's
# size), the depth cut active at the low frequencies, a symmetric distance
# grid (the bf16 forms' mirror), receivers from 3716 km to the surface
GRID = dict(tlen=1638.4, imax=40, rmin_=3716., rmax_=6371., rdelta_=265.5, thetamin=0., thetamax=180.,
            thetadelta=2.0, maxlmax=1500)
FREQS = (1, 2, 17, 33, 40)


def _catalog(eng, freqs=FREQS):
    return {i: eng.fetch_all_fp64(i) for i in freqs}


def _same(a, b):
    for i in a:
        for x, y in zip(a[i], b[i]):
            assert np.array_equal(x, y), f"frequency {i} differs: max {np.abs(x - y).max():.3e}"


def _input():
    import dsm_sgt
    return dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(**GRID))


def test_graph_mode_equals_eager():
    import dsm_sgt
    inp = _input()
    with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6) as eager:
        eager.run(1, GRID["imax"])
        want = _catalog(eager)
        lc = [eager.lcut(i) for i in FREQS]
    with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning={"graph": 1, "nslot": 16}) as g:
        assert g.tuning().graph == 1 and g.tuning().nslot == 16
        for rep in range(3):
            g.clear()
            g.run(1, GRID["imax"])
            _same(_catalog(g), want)
            assert [g.lcut(i) for i in FREQS] == lc
            # eager work of the process between the graph runs
            with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning={"nslot": 16}) as e2:
                e2.run(GRID["imax"] - 3, GRID["imax"])
                _same(_catalog(e2, (GRID["imax"],)), {GRID["imax"]: want[GRID["imax"]]})
        print(f"graph mode: {len(FREQS)} frequencies bit-identical to the eager launches over 3 runs")


@pytest.mark.parametrize("plan", [{"xcd_group": -1}, {"xcd_group": 2}, {"xcd_group": 8}, {"nslot": 24},
                                  {"nslot": 24, "xcd_group": -1, "graph": 1}])
def test_plans_give_the_same_catalog(plan):
    import dsm_sgt
    inp = _input()
    with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6) as ref:
        ref.run(1, GRID["imax"])
        want = _catalog(ref)
    with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning=plan) as e:
        t = e.tuning().as_dict()
        for k, v in plan.items():
            assert t[k] == v, (k, t)
        e.run(1, GRID["imax"])
        _same(_catalog(e), want)


def test_fp64_and_sh_plans_give_the_same_catalog():
    import dsm_sgt
    inp = _input()
    for mode, prec, plan in ((dsm_sgt.PSV, dsm_sgt.FP64, {"nslot": 8}),
                             (dsm_sgt.SH, dsm_sgt.FP64, {"nslot": 8, "lchunk": 256}),
                             (dsm_sgt.SH, dsm_sgt.BF16X6, {"nslot": 8, "lchunk": 1024})):
        with dsm_sgt.Engine(inp, mode=mode, precision=prec) as ref:
            ref.run(1, 20)
            want = _catalog(ref, (1, 9, 20))
        with dsm_sgt.Engine(inp, mode=mode, precision=prec, tuning=plan) as e:
            assert e.tuning().nslot == plan["nslot"]
            e.run(1, 20)
            _same(_catalog(e, (1, 9, 20)), want)


def test_autotune_and_cache(tmp_path):
    import dsm_sgt
    inp = _input()
    cache = str(tmp_path / "tune.txt")
    plan, ms = dsm_sgt.autotune(inp, dsm_sgt.PSV, dsm_sgt.BF16X6, cache_path=cache)
    assert ms > 0.0
    assert plan.nslot in (64, 96, 128) and plan.xcd_group in (0, 2, 8, -1) and plan.graph in (0, 1)
    again, ms2 = dsm_sgt.autotune(inp, dsm_sgt.PSV, dsm_sgt.BF16X6, cache_path=cache)
    assert ms2 == 0.0 and again.as_dict() == plan.as_dict()
    lines = open(cache).read().split("\n")
    assert len([x for x in lines if x.strip()]) == 1
    # another precision is another key
    p64, ms3 = dsm_sgt.autotune(inp, dsm_sgt.PSV, dsm_sgt.FP64, cache_path=cache)
    assert ms3 > 0.0 and p64.nslot in (16, 32, 48)
    with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning=plan) as e, \
            dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6) as ref:
        e.run(1, GRID["imax"])
        ref.run(1, GRID["imax"])
        _same(_catalog(e), _catalog(ref))
    print(f"autotune: plan {plan.as_dict()} in {ms:.1f} ms probe; FP64 plan {p64.as_dict()}")


@pytest.mark.parametrize("mode_prec", ["PSV/BF16X6", "PSV/FP64", "SH/BF16X6"])
def test_catalog_window(mode_prec):
    """A catalog window of 12 frequencies: each run replaces the resident
    catalog; records equal the full catalog's bit for bit; frequencies of
    an earlier run are no longer fetchable but keep their l cut-off; a run
    longer than the window is refused."""
    import dsm_sgt
    mode, prec = (getattr(dsm_sgt, x) for x in mode_prec.split("/"))
    inp = _input()
    with dsm_sgt.Engine(inp, mode=mode, precision=prec) as full:
        full.run(1, 24)
        want = _catalog(full, (1, 12, 13, 24))
        lc = {i: full.lcut(i) for i in (1, 12, 13, 24)}
    with dsm_sgt.Engine(inp, mode=mode, precision=prec, tuning={"catalog_window": 12, "nslot": 8}) as w:
        assert w.tuning().catalog_window == 12
        with pytest.raises(dsm_sgt.DsmSgtError) as e:
            w.run(1, 13)
        assert e.value.code == 1  # EINVAL
        w.run(0, 12)  # i = 0 is skipped, so 12 frequencies
        _same(_catalog(w, (1, 12)), {i: want[i] for i in (1, 12)})
        w.run(13, 24)
        _same(_catalog(w, (13, 24)), {i: want[i] for i in (13, 24)})
        with pytest.raises(dsm_sgt.DsmSgtError) as e:
            w.fetch(1, dsm_sgt.TSGT, 0)
        assert e.value.code == 2  # ESTATE: replaced by the second run
        assert {i: w.lcut(i) for i in lc} == lc


def _env_engine(inp, env, **kw):
    import os
    import dsm_sgt
    old = os.environ.get("DSM_SGT_LBOUND")
    os.environ["DSM_SGT_LBOUND"] = env
    try:
        return dsm_sgt.Engine(inp, **kw)
    finally:
        if old is None:
            os.environ.pop("DSM_SGT_LBOUND", None)
        else:
            os.environ["DSM_SGT_LBOUND"] = old


@pytest.mark.parametrize("factor", ["1.1", "0.05"])
def test_l_bound_deep_source(factor):
    """Deep source (600 km): the batched form bounds each batch's l range by
    the previous batch's cut-off (factor 1.1, no retry expected), or by a
    deliberately short bound (factor 0.05: batches are sent back for the
    full range).  Either way the catalog and the cut-offs equal those of
    full-range launches bit for bit, and the run is reported in
    solve_retries."""
    import os
    import dsm_sgt
    g = dict(GRID, r0=6371.0 - 600.0, imax=48)
    inp = dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(**g))
    freqs = (1, 8, 9, 30, 48)
    with _env_engine(inp, "0", mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning={"nslot": 8}) as full:
        full.run(1, g["imax"])
        want = _catalog(full, freqs)
        lc = [full.lcut(i) for i in freqs]
        assert full.stats().solve_retries == 0
    assert max(lc) < g["maxlmax"] // 2, lc  # the deep source's cut-off is far below lmax
    os.environ["DSM_SGT_LBOUND"] = factor
    try:
        with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning={"nslot": 8}) as b:
            b.run(1, g["imax"])
            _same(_catalog(b, freqs), want)
            assert [b.lcut(i) for i in freqs] == lc
            retries = b.stats().solve_retries
    finally:
        os.environ.pop("DSM_SGT_LBOUND", None)
    if factor == "0.05":
        assert retries > 0
    print(f"l bound x{factor}: cut-offs {lc}, {retries} batch(es) sent back for the full range")


@pytest.mark.parametrize("depth_km", [0.0, 600.0])
def test_sh_depth_cut(depth_km):
    """SH mixed form with the depth cut (default) against the same form
    with it off (DSM_SGT_DEPTH_CUT=0): the rows dropped below the source
    carry less than e^-40 of the solution, so the catalogs agree to
    round-off and the l cut-offs are the same."""
    import os
    import dsm_sgt
    from parity import component_errors
    g = dict(GRID, r0=6371.0 - depth_km, imax=24, rmin_=4421.0, rdelta_=195.0)
    inp = dsm_sgt.DsmSgtInput.from_buffer_copy(di.synthetic_input(**g))
    freqs = (1, 12, 24)
    os.environ["DSM_SGT_DEPTH_CUT"] = "0"
    try:
        with dsm_sgt.Engine(inp, mode=dsm_sgt.SH, precision=dsm_sgt.BF16X6) as full:
            full.run(1, g["imax"])
            want = _catalog(full, freqs)
            lc = [full.lcut(i) for i in freqs]
    finally:
        os.environ.pop("DSM_SGT_DEPTH_CUT", None)
    with dsm_sgt.Engine(inp, mode=dsm_sgt.SH, precision=dsm_sgt.BF16X6) as cut:
        cut.run(1, g["imax"])
        got = _catalog(cut, freqs)
        assert [cut.lcut(i) for i in freqs] == lc
    worst = max(float(component_errors(x, y).max()) for i in freqs for x, y in zip(got[i], want[i]))
    assert worst <= 1e-10, worst
    print(f"SH depth cut, source {depth_km} km: worst component difference {worst:.1e}")


@pytest.mark.parametrize("plan", [{"overlap": 1, "nslot": 8}, {"overlap": 1, "nslot": 16, "catalog_window": 20}])
def test_overlap_plan_gives_the_same_catalog(plan):
    """The overlap plan (the synthesis of a batch on its own stream beside
    the next batch's solves, with two sets of the buffers between them):
    every frequency's records bit-identical to the one-stream plan's, over
    two runs (with a catalog window the second run replaces the first's
    frequencies; without, it recomputes them in place)."""
    import dsm_sgt
    inp = _input()
    imax = GRID["imax"]
    freqs = tuple(range(1, imax + 1))
    with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning={"nslot": plan["nslot"]}) as ref:
        ref.run(1, imax)
        want = _catalog(ref, freqs)
        lc = [ref.lcut(i) for i in freqs]
    with dsm_sgt.Engine(inp, mode=dsm_sgt.PSV, precision=dsm_sgt.BF16X6, tuning=plan) as e:
        assert e.tuning().overlap == 1 and e.tuning().nslot == plan["nslot"]
        runs = [(1, 20), (21, imax)] if plan.get("catalog_window") else [(1, imax), (1, imax)]
        for a, b in runs:
            e.run(a, b)
            _same(_catalog(e, range(a, b + 1)), {i: want[i] for i in range(a, b + 1)})
        assert [e.lcut(i) for i in freqs] == lc
    print(f"overlap plan {plan}: {imax} frequencies bit-identical to the one-stream plan")
