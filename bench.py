#!/usr/bin/env python3
"""SGT catalog benchmark on one MI355X.

Workload (one "step"): the full P-SV strain-Green's-tensor catalog of the
5 s configuration -- synthetic 4-zone sphere with a fluid outer core,
tlen 3276.8 s, frequencies i = 1..656, re = 0.1 (3727 band rows), lmax 8000,
60 receiver radii (3716..6371 km, 45 km), 1801 distances (0..180 deg,
0.1 deg), source at the surface -- computed into HBM (FP64 accumulators of
TSGT_PSV, RSGT_PSV and SYNN_PSV), including the omega-independent set-up
and the Legendre table.  The timed region starts with the parsed model in
host memory and ends with the catalog resident in HBM; writing the files is
not timed.  `--mode sh` runs the 10 s SH catalog instead (i = 1..328,
531 rows, lmax 4000, 40 radii 4421..6371 km).

Times the batched mixed form by default (--form bf16x6: split-bf16
synthesis on MFMA, fused FP64 band solve with the depth cut); the plain
FP64 form (the parity form) is timed once beside it on the same input and
the mixed catalog is compared with it, per component, at 17 frequencies
spread over the catalog (bar: relative L2 <= 1e-6).

Prints one JSON line with the metric of the reference's headline benchmark,
the roofline of the dominant kernel of the timed form and a CPU baseline:
the oracle (C restatement of the reference) timed on a bounded sample of
the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "dsm-kernel_amd"))
sys.path.insert(0, os.path.join(HERE, "tests"))

METRIC = "SGT catalog wall time: batched/mixed vs plain FP64 on 1 GPU at ≤1e-6 rel err"
FP64_MFMA_PEAK = 78.6e12
HBM_PEAK = 8.0e12

WORKLOADS = {
    "psv": dict(tlen=3276.8, imax=656, rmin_=3716.0, rmax_=6371.0, rdelta_=45.0, thetamin=0.0,
                thetamax=180.0, thetadelta=0.1, maxlmax=8000, re=1e-1, r0=6371.0),
    "sh": dict(tlen=3276.8, imax=328, rmin_=4421.0, rmax_=6371.0, rdelta_=50.0, thetamin=0.0,
               thetamax=180.0, thetadelta=0.1, maxlmax=4000, re=1e-1, r0=6371.0),
}


def make_input(p, rsgt=1):
    from dsm_sgt.inputs import read_model, make_input as mk
    with open(os.path.join(HERE, "dsm-kernel_amd", "configs", "model4.psv.model")) as f:
        zones = read_model(f.read())
    conf = dict(re=p["re"], ratc=1e-10, ratl=1e-5, omegai_seed=2e-2, maxlmax=p["maxlmax"])
    params = dict(tlen=p["tlen"], rmin_=p["rmin_"], rmax_=p["rmax_"], rdelta_=p["rdelta_"],
                  r0min=p["r0"], thetamin=p["thetamin"], thetamax=p["thetamax"],
                  thetadelta=p["thetadelta"], imin=0, imax=p["imax"], rsgtswitch=rsgt,
                  tsgtswitch=1, synnswitch=1, psgtswitch=0)
    return mk(zones, conf, params)


def _oracle_sample(args):
    """One frequency of the CPU sample (run in a worker process): the
    oracle on the first `lsample` degrees with 2 and then 12 distances."""
    mode, p, i, lsample = args
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib as ol
    import dsm_inputs as di
    pts = []
    for nth in (2, 12):
        inp = di.synthetic_input(tlen=p["tlen"], imax=p["imax"], rmin_=p["rmin_"], rmax_=p["rmax_"],
                                 rdelta_=p["rdelta_"], thetamin=10.0, thetamax=10.0 + (nth - 1) * 0.1 + 0.01,
                                 thetadelta=0.1, maxlmax=lsample, re=p["re"], r0=p["r0"])
        t0 = time.perf_counter()
        (ol.psv_omega if mode == "psv" else ol.sh_omega)(inp, i, fast=True)
        pts.append((nth, time.perf_counter() - t0))
    return i, pts


def cpu_baseline(mode, p, r_n, lsample=1000):
    """Time the oracle (the C restatement of the reference's driver) on the
    host's cores, one frequency per core as the reference gives one
    frequency to each MPI rank: on each of C frequencies spread over the
    catalog, the same radial grid with all receivers, the first `lsample`
    degrees and two small distance subsets.  Per frequency, fit
    t = L*(a + b*theta_n) (every per-l step costs the same) and extrapolate
    to the full l range and distances; the catalog's core time is the mean
    over the sampled frequencies times their number, its wall time on C
    cores that divided by C (frequencies are independent).  Every
    frequency of the surface-source catalog runs the full l range (the GPU
    run reports the l cut-off of the first and last frequency).  Worker
    processes are started fresh (spawn): they never touch the GPU."""
    import multiprocessing as mp
    # the GPU box gives one GPU's job a share of 16 host cores, whatever
    # os.cpu_count() shows there
    cores = max(1, min(16, os.cpu_count() or 1))
    freqs = sorted(set(max(1, int(round(p["imax"] * (k + 0.5) / cores))) for k in range(cores)))
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(len(freqs)) as pool:
        res = pool.map(_oracle_sample, [(mode, p, i, lsample) for i in freqs])
    wall = time.perf_counter() - t0
    nfull = int(round((p["thetamax"] - p["thetamin"]) / p["thetadelta"])) + 1
    per_omega = []
    for i, ((n1, t1), (n2, t2)) in res:
        b = (t2 - t1) / (n2 - n1)
        a = max(t1 - b * n1, 0.0)
        per_omega.append((a + b * nfull) * (p["maxlmax"] + 1) / (lsample + 1))
    core_s = sum(per_omega) / len(per_omega) * p["imax"]
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    ratio = REF_ROUTINE_RATIO[mode]
    return dict(value=core_s / len(freqs), unit="s", cores=len(freqs), kind="port", core_seconds=core_s,
                nproc=os.cpu_count(), cpu_model=model,
                reference_routines_ratio=ratio, reference_estimate_s=ratio * core_s / len(freqs),
                reference_note=("per-frequency time of the same driver with the reference's own routines "
                                "(flang -O2 builds of src/SGT*/ *.f90) over the restated ones (gcc -O3), "
                                "measured on one core of the build container at l <= 400, 12 distances, "
                                "frequencies 100 and 500; the restated port is the faster baseline"),
                sample=f"oracle (gcc -O3, one frequency per core) on frequencies {freqs} of the same grid, "
                       f"{r_n} radii, l=0..{lsample}, 2 and 12 distances ({wall:.1f} s wall); per frequency "
                       f"t = L*(a + b*theta_n) extrapolated to l=0..{p['maxlmax']}, {nfull} distances; mean x "
                       f"{p['imax']} frequencies / {len(freqs)} cores")


# Time of one frequency with the reference's own numerical routines in the
# oracle's driver (oracle/_ref, flang -O2) over the restated C routines
# (gcc -O3, the timed baseline), measured in the build container on the
# 5 s grid at l <= 400 and 12 distances, frequencies 100 and 500:
# P-SV 2.38 and 2.54, SH 2.89 and 2.76.
REF_ROUTINE_RATIO = {"psv": 2.46, "sh": 2.82}


# Multi-depth 2.5 s workload: P-SV and SH catalogs for 16 source depths
# drawn uniformly in 0..700 km, 60 radii x 3601 distances, lmax 16000,
# i = 1..1311 (tlen 3276.8 s).  The catalog (~1 TB in complex64) does not fit
# in device memory: each engine holds a window of C5_WINDOW frequencies, and
# every run replaces it.  64 spot frequencies (4 per depth) are compared
# with the plain FP64 form of the same depth.
C5 = dict(tlen=3276.8, imax=1311, rmin_=3716.0, rmax_=6371.0, rdelta_=45.0, thetamin=0.0, thetamax=180.0,
          thetadelta=0.05, maxlmax=16000, re=1e-1, r0=6371.0)
C5_SEED = 5
C5_WINDOW = 128


def run_c5(args):
    """Time the batched mixed form (args.form) over the multi-depth 2.5 s
    workload; one JSON line.  Per (depth, mode): engine create + runs of
    C5_WINDOW frequencies with the catalog resident in HBM after each run
    (records fetched only for the spot checks, outside the timed region).
    The spot frequencies' records are kept in host memory (complex64, the
    mixed forms' own values) and compared with the plain FP64 form after
    the mode's timed depths, so that no reference engine is built between
    two timed ones.
    RSGT does not depend on the source (its files carry no source radius in
    their names, SGTpsv-MPI.f90:1121): each mode computes it at its
    first (deepest) source only, rsgtswitch = 0 for the others, as a user
    of the reference would set it; the catalog counts it once."""
    import numpy as np
    import dsm_sgt
    from parity import component_errors
    rng = np.random.default_rng(C5_SEED)
    depths = rng.uniform(0.0, 700.0, 16)
    spots = rng.choice(np.arange(1, C5["imax"] + 1), 64, replace=False)
    prec = getattr(dsm_sgt, FORMS[args.form][0])
    imax = args.imax or C5["imax"]
    spots = [int(x) for x in spots if x <= imax] if args.imax else [int(x) for x in spots]
    modes = (("psv", dsm_sgt.PSV), ("sh", dsm_sgt.SH))
    plans, tune_s, warmup_s = {}, {}, {}
    total = {"psv": 0.0, "sh": 0.0}
    worst = {"psv": {"tsgt": 0.0, "rsgt": 0.0, "synn": 0.0}, "sh": {"tsgt": 0.0, "rsgt": 0.0, "synn": 0.0}}
    bytes64, checked = 0.0, []
    chosen = list(enumerate(depths[:args.c5_depths]))
    deepest = max(chosen, key=lambda x: x[1])[0]  # the depth that computes RSGT
    rows = {d: {"depth_km": float(depth)} for d, depth in chosen}
    # one mode at a time, tuned right before its depths, deepest source first:
    # the engines of the deepest source need the largest buffers, which the
    # shallower sources' engines of the same mode then reuse (device blocks
    # are cached across engines) instead of allocating ~100 GB again
    for name, mode in modes:
        p = dict(C5, imax=imax, r0=6371.0 - float(depths.max()))
        t0 = time.perf_counter()
        if args.plan:
            plan = dsm_sgt.Tuning(**json.loads(args.plan).get(name, {}))
        elif args.no_tune:
            plan = dsm_sgt.Tuning()
        else:
            plan, _ = dsm_sgt.autotune(make_input(p), mode, prec, cache_path=args.tune_cache)
        plan.catalog_window = C5_WINDOW
        plans[name], tune_s[name] = plan, time.perf_counter() - t0
        print(f"c5: {name} plan {plan.as_dict()} ({tune_s[name]:.1f} s)", file=sys.stderr, flush=True)
        # warm-up (untimed, as the warmup steps of the other workloads): one
        # engine of the plan at the deepest source runs the last window, so
        # that the device blocks the timed engines reuse are allocated once
        # (the blocks cached from the other mode's engines are released first:
        # their sizes are not this mode's)
        t0 = time.perf_counter()
        dsm_sgt.release_cache()
        p = dict(C5, imax=imax, r0=6371.0 - float(depths[deepest]))
        with dsm_sgt.Engine(make_input(p), mode=mode, precision=prec, tuning=plans[name]) as warm:
            warm.run(max(1, imax - C5_WINDOW + 1), imax)
        warmup_s[name] = time.perf_counter() - t0
        kept = {}  # (depth index, frequency) -> the mixed form's records
        for k_d, (d, depth) in enumerate(sorted(chosen, key=lambda x: -x[1])):
            p = dict(C5, imax=imax, r0=6371.0 - float(depth))
            rsgt = 1 if k_d == 0 else 0
            kinds = (dsm_sgt.TSGT, dsm_sgt.RSGT, dsm_sgt.SYNN) if rsgt else (dsm_sgt.TSGT, dsm_sgt.SYNN)
            inp = make_input(p, rsgt)
            my_spots = spots[4 * d:4 * d + 4]
            row = rows[d]
            t0 = time.perf_counter()
            eng = dsm_sgt.Engine(inp, mode=mode, precision=prec, tuning=plans[name])
            t_run = time.perf_counter() - t0
            row[name + "_create_s"] = t_run
            info = eng.info
            lcut_max = 0
            for i0 in range(1, imax + 1, C5_WINDOW):
                i1 = min(imax, i0 + C5_WINDOW - 1)
                t1 = time.perf_counter()
                eng.run(i0, i1)
                t_run += time.perf_counter() - t1
                lcut_max = max(lcut_max, max(eng.lcut(i) for i in (i0, i1)))
                for i in my_spots:
                    if i0 <= i <= i1:
                        kept[(d, i)] = tuple(None if x is None else x.astype(np.complex64)
                                             for x in eng.fetch_all_fp64(i, kinds))
            bytes64 += 8.0 * imax * info.theta_n * (info.r_n * (info.num_tsgt + rsgt * info.num_rsgt) + info.num_synn)
            row[name + "_rsgt"] = rsgt
            st = eng.stats()
            row[name + "_retries"] = st.solve_retries
            row[name + "_stages_ms"] = {"solve": st.solve_ms, "synth": st.synth_ms, "other": st.other_ms,
                                        "legendre": st.legendre_ms}
            eng.close()
            total[name] += t_run
            row[name + "_s"] = t_run
            row[name + "_lcut_max"] = lcut_max
            print(f"c5: {name} depth {d + 1}/{args.c5_depths} ({depth:.1f} km): {t_run:.2f} s "
                  f"(create {row[name + '_create_s']:.2f} s)", file=sys.stderr, flush=True)
        # the spot checks of the mode: the plain FP64 form of each depth, one
        # frequency at a time (its window), against the kept records
        for d, depth in chosen:
            my_spots = spots[4 * d:4 * d + 4]
            if not my_spots:
                continue
            rsgt = 1 if d == deepest else 0
            kinds = (dsm_sgt.TSGT, dsm_sgt.RSGT, dsm_sgt.SYNN) if rsgt else (dsm_sgt.TSGT, dsm_sgt.SYNN)
            p = dict(C5, imax=imax, r0=6371.0 - float(depth))
            with dsm_sgt.Engine(make_input(p, rsgt), mode=mode, precision=dsm_sgt.FP64,
                                tuning={"nslot": 4, "catalog_window": 4}) as spot_ref:
                for i in my_spots:
                    spot_ref.run(i, i)
                    got, want = kept.pop((d, i)), spot_ref.fetch_all_fp64(i, kinds)
                    checked.append(i)
                    for k, kind in enumerate(("tsgt", "rsgt", "synn")):
                        if got[k] is not None:
                            worst[name][kind] = max(worst[name][kind], float(component_errors(got[k], want[k]).max()))
            print(f"c5: {name} spot checks of depth {d + 1}/{args.c5_depths}: worst so far {worst[name]}",
                  file=sys.stderr, flush=True)
    per_depth = [rows[d] for d, _ in chosen]
    value = total["psv"] + total["sh"]
    out = {
        "metric": "SGT catalog wall time, multi-depth 2.5 s P-SV + SH (batched mixed form)",
        "value": value, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * value,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": FORMS[args.form][1],
        "data": "synthetic",
        "config": {"workload": f"SGTpsv + SGTsh 2.5 s catalogs for 16 source depths (uniform 0..700 km, seed "
                               f"{C5_SEED}): 4-zone sphere, tlen 3276.8 s, i=1..{imax}, re 0.1, lmax 16000, "
                               f"60 radii (3716..6371 km) x 3601 distances; catalog window {C5_WINDOW} "
                               f"frequencies", "depths_km": [float(x) for x in depths[:args.c5_depths]]},
        "seconds": total, "catalog_bytes_complex64": bytes64,
        "spot_check": {"frequencies": sorted(checked), "per_depth": 4, "against": "plain FP64 form (GPU)",
                       "max_rel_l2": worst, "bar": 1e-6},
        "plans": {k: v.as_dict() for k, v in plans.items()}, "tune_s": tune_s, "warmup_s": warmup_s,
        "per_depth": per_depth,
    }
    print(json.dumps(out), flush=True)
    return 0


FORMS = {"bf16x6": ("BF16X6", "bf16x6+f64"), "bf16x3": ("BF16X3", "bf16x3+f64"), "fp32": ("FP32", "f32+f64"),
         "fp64": ("FP64", "f64")}
BF16_PEAK = 2.5e15
FP32_MFMA_PEAK = 157.3e12
PSV_ROWS, SH_ROWS = 63, 36  # coefficient rows per receiver of the synthesis (psv_jobs.h, sh_jobs.h)


def catalog_errors(got_all, ref, freqs):
    """Worst relative L2 per catalog component (parity.py) of the fetched
    records `got_all[i]` against the engine `ref` over the frequencies
    `freqs`, per kind; and the worst element error, max |x - y| over a
    component's values divided by max |y| of that component."""
    import numpy as np
    from parity import component_errors
    worst = {"tsgt": 0.0, "rsgt": 0.0, "synn": 0.0}
    worst_max = {"tsgt": 0.0, "rsgt": 0.0, "synn": 0.0}
    for i in freqs:
        got, want = got_all[i], ref.fetch_all_fp64(i)
        for k, kind in enumerate(("tsgt", "rsgt", "synn")):
            worst[kind] = max(worst[kind], float(component_errors(got[k], want[k]).max()))
            x, y = np.asarray(got[k]), np.asarray(want[k])
            top = np.abs(y).reshape(-1, y.shape[-1]).max(axis=0)
            diff = np.abs(x - y).reshape(-1, y.shape[-1]).max(axis=0)
            ok = top > 1e-8 * top.max()  # round-off-zero components as in parity.py
            if ok.any():
                worst_max[kind] = max(worst_max[kind], float((diff[ok] / top[ok]).max()))
    return worst, worst_max


# HBM bytes of the solve stage (excitation, band factorization and
# substitutions, coefficient rows, l = 0 systems, cut-off, depth cut) over
# the whole 5 s P-SV catalog: rocprofv3 FETCH_SIZE + WRITE_SIZE summed over
# the stage's dispatches, scaled by a plain copy's counters (x1.991,
# x0.97); profiles/band_solve_psv_5s_catalog.md
SOLVE_STAGE_BYTES = {("psv", "bf16x6"): 2.026e11, ("psv", "fp64"): 1.6449e12}


def band_solve(mode, form, st, full=True):
    """HBM roofline of the solve stage: the algorithmic bytes of the
    reference's systems, 16 (w n + 4 R r_n) per (omega, l >= 1) system below
    the cut-off with n the rows solved after the depth cut (w = 4, R = 18
    for P-SV; w = 2, R = 14 for SH; dsm_sgt_stats_t.solve_alg_bytes), over
    the stage's live device time (HIP events on the engine stream).  The
    rocprof-measured HBM bytes of the stage (FETCH_SIZE + WRITE_SIZE) are
    given beside it where a profile of the same form and workload exists."""
    out = {"systems": st.systems, "stage_ms": st.solve_ms, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "alg_bytes": st.solve_alg_bytes, "design_bytes": st.solve_bytes}
    if st.solve_ms > 0:
        a = st.solve_alg_bytes / (st.solve_ms * 1e-3)
        out.update(achieved=a / 1e9, frac=a / HBM_PEAK)
    b = SOLVE_STAGE_BYTES.get((mode, form)) if full else None
    out["rocprof_bytes"] = b
    if b is not None and st.solve_ms > 0:
        out["rocprof_gbs"] = b / (st.solve_ms * 1e-3) / 1e9
    return out


def vendor_gemm(shapes, reps=10):
    """bf16 GEMM rate of the vendor library (torch.matmul -> hipBLASLt) on
    this device, fp32 accumulation, for (M, N, K) shapes: the practical MFMA
    rate a plain bf16 GEMM of the synthesis's shape reaches here, beside the
    spec-sheet peak."""
    import torch
    out = []
    for (m, n, k) in shapes:
        a = torch.randn(m, k, device="cuda", dtype=torch.bfloat16)
        b = torch.randn(k, n, device="cuda", dtype=torch.bfloat16)
        for _ in range(3):
            c = a @ b
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            c = a @ b
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out.append({"shape": [m, n, k], "ms": ms, "tflops": 2.0 * m * n * k / ms / 1e9})
        del a, b, c
    torch.cuda.empty_cache()
    return out


def profile_traffic(mode, form, plan):
    """HBM bytes per launch of the dominant kernel for this workload, form
    and launch plan, from the rocprof summaries under profiles/: a line
    `traffic per launch: <mode> <form> nslot=<n> xcd_group=<x> graph=<g> = <bytes> B`
    (FETCH_SIZE + WRITE_SIZE scaled by a plain copy's counters); None when
    no profile of this plan exists."""
    import re
    want = (mode, form, plan.get("nslot"), plan.get("xcd_group"), plan.get("graph"))
    pat = re.compile(r"traffic per launch: (\w+) (\w+) nslot=(\d+) xcd_group=(-?\d+) graph=(\d+) = ([0-9.eE+]+) B")
    d = os.path.join(HERE, "profiles")
    for name in sorted(os.listdir(d)) if os.path.isdir(d) else []:
        with open(os.path.join(d, name)) as f:
            for m in pat.finditer(f.read()):
                if (m.group(1), m.group(2), int(m.group(3)), int(m.group(4)), int(m.group(5))) == want:
                    return float(m.group(6))
    return None


def km_leg(eng, p, npoints=4096, reps=3):
    """KernelMaker's catalog stage and transform on the resident catalog of
    the timed engine (P-SV records only): npoints seeded scattering points at
    the middle receiver radius, frequencies 0..imax, vertical component,
    window 2 degrees (21 distances), reduction slowness 13.9 s / degree; then
    the h3 spectra of every point and component to the time domain
    of 1024 points (np1 = 4096, the reference's FFT length for this catalog).  Device time
    from the engine's HIP events; the unique HBM bytes of the stage (one
    radius's catalog read, the reduced catalog written and read once, the
    outputs written) over it."""
    import numpy as np
    import dsm_sgt
    rng = np.random.default_rng(0)
    nf = p["imax"] + 1
    ir = eng.info.r_n // 2
    dr, ds = rng.uniform(0.0, 180.0, npoints), rng.uniform(0.0, 180.0, npoints)
    a = rng.uniform(0.0, 2 * np.pi, (npoints, 3))
    trig = np.stack([f(m * a[:, k]) for k in range(3) for m, f in ((1, np.cos), (1, np.sin), (2, np.cos),
                                                                    (2, np.sin))], axis=1)
    mt = np.array([1.0, -0.5, -0.5, 0.3, 0.2, 0.1])
    args = (eng, None, ir, 0, p["imax"], 0, 13.9, 2.0, mt, dr, ds, trig)
    dsm_sgt.km_points(*args)  # warm-up
    wall, dev = [], []
    for _ in range(reps):
        eng.reset_stats()
        t0 = time.perf_counter()
        h3, h4 = dsm_sgt.km_points(*args)
        wall.append(time.perf_counter() - t0)
        dev.append(eng.stats().other_ms)
    T = eng.info.theta_n
    byts = (16.0 * nf * T * 30 * (eng.info.r_n > 0)  # catalog of one radius read
            + 2 * 16.0 * nf * T * 30  # reduced catalog written and read
            + 2 * 16.0 * npoints * nf * 6)  # outputs
    dms = min(dev)
    spec = h3[:1024].transpose(0, 2, 1).reshape(-1, nf)  # 1024 points x 6 components
    np1 = 4096
    dsm_sgt.km_vectorfft(eng, spec[:64], 0, np1, np.log(50.0) / p["tlen"], p["tlen"], 0, 2 * np1 - 1)
    eng.reset_stats()
    t0 = time.perf_counter()
    dsm_sgt.km_vectorfft(eng, spec, 0, np1, np.log(50.0) / p["tlen"], p["tlen"], 0, 2 * np1 - 1)
    fft_wall = time.perf_counter() - t0
    fft_dev = eng.stats().other_ms
    nvec = spec.shape[0]
    return {"points": npoints, "frequencies": nf, "window_distances": 21, "receiver_radius_km": float(eng.radii[ir]),
            "points_wall_ms": 1e3 * min(wall), "points_device_ms": dms,
            "point_frequencies_per_s": npoints * nf / (dms * 1e-3),
            "unique_hbm_bytes": byts, "hbm_gbs": byts / (dms * 1e-3) / 1e9,
            "fft": {"spectra": nvec, "np1": np1, "wall_ms": 1e3 * fft_wall, "device_ms": fft_dev,
                    "spectra_per_s": nvec / (fft_dev * 1e-3),
                    "hbm_gbs": (16.0 * nvec * nf + 8.0 * nvec * 2 * np1) / (fft_dev * 1e-3) / 1e9},
            "note": "P-SV records only; device time from HIP events on the engine stream; wall time includes "
                    "the host phase tables and the PCIe copies of points and outputs"}


def output_legs(inp, mode, prec, plan, imax, out_dir=None):
    """The catalog out of HBM, timed from engine creation as the timed step
    is: (1) into pinned host memory (dsm_sgt_copy_catalog_async per window
    of 128 frequencies, each window's copy beside the next window's
    compute); (2) into the catalog files (dsm_sgt_write_catalogs_async per
    window, the writer thread beside the next window's compute), in a
    scratch directory removed afterwards.  The pinned buffer is allocated
    before the timed region."""
    import shutil
    import tempfile
    import numpy as np
    import dsm_sgt
    W = 128
    t = dsm_sgt.Tuning(**dict(plan.as_dict() if plan is not None else {}, catalog_window=W))
    out = {"window": W}
    probe = dsm_sgt.Engine(inp, mode=mode, precision=prec, tuning=t)
    blk = probe.record_block()
    probe.close()
    total = 8 * blk * imax
    out["catalog_bytes"] = total
    buf = dsm_sgt.HostBuffer(total)
    t0 = time.perf_counter()
    eng = dsm_sgt.Engine(inp, mode=mode, precision=prec, tuning=t)
    for i0 in range(1, imax + 1, W):
        i1 = min(imax, i0 + W - 1)
        eng.run(i0, i1)
        eng.copy_catalog_async(i0, i1, buf.view(8 * blk * (i0 - 1), 8 * blk * (i1 - i0 + 1)))
    eng.wait()
    out["host_memory_s"] = time.perf_counter() - t0
    # spot check: the last frequency's SYNN record in host memory equals the fetched one
    a = buf.array[blk * (imax - 1):blk * imax]
    n = eng.info
    syn = a[n.theta_n * n.r_n * (n.num_tsgt + n.num_rsgt):].reshape(n.theta_n, n.num_synn)
    out["host_memory_check"] = bool(np.array_equal(syn, eng.fetch(imax, dsm_sgt.SYNN)))
    eng.close()
    buf.close()
    base = out_dir or tempfile.gettempdir()
    free = shutil.disk_usage(base).free
    if free < 1.5 * total:
        out["files_s"] = None
        out["files_note"] = f"skipped: {free / 1e9:.1f} GB free under {base}, the catalog is {total / 1e9:.1f} GB"
        return out
    d = tempfile.mkdtemp(prefix="dsm_catalog_", dir=base)
    try:
        t0 = time.perf_counter()
        eng = dsm_sgt.Engine(inp, mode=mode, precision=prec, tuning=t)
        for i0 in range(1, imax + 1, W):
            i1 = min(imax, i0 + W - 1)
            eng.run(i0, i1)
            eng.write_catalogs_async(i0, i1, d, "model4")
        eng.wait()
        out["files_s"] = time.perf_counter() - t0
        nfiles = sum(len(fs) for _, _, fs in os.walk(d))
        out["files"] = nfiles
        out["files_dir_fs"] = base
        eng.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--vendor-gemm":
        print(json.dumps(vendor_gemm([tuple(x) for x in json.loads(sys.argv[2])])), flush=True)
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=("psv", "sh"), default="psv")
    ap.add_argument("--form", choices=tuple(FORMS), default="bf16x6",
                    help="arithmetic of the timed catalog (the plain FP64 form is timed once beside it)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 reference run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--imax", type=int, default=None, help="(testing) fewer frequencies")
    ap.add_argument("--no-tune", action="store_true", help="default launch plan instead of the autotuned one")
    ap.add_argument("--tune-cache", default=os.path.join(os.environ.get("XDG_CACHE_HOME") or
                                                         os.path.expanduser("~/.cache"), "dsm_sgt_tune.txt"),
                    help="autotuner cache file (plans keyed by input, form and device)")
    ap.add_argument("--plan", default=None,
                    help='launch plan as JSON (e.g. \'{"nslot": 64, "xcd_group": 2}\') instead of the autotuner')
    ap.add_argument("--c5-depths", type=int, default=16, help="(profiling) first N of the 16 source depths")
    ap.add_argument("--no-output", action="store_true",
                    help="skip the catalog output legs (host memory, files)")
    ap.add_argument("--out-dir", default=None, help="directory for the files leg's scratch catalog")
    ap.add_argument("--no-km", action="store_true", help="skip the KernelMaker stage leg")
    ap.add_argument("--no-vendor", action="store_true",
                    help="skip the vendor bf16 GEMM (torch.matmul) calibration of the MFMA roofline")
    ap.add_argument("--workload", choices=("catalog", "c5"), default="catalog",
                    help="catalog: the 5 s P-SV (or --mode sh 10 s SH) catalog; c5: 16 source depths at 2.5 s, "
                         "P-SV + SH")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the catalog's frequencies are independent "
              "and would be split across GPUs, but no multi-GPU run is part of it. Use --gpus 1.",
              file=sys.stderr)
        return 2
    import dsm_sgt
    if args.workload == "c5":
        return run_c5(args)
    p = dict(WORKLOADS[args.mode])
    if args.imax:
        p["imax"] = args.imax
    inp = make_input(p)
    mode = dsm_sgt.PSV if args.mode == "psv" else dsm_sgt.SH
    prec = getattr(dsm_sgt, FORMS[args.form][0])

    # launch plan: the library's autotuner, once per input shape and form
    # (cached; not part of the timed steps)
    plan, tune = None, None
    if args.plan:
        plan = dsm_sgt.Tuning(**json.loads(args.plan))
    elif not args.no_tune:
        os.makedirs(os.path.dirname(os.path.abspath(args.tune_cache)), exist_ok=True)
        t0 = time.perf_counter()
        plan, probe_ms = dsm_sgt.autotune(inp, mode, prec, cache_path=args.tune_cache)
        tune = {"plan": plan.as_dict(), "probe_ms": probe_ms, "tune_s": time.perf_counter() - t0,
                "cached": probe_ms == 0.0}
        print("autotune: %s" % tune, file=sys.stderr, flush=True)

    detail = []

    def step(pr):
        t0 = time.perf_counter()
        eng = dsm_sgt.Engine(inp, mode=mode, precision=pr, tuning=plan if pr == prec else None)
        t1 = time.perf_counter()
        eng.run(1, p["imax"])  # returns with the device synchronised
        st = eng.stats()
        detail.append({"create_s": t1 - t0, "run_s": time.perf_counter() - t1,
                       "prep_wait_s": st.prep_wait_ms * 1e-3, "device_wait_s": st.device_wait_ms * 1e-3})
        return eng

    for _ in range(args.warmup):
        step(prec).close()
    times, eng = [], None
    for _ in range(args.steps):
        if eng is not None:
            eng.close()
        t0 = time.perf_counter()
        eng = step(prec)
        times.append(time.perf_counter() - t0)
        print("step %.3f s: %s" % (times[-1], detail[-1]), file=sys.stderr, flush=True)
    steps_detail = detail[-args.steps:]
    st = eng.stats()
    cuts = [eng.lcut(i) for i in (1, p["imax"])]
    info = eng.info
    eng_plan = eng.tuning().as_dict()
    ms = 1e3 * sum(times) / len(times)
    freqs = sorted(set(int(round(x)) for x in [1 + k * (p["imax"] - 1) / 16 for k in range(17)]))
    compare = args.form != "fp64" and not args.no_fp64
    got_all = {i: eng.fetch_all_fp64(i) for i in freqs} if compare else {}
    km = None if (args.no_km or args.mode != "psv") else km_leg(eng, p)
    eng.close()

    # the plain FP64 form once, and the form's error against it
    fp64 = None
    if compare:
        t0 = time.perf_counter()
        ref = step(dsm_sgt.FP64)
        fp64_s = time.perf_counter() - t0
        fst = ref.stats()
        errs, errs_max = catalog_errors(got_all, ref, freqs)
        fp64 = {"wall_s": fp64_s, "speedup": fp64_s / (ms / 1e3),
                "stages_ms": {"solve+coef": fst.solve_ms, "synth": fst.synth_ms, "other": fst.other_ms},
                "band_solve": band_solve(args.mode, "fp64", fst, not args.imax),
                "max_rel_l2_vs_fp64": errs, "max_elem_err_vs_fp64": errs_max, "error_frequencies": freqs, "bar": 1e-6}
        ref.close()

    r_n, th_n, nn = info.r_n, info.theta_n, info.nn
    launches = max(st.synth_launches, 1)
    launch_ms = st.synth_ms / launches
    # Synthesis flops, three counts over the launches of the dominant kernel
    # (the TSGT/RSGT contraction; main and accumulate launches):
    #  alg:    the algorithmic count 4 sum_omega L(omega) sum_s P_s r_n theta_n,
    #          P_s the real (coefficient, basis) products per right-hand side and
    #          (r, theta): 8 per solid receiver over 18 P-SV right-hand sides
    #          (144), 5 over 14 SH right-hand sides (70);
    #  useful: this formulation's count, 4 flops per (coefficient row, l, r,
    #          theta) with 63 (P-SV) / 36 (SH) rows per receiver, halved where the
    #          distance symmetry computes theta and pi - theta together;
    #  issued: the MFMA flops the bf16 kernels issue (all split cross products
    #          over the tiles' degree extents); FP64: the useful count over the
    #          distance columns the tiles cover.
    rows, p_s = (PSV_ROWS, 144.0) if args.mode == "psv" else (SH_ROWS, 70.0)
    sum_l = st.synth_flops / (4.0 * th_n * rows * r_n)
    alg = 4.0 * sum_l * p_s * r_n * th_n
    mirror = args.form in ("bf16x6", "bf16x3", "fp64")  # the FP32 kernel has no distance symmetry
    useful = st.synth_flops / (2.0 if mirror else 1.0)
    if args.form in ("bf16x6", "bf16x3"):
        kernel, peak = "synth_b16_kernel<%d, 0>" % (3 if args.form == "bf16x6" else 2), BF16_PEAK
        issued = st.synth_mfma_flops
    elif args.form == "fp32":
        kernel, peak, issued = "synth_f32_kernel<0>", FP32_MFMA_PEAK, st.synth_flops
    else:
        # the distance tiles (128 wide) cover the first (theta_n + 1) / 2 distances
        cols = 128 * -(-((th_n + 1) // 2) // 128)
        kernel, peak, issued = "synth_f64_kernel<0, 1>", FP64_MFMA_PEAK, st.synth_flops * cols / th_n
    flops = alg / launches
    achieved = alg / (st.synth_ms * 1e-3) if st.synth_ms > 0 else 0.0
    issued_rate = issued / (st.synth_ms * 1e-3) if st.synth_ms > 0 else 0.0
    bound_note = ("achieved: the algorithmic flops per launch (4 sum L sum_s P_s r_n theta_n) over the average "
                  "launch time; this formulation does alg/useful times fewer flops (63 coefficient rows per "
                  "receiver instead of 144 products, and the distance symmetry), so frac can exceed 1; "
                  "issued_frac is the MFMA issue rate of the kernel against the same peak")
    # HBM bytes per launch of the dominant kernel, P-SV bf16x6 form, 128
    # frequencies per batch, by the synthesis tile grouping of the plan:
    # rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch of
    # synth_b16_kernel<3, 0>, scaled by the counters of a plain 1 GiB copy
    # (x1.991, x0.973); profiles/synth_b16_psv_5s_catalog_mixed.md
    traffic = None if args.imax else profile_traffic(args.mode, args.form, eng_plan)
    form_txt = {"bf16x6": "batched mixed: bf16x6 synthesis, FP64 solve with the depth cut",
                "bf16x3": "batched mixed: bf16x3 synthesis, FP64 solve with the depth cut",
                "fp32": "mixed: FP32 synthesis, FP64 solve with the depth cut", "fp64": "plain FP64"}[args.form]
    if args.mode == "psv":
        workload = (f"SGTpsv 5 s catalog: 4-zone sphere with fluid core, tlen 3276.8 s, i=1..{p['imax']}, "
                    f"re 0.1 ({nn} rows), lmax {p['maxlmax']}, {r_n} radii (3716..6371 km) x {th_n} distances, "
                    f"surface source")
    else:
        workload = (f"SGTsh 10 s catalog: 4-zone sphere, tlen 3276.8 s, i=1..{p['imax']}, re 0.1 ({nn} rows), "
                    f"lmax {p['maxlmax']}, {r_n} radii x {th_n} distances, surface source")
    out = {
        "metric": METRIC,
        "value": ms / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": FORMS[args.form][1],
        "data": "synthetic",
        "config": {"workload": workload, "form": form_txt, "frequencies": p["imax"],
                   "lmax": p["maxlmax"], "band_rows": nn, "radii": r_n, "distances": th_n, "l_cut": cuts},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic, "kernel": kernel, "avg_launch_ms": launch_ms,
                     "launches": launches, "flops_per_launch": flops, "alg_flops": alg, "useful_flops": useful,
                     "issued_flops": issued, "alg_over_useful": alg / useful if useful else None,
                     "issued_tflops": issued_rate / 1e12, "issued_frac": issued_rate / peak, "note": bound_note},
        "stages_ms": {"solve+coef": st.solve_ms, "legendre": st.legendre_ms, "synth": st.synth_ms,
                      "other": st.other_ms},
        "steps_detail": steps_detail,
        "plan": eng_plan,
        "band_solve": band_solve(args.mode, args.form, st, not args.imax),
    }
    if not args.no_vendor and args.form in ("bf16x6", "bf16x3"):
        # the synthesis GEMM's rows (128 frequencies x receivers x re|im),
        # its computed distance columns and a two-function l series; and a
        # large square GEMM.  In a child process: torch's HIP runtime does
        # not open the device beside this process's.
        import subprocess
        shapes = [(256 * r_n, 1024, 16384), (16384, 16384, 16384)]
        res = subprocess.run([sys.executable, os.path.abspath(__file__), "--vendor-gemm", json.dumps(shapes)],
                             capture_output=True, text=True, timeout=300)
        if res.returncode != 0:
            raise RuntimeError("vendor GEMM leg failed: " + res.stderr[-2000:])
        g = json.loads(res.stdout.strip().splitlines()[-1])
        out["roofline"]["vendor_bf16_gemm"] = {
            "note": "torch.matmul (hipBLASLt) bf16 in, fp32 accumulation, on this device in the same run",
            "runs": g, "frac_of_peak": [x["tflops"] * 1e12 / peak for x in g],
            "issued_over_vendor_same_shape": issued_rate / (g[0]["tflops"] * 1e12)}
    if km is not None:
        out["kernelmaker_stage"] = km
    if not args.no_output:
        o = output_legs(inp, mode, prec, plan, p["imax"], args.out_dir)
        o["in_hbm_s"] = ms / 1e3
        out["catalog_output"] = o
        print("catalog output: %s" % o, file=sys.stderr, flush=True)
    if tune is not None:
        out["autotune"] = tune
    if fp64 is not None:
        out["plain_fp64"] = fp64
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.mode, p, r_n)
        out["speedup_vs_cpu"] = out["cpu_baseline"]["value"] / out["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
