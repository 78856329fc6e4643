#!/usr/bin/env python3
"""Benchmark of the weighted-random-sampling hot path on one MI355X.

One step = one PSA alias-table build of N = 2^26 Zipf(1.1) FP32 weights (ranks
a seeded Feistel permutation of 1..N, seed 2) with compensated FP32 tile sums,
the exact fixed-point carry and FP64 mean/split arithmetic, split size 32,
inputs resident in HBM: FP64 total -> heavy/light prefix-partition (tile
aggregates, tile scan, emit) -> block-boundary splits -> merge pack, in the
launch geometry and pack form the autotuner picks.  `value` is the build
throughput in GB/s of the reference pipeline's algorithmic bytes
B_build = 32N + 12K (K = N/32): mean read 4N, partition read 4N + index 4N +
prefix 4N (this build writes the partition-ordered weights in its place),
split 12K, pack read index 4N + weight 4N + write table 8N.  The direct pack
form moves fewer bytes for the same table (the partition writes only the
heavies, the pack reads the lights from the weights); `value` keeps the fixed
B_build so that forms and runs compare by time.

Also reported on the same line:
  roofline      the dominant kernel's algorithmic bytes per launch / its average
                duration, from HIP events around each kernel on the stream the
                kernels run on (a second pass of K builds); `traffic` is the HBM
                bytes per launch of that kernel from the rocprofv3 FETCH_SIZE /
                WRITE_SIZE summary committed under profiles/ (null if absent);
  precision     the same build in all four arithmetic modes: time and the
                implied distribution's max abs/rel error vs FP64;
  sampling      alias sampling, S = 2^26 Philox samples, and their histogram's
                pooled chi-square and RMSE ratio, computed on the device;
  its           ITS on N = 2^24 LogNormal(0, 2): normalised scan (8N bytes)
                and S = 2^24 binary-search samples, CDF error vs the FP64 prefix;
                the cutpoint guide table (N entries) over the same CDF and S = 2^24
                guided samples (bit-identical to the ITS samples);
  spike         N = 2^28 spike mixture build (default and tuned geometry) and
                S = 2^30 alias samples, with the same on-device statistics of the
                2^30 samples and the implied distribution's error vs FP64;
  sweep         S = 1e7 samples at N = 1e5 .. 1e8 (U(0,1) weights), the shape of
                the reference's sampling plots: alias build + sample, ITS and
                cutpoint sampling throughput, and the windowed samplers
                (window 128; groups of 64 samples share a window); `cold`:
                the same with a 512 MiB write before every call (the
                reference flushes its L2 between runs);
  batched       4096 log-uniform-size LogNormal(0, 3) distributions built and
                sampled (64 each) by one hipGraph vs one launch sequence per
                distribution;
  cold          the headline build with a 512 MiB write before every build (the
                reference flushes its L2 between measured runs): median ms and GB/s;
  c1            N = 2^20 U(0,1) weights (seed 0): build, S = 2^20 alias samples,
                their chi-square and RMSE, the implied distribution's error vs FP64;
  scan_2e26     the normalised scan (ITS build) of the headline weights in FP32 and
                compensated modes, GB/s of 8N bytes;
  split_pack    the reference's own split + pack strategies on the headline
                weights, as pack forms: SERIAL (split records + one thread per
                split interval) and SUBGROUP (the same records, 16 lanes per
                interval) at split sizes 2 .. 1024, and INLINE, with the time of
                their split and pack kernels;
  cpu_baseline  the reference CPU path (the oracle's single-threaded C
                restatement) on the same N = 2^26 weights, best of 5.

Every figure beside the headline is timed after at least 10 ms of untimed
calls of the same operation (the first calls after other work run below the
steady rate); the headline takes exactly W warm-up builds and K timed ones.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--only SECTION,...] [--no-tune]
"""
import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "merian-wrs_amd"))

HBM_PEAK_GBS = 8000.0
PREC_NAMES = ["fp32", "fp64-carry", "compensated-fp32", "fp64-tiles"]
METRIC = "PSA build GB/s (% HBM peak) + samples/s; batched vs per-launch at fixed error"
PMC_SUMMARY = os.path.join(ROOT, "profiles", "psa_build_zipf_n2e26_hbm_traffic.md")
SECTIONS = ("precision", "c1", "scan26", "split_pack", "its", "spike", "batched", "sweep")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Implied:
    """Per-element error of an alias table's implied distribution vs FP64."""

    def __init__(self, w: np.ndarray):
        w64 = w.astype(np.float64)
        # sorted summation (relative error ~1e-16); above 2^26 numpy's pairwise
        # FP64 sum (relative error ~1e-15, far below the errors measured)
        total = np.sum(np.sort(w64)) if w64.size <= (1 << 26) else np.sum(w64)
        self.t = w64 / total
        self.nz = self.t > 0

    def error(self, table: np.ndarray) -> dict:
        n = self.t.size
        p = np.clip(table["p"].astype(np.float64), 0.0, 1.0)
        q = p.copy()
        q += np.bincount(table["a"].astype(np.int64), weights=1.0 - p, minlength=n)
        q /= n
        d = np.abs(q - self.t)
        return dict(max_abs=float(d.max()), max_rel=float((d[self.nz] / self.t[self.nz]).max()))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def pmc_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary."""
    try:
        text = open(PMC_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"```json\s*(\{.*?\})\s*```", text, re.S)
    if not m:
        return None
    row = json.loads(m.group(1)).get("kernels", {}).get(kernel)
    return None if row is None else int(row["traffic_bytes"])


def sample_statistics(st: dict) -> dict:
    """On-device histogram statistics -> the reported figures (p-value by scipy)."""
    from scipy import stats as sps

    p = float(sps.chi2.sf(st["chi2"], st["df"])) if st["df"] > 0 else 1.0
    return dict(chi2=round(st["chi2"], 2), df=st["df"], chi2_p=p,
                rmse=st["rmse"], expected_rmse=st["expected_rmse"],
                rmse_ratio=round(st["rmse"] / st["expected_rmse"], 5) if st["expected_rmse"] > 0 else None,
                zero_mass_hits=st["zero_mass_hits"])


def time_it(fn, st, wrs, reps: int) -> float:
    """ms per call of fn() enqueued `reps` times back to back on stream st."""
    e0, e1 = wrs.Event(), wrs.Event()
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    st.synchronize()
    return e0.elapsed_ms(e1) / reps


def time_warm(fn, st, wrs, reps: int, warm_ms: float = 10.0) -> float:
    """time_it after at least warm_ms of untimed calls of fn (at least two):
    the first calls after other work run below the steady rate."""
    t0 = time.perf_counter()
    calls = 0
    while calls < 2 or (time.perf_counter() - t0) * 1e3 < warm_ms:
        fn()
        st.synchronize()
        calls += 1
    return time_it(fn, st, wrs, reps)


def time_cold(fn, st, wrs, reps: int, flush) -> float:
    """median ms of fn() with the caches flushed before every call: a 512 MiB
    copy (flush = (dst, src, n)) evicts the L2s and the 256 MB MALL, as the
    reference flushes its L2 between measured runs."""
    e0, e1 = wrs.Event(), wrs.Event()
    t = []
    for _ in range(reps):
        wrs.copy_f32(flush[0], flush[1], flush[2], stream=st)
        e0.record(st)
        fn()
        e1.record(st)
        st.synchronize()
        t.append(e0.elapsed_ms(e1))
    return float(np.median(t))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--split-size", type=int, default=32)
    ap.add_argument("--precision", type=int, default=2, choices=range(4))
    ap.add_argument("--log2s", type=int, default=26)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline build (no precision study, ITS, spike)")
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch geometry instead of the autotuner's choice")
    ap.add_argument("--only", default=None,
                    help="comma-separated sections to run besides the headline build, of "
                         + ",".join(SECTIONS))
    args = ap.parse_args()
    if args.only is not None:
        sections = {x for x in args.only.split(",") if x}
        if not sections <= set(SECTIONS):
            ap.error(f"--only takes sections of {','.join(SECTIONS)}")
    else:
        sections = set(SECTIONS) if not args.no_extras else {"batched"}
        if args.no_batched:
            sections.discard("batched")
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path has no multi-GPU form here); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    import pywrs as wrs

    info = wrs.device_info()
    log(f"device: {info}")
    n = 1 << args.log2n
    S = 1 << args.log2s
    st = wrs.Stream()

    # ---- synthetic input, generated on the device --------------------------
    w = wrs.empty(n, np.float32)
    wrs.generate(wrs.GEN_ZIPF, n, 2, w, a=1.1, stream=st)
    st.synchronize()
    # the plan, the table and the flush buffers first (allocation leaves the
    # device idle); the plan takes the tuner's geometry afterwards
    plan = wrs.Plan(n, args.split_size, args.precision)
    K = plan.split_count(n)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    fn_ = 1 << 27  # 512 MiB of floats
    flush = (wrs.empty(fn_, np.float32), wrs.empty(fn_, np.float32), fn_)
    st.synchronize()
    # ---- launch geometry: the autotuner times every merge shape on these weights
    tune, tune_log = None, []
    if not args.no_tune:
        tune, tune_log = wrs.autotune_psa(w, n, precision=args.precision,
                                          split_size=args.split_size, stream=st)
        log(f"autotune (PSA, N=2^{args.log2n}): {tune}; candidates (ms, partition/direct/tile/per): "
            + ", ".join(f"{e['partition']}/{e['direct']}/{e['merge_tile']}/{e['merge_per']}: "
                        f"{e['ms']:.4f}" for e in tune_log))
        plan.set_tuning(tune)
    B_build = 32 * n + 12 * K

    # ---- warmup + timed region: K builds back to back ----------------------
    for _ in range(args.warmup):
        plan.psa_build(w, n, table, stream=st)
    st.synchronize()
    ms_step = time_it(lambda: plan.psa_build(w, n, table, stream=st), st, wrs, args.steps)
    gbs = B_build / (ms_step * 1e-3) / 1e9
    log(f"PSA build N=2^{args.log2n} ({PREC_NAMES[args.precision]}): {ms_step:.4f} ms/step, "
        f"{gbs:.1f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}% of peak)")
    # ---- the same builds with the caches flushed before each one (median) --
    cold_ms = time_cold(lambda: plan.psa_build(w, n, table, stream=st), st, wrs, args.steps, flush)
    cold_gbs = B_build / (cold_ms * 1e-3) / 1e9
    log(f"PSA build, 512 MiB flush before each: {cold_ms:.4f} ms, {cold_gbs:.1f} GB/s")

    # ---- per-kernel durations: K more builds, HIP events around each kernel
    plan.set_timing(True)
    acc = np.zeros(len(wrs.Plan.PHASES))
    for _ in range(args.steps):
        plan.psa_build(w, n, table, stream=st)
        acc += np.array(plan.phase_times())
    plan.set_timing(False)
    phase = dict(zip(wrs.Plan.PHASES, (acc / args.steps).tolist()))
    # heavy count H of these weights (one debug build)
    hbuf = wrs.empty(1, np.uint32)
    plan.psa_build_debug(w, n, table, dict(heavy_count=hbuf), stream=st)
    st.synchronize()
    H = int(hbuf.numpy()[0])
    p, t_ = args.precision, (tune or {})
    stream = t_.get("direct", 0) == wrs.PACK_STREAM
    direct = t_.get("direct", 0) != 2  # WRS_PACK_CLASSIC: 2; the default is direct
    lookback = bool(t_.get("partition")) and not stream
    # bytes each kernel must move.  Classic form: the emit reads w and writes
    # index + weight of every rank, the merge pack reads them back and writes
    # the 8-byte table.  Direct form: the emit reads w and writes the heavies'
    # index + weight and every 64th light's position, the merge pack reads w,
    # the heavies' index + weight and writes the table.  Aggregate and reduce
    # read w.
    alg = {"reduce": 4 * n, "finalize": 0, "tile_aggregate": 4 * n, "tile_scan": 0,
           "partition_emit": (4 * n + 8 * H + 4 * ((n - H) // 64)) if direct else 12 * n,
           "boundary_splits": 0,
           "merge_pack": (12 * n + 8 * H) if direct else 16 * n}
    # kernel instances (template arguments: precision, forms, merge shape) as
    # rocprofv3 names them
    mt = t_.get("merge_tile") or 1024
    mp = t_.get("merge_per") or 4
    d_ = "true" if direct else "false"
    kname = {"reduce": f"k_reduce_final<{'false' if p == 0 else 'true'}>",
             "tile_aggregate": f"k_part_agg<{p}>",
             "partition_emit": (f"k_part_lookback<{p}, false, {d_}>" if lookback
                                else f"k_part_emit<{p}, false, {d_}>"),
             "merge_pack": f"k_mergepack<{0 if p == 0 else 1}, {mt}, {mt // mp}, {d_}>"}
    if lookback:  # one kernel: the aggregate pass is empty
        del kname["tile_aggregate"]
    # every kernel instance of one build, as rocprofv3 names them
    fp = "false" if p == 0 else "true"
    build_kernels = [f"k_reduce_final<{fp}>"]  # the reduce's last block finalizes
    if not lookback:
        build_kernels += [f"k_part_agg<{p}>", f"k_part_tilescan<{p}>"]
    build_kernels += [kname["partition_emit"], f"k_coarse<{0 if p == 0 else 1}, {d_}>",
                      f"k_repair_bounds_lb<{d_}>", kname["merge_pack"]]
    if stream:
        # the streamed form: the aggregate reads w and writes the heavies'
        # slots (index + weight); the heavy keys read them and write a 16-byte
        # key and an index per heavy; the rows pass reads w and writes the
        # heavies' rows and crossing offsets; the table pass reads w, the rows,
        # indices and crossings of its tile and writes the table
        phase = dict(zip(("reduce", "finalize", "tile_aggregate", "tile_scan",
                          "heavy_keys_and_ranges", "heavy_rows", "table"), phase.values()))
        alg = {"reduce": 4 * n, "finalize": 0, "tile_aggregate": 4 * n + 8 * H, "tile_scan": 0,
               "heavy_keys_and_ranges": 8 * H + 20 * H, "heavy_rows": 4 * n + 10 * H,
               "table": 12 * n + 14 * H}
        f32 = "true" if p == 0 else "false"
        kname = {"reduce": f"k_reduce_final<{fp}>", "tile_aggregate": "k_sp_agg",
                 "heavy_keys_and_ranges": "k_sp_heavy", "heavy_rows": f"k_sp_rows<{f32}, 512>",
                 "table": f"k_sp_table<{f32}, 256>"}
        build_kernels = [f"k_reduce_final<{fp}>", "k_sp_agg", "k_part_tilescan<1, true>",
                         "k_sp_heavy", "k_sp_qrange", kname["heavy_rows"], kname["table"]]
    log("phases (ms): " + ", ".join(f"{k} {v:.4f}" for k, v in phase.items()))
    dom = max(kname, key=lambda k: phase[k])
    dom_ms = phase[dom]
    dom_gbs = alg[dom] / (dom_ms * 1e-3) / 1e9
    traffic = pmc_traffic(kname[dom])

    # ---- copy calibration ---------------------------------------------------
    cdst = wrs.empty(n, np.float32)
    wrs.copy_f32(cdst, w, n, stream=st)
    copy_ms = time_warm(lambda: wrs.copy_f32(cdst, w, n, stream=st), st, wrs, 10)
    copy_gbs = 2 * 4 * n / (copy_ms * 1e-3) / 1e9
    del cdst
    log(f"copy kernel: {copy_gbs:.1f} GB/s")

    # ---- alias sampling -----------------------------------------------------
    smp = wrs.empty(S, np.uint32)
    wrs.alias_sample(table, n, S, 12345, smp, stream=st)
    ctr = [0]

    def sample_once():
        ctr[0] += 1
        wrs.alias_sample(table, n, S, 12345, smp, counter_base=ctr[0] * S, stream=st)

    sample_ms = time_warm(sample_once, st, wrs, 5)
    samples_per_s = S / (sample_ms * 1e-3)
    log(f"alias sampling S=2^{args.log2s}: {sample_ms:.4f} ms, {samples_per_s / 1e9:.2f} Gsamples/s")
    # statistics of the last S samples, on the device (histogram, pooled chi-square, RMSE)
    hist = wrs.empty(n + 1, np.uint64)
    wrs.histogram(smp, S, n, hist, stream=st)
    sstat = sample_statistics(plan.sample_stats(w, n, hist, S, stream=st))
    log(f"sample statistics: {sstat}")
    del smp, hist

    # ---- accuracy, and the precision study on the same weights -------------
    w_host = w.numpy()
    implied = None if args.no_error else Implied(w_host)
    err = None
    if implied is not None:
        err = implied.error(table.numpy())
        log(f"implied distribution vs FP64: max_abs {err['max_abs']:.3e} max_rel {err['max_rel']:.3e}")
    study = None
    if "precision" in sections:
        study = []
        for prec in range(4):
            pp = plan if prec == args.precision else wrs.Plan(n, args.split_size, prec, tune=tune)
            tt = wrs.empty(n, wrs.ALIAS_DTYPE)
            # warm-up: the device idled during the host-side error computation
            for _ in range(20):
                pp.psa_build(w, n, tt, stream=st)
            ms = time_warm(lambda: pp.psa_build(w, n, tt, stream=st), st, wrs, 10)
            row = dict(mode=PREC_NAMES[prec], ms=round(ms, 4),
                       gbs=round(B_build / (ms * 1e-3) / 1e9, 1))
            if implied is not None:
                e = implied.error(tt.numpy())
                row.update(max_abs=e["max_abs"], max_rel=e["max_rel"])
            study.append(row)
            log(f"precision {row}")
            del tt

    # ---- N = 2^20 U(0,1) weights (seed 0): build + S = 2^20 alias samples ----
    c1 = None
    if "c1" in sections:
        n1 = s1 = 1 << 20
        w1 = wrs.empty(n1, np.float32)
        wrs.generate(wrs.GEN_UNIFORM01, n1, 0, w1, stream=st)
        st.synchronize()
        t1 = None
        if not args.no_tune:
            t1, _ = wrs.autotune_psa(w1, n1, precision=args.precision,
                                     split_size=args.split_size, stream=st)
        p1 = wrs.Plan(n1, args.split_size, args.precision, tune=t1)
        tb1 = wrs.empty(n1, wrs.ALIAS_DTYPE)
        for _ in range(20):
            p1.psa_build(w1, n1, tb1, stream=st)
        b1_ms = time_warm(lambda: p1.psa_build(w1, n1, tb1, stream=st), st, wrs, 20)
        b1_cold = time_cold(lambda: p1.psa_build(w1, n1, tb1, stream=st), st, wrs, 9, flush)
        o1 = wrs.empty(s1, np.uint32)
        wrs.alias_sample(tb1, n1, s1, 0, o1, stream=st)
        a1_ms = time_warm(lambda: wrs.alias_sample(tb1, n1, s1, 0, o1, stream=st), st, wrs, 20)
        h1 = wrs.empty(n1 + 1, np.uint64)
        wrs.histogram(o1, s1, n1, h1, stream=st)
        st1 = sample_statistics(p1.sample_stats(w1, n1, h1, s1, stream=st))
        e1 = Implied(w1.numpy()).error(tb1.numpy())
        bb1 = 32 * n1 + 12 * p1.split_count(n1)
        c1 = dict(n=n1, samples=s1, tuning=t1, build_ms=round(b1_ms, 4),
                  build_gbs=round(bb1 / (b1_ms * 1e-3) / 1e9, 1),
                  cold_build_ms=round(b1_cold, 4), sample_ms=round(a1_ms, 4),
                  samples_per_s=round(s1 / (a1_ms * 1e-3), 1),
                  max_abs_error_vs_fp64=e1["max_abs"], max_rel_error_vs_fp64=e1["max_rel"],
                  sample_statistics=st1)
        log(f"c1: {c1}")
        del w1, tb1, o1, h1, p1

    # ---- normalised scan of the headline weights, FP32 and compensated -----
    scan26 = None
    if "scan26" in sections:
        scan26 = []
        hi26 = wrs.empty(n, np.float32)
        lo26 = wrs.empty(n, np.float32)
        w64 = w_host.astype(np.float64)
        exact26 = np.cumsum(w64) / np.sum(np.sort(w64))
        del w64
        for prec in (0, 2):
            pl = wrs.Plan(n, args.split_size, prec)
            lo_ = lo26 if prec >= 2 else None
            pl.its_build(w, n, hi26, lo_, stream=st)
            ms = time_warm(lambda: pl.its_build(w, n, hi26, lo_, stream=st), st, wrs, 10)
            got = hi26.numpy().astype(np.float64)
            if lo_ is not None:
                got += lo26.numpy().astype(np.float64)
            scan26.append(dict(mode=PREC_NAMES[prec], n=n, ms=round(ms, 4),
                               gbs_8N=round(8 * n / (ms * 1e-3) / 1e9, 1),
                               cdf_max_abs_err=float(np.abs(got - exact26).max())))
            log(f"scan 2^{args.log2n}: {scan26[-1]}")
            del pl, got
        del hi26, lo26, exact26

    # ---- the reference's split + pack strategies on the headline weights ---
    split_pack = None
    if "split_pack" in sections:
        split_pack = []
        tsp_ = wrs.empty(n, wrs.ALIAS_DTYPE)
        for form, split in ((3, 2), (3, 4), (3, 8), (3, 32), (3, 128), (3, 1024),
                            (5, 2), (5, 8), (5, 32), (5, 128), (5, 1024),
                            (4, 32), (4, 128), (4, 1024)):
            pl = wrs.Plan(n, split, args.precision, tune=dict(direct=form))
            pl.psa_build(w, n, tsp_, stream=st)  # allocates its prefix and split arrays
            ms = time_warm(lambda: pl.psa_build(w, n, tsp_, stream=st), st, wrs, 5)
            pl.set_timing(True)
            pl.psa_build(w, n, tsp_, stream=st)
            ph = dict(zip(wrs.Plan.PHASES, pl.phase_times()))
            pl.set_timing(False)
            K_ = pl.split_count(n)
            row = dict(form={3: "serial", 4: "inline", 5: "subgroup"}[form], split_size=split,
                       ms=round(ms, 4),
                       gbs=round((32 * n + 12 * K_) / (ms * 1e-3) / 1e9, 1),
                       partition_ms=round(ph["tile_aggregate"] + ph["tile_scan"]
                                          + ph["partition_emit"], 4),
                       split_ms=round(ph["boundary_splits"], 4),
                       pack_ms=round(ph["merge_pack"], 4))
            split_pack.append(row)
            log(f"split+pack {row}")
            del pl
        del tsp_

    # ---- ITS: normalised scan + binary-search sampling ----------------------
    its = None
    if "its" in sections:
        ni = 1 << 24
        wi = wrs.empty(ni, np.float32)
        wrs.generate(wrs.GEN_LOGNORMAL, ni, 1, wi, a=0.0, b=2.0, stream=st)
        iplan = wrs.Plan(ni, args.split_size, args.precision)
        hi = wrs.empty(ni, np.float32)
        lo = wrs.empty(ni, np.float32)
        iplan.its_build(wi, ni, hi, lo, stream=st)
        scan_ms = time_warm(lambda: iplan.its_build(wi, ni, hi, lo, stream=st), st, wrs, 10)
        outi = wrs.empty(ni, np.uint32)
        wrs.its_sample(hi, lo, ni, ni, 7, outi, stream=st)
        its_ms = time_warm(lambda: wrs.its_sample(hi, lo, ni, ni, 7, outi, stream=st), st, wrs, 5)
        # cutpoint: guide table (m = N entries) over the same CDF, guided samples
        gd = wrs.empty(ni + 1, np.uint32)
        wrs.cutpoint_build(hi, lo, ni, gd, ni, stream=st)
        guide_ms = time_warm(lambda: wrs.cutpoint_build(hi, lo, ni, gd, ni, stream=st), st, wrs, 5)
        outc = wrs.empty(ni, np.uint32)
        wrs.cutpoint_sample(hi, lo, ni, gd, ni, ni, 7, outc, stream=st)
        cut_ms = time_warm(lambda: wrs.cutpoint_sample(hi, lo, ni, gd, ni, ni, 7, outc, stream=st),
                         st, wrs, 5)
        st.synchronize()
        cut_same = bool(np.array_equal(outc.numpy(), outi.numpy()))
        wi64 = wi.numpy().astype(np.float64)
        exact = np.cumsum(wi64) / np.sum(np.sort(wi64))
        got = hi.numpy().astype(np.float64) + lo.numpy().astype(np.float64)
        its = dict(n=ni, scan_ms=round(scan_ms, 4), scan_gbs_8N=round(8 * ni / (scan_ms * 1e-3) / 1e9, 1),
                   sample_ms=round(its_ms, 4), samples_per_s=round(ni / (its_ms * 1e-3), 1),
                   cdf_max_abs_err=float(np.abs(got - exact).max()),
                   cutpoint=dict(guide_entries=ni, guide_build_ms=round(guide_ms, 4),
                                 sample_ms=round(cut_ms, 4),
                                 samples_per_s=round(ni / (cut_ms * 1e-3), 1),
                                 identical_to_its=cut_same))
        log(f"its: {its}")
        del wi, hi, lo, outi, iplan, gd, outc

    # ---- spike mixture at N = 2^28, S = 2^30 --------------------------------
    spike = None
    if "spike" in sections:
        ns, ss = 1 << 28, 1 << 30
        wsp = wrs.empty(ns, np.float32)
        wrs.generate(wrs.GEN_SPIKE, ns, 4, wsp, stream=st)
        splan = wrs.Plan(ns, args.split_size, args.precision)
        tsp = wrs.empty(ns, wrs.ALIAS_DTYPE)
        splan.psa_build(wsp, ns, tsp, stream=st)
        sp_ms = time_warm(lambda: splan.psa_build(wsp, ns, tsp, stream=st), st, wrs, 5)
        osp = wrs.empty(ss, np.uint32)
        wrs.alias_sample(tsp, ns, ss, 99, osp, stream=st)
        sps_ms = time_warm(lambda: wrs.alias_sample(tsp, ns, ss, 99, osp, stream=st), st, wrs, 3)
        hsp = wrs.empty(ns + 1, np.uint64)
        wrs.histogram(osp, ss, ns, hsp, stream=st)
        spstat = sample_statistics(splan.sample_stats(wsp, ns, hsp, ss, stream=st))
        del hsp
        b_sp = 32 * ns + 12 * splan.split_count(ns)
        del osp
        # the tuned geometry, and the implied distribution's error vs FP64
        sp_tune, sp_tms = None, None
        if not args.no_tune:
            sp_tune, _ = wrs.autotune_psa(wsp, ns, precision=args.precision,
                                          split_size=args.split_size, stream=st)
            tplan = wrs.Plan(ns, args.split_size, args.precision, tune=sp_tune)
            tplan.psa_build(wsp, ns, tsp, stream=st)
            sp_tms = time_warm(lambda: tplan.psa_build(wsp, ns, tsp, stream=st), st, wrs, 5)
            del tplan
        esp = None if args.no_error else Implied(wsp.numpy()).error(tsp.numpy())
        spike = dict(n=ns, build_ms=round(sp_ms, 4), build_gbs=round(b_sp / (sp_ms * 1e-3) / 1e9, 1),
                     frac_hbm_peak=round(b_sp / (sp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     tuning=sp_tune,
                     tuned_build_ms=None if sp_tms is None else round(sp_tms, 4),
                     tuned_frac_hbm_peak=None if sp_tms is None else
                     round(b_sp / (sp_tms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     max_abs_error_vs_fp64=None if esp is None else esp["max_abs"],
                     max_rel_error_vs_fp64=None if esp is None else esp["max_rel"],
                     samples=ss, sample_ms=round(sps_ms, 3),
                     samples_per_s=round(ss / (sps_ms * 1e-3), 1), sample_statistics=spstat)
        log(f"spike: {spike}")
        del wsp, tsp, splan

    # ---- batched 4096-distribution build + sample --------------------------
    batched = None
    if "batched" in sections:
        nseg, per = 4096, 64
        sizes = wrs.segment_sizes(3, nseg)
        off = wrs.segment_offsets(sizes)
        tot = int(off[-1])
        bw = wrs.empty(tot, np.float32)
        wrs.generate(wrs.GEN_LOGNORMAL, tot, 3, bw, a=0.0, b=3.0, stream=st)
        doff = wrs.to_device(off)
        btune, btune_log = None, []
        if not args.no_tune:
            btune, btune_log = wrs.autotune_batched(off, bw, precision=args.precision,
                                                    split_size=args.split_size, stream=st)
            log(f"autotune (batched): {btune}; candidates (ms): " + ", ".join(
                f"{e['partition']}/{e['merge_tile']}/{e['merge_per']}: {e['ms']:.4f}"
                for e in btune_log))
        bp = wrs.BatchPlan(off, args.split_size, args.precision, tune=btune)
        tb = wrs.empty(tot, wrs.ALIAS_DTYPE)
        sb = wrs.empty(nseg * per, np.uint32)
        g = wrs.Graph(st)
        with g:
            bp.build(bw, tb, stream=st)
            wrs.alias_sample_batched(tb, doff, nseg, per, 42, sb, stream=st)
        g.launch()
        st.synchronize()
        batched_ms = time_it(g.launch, st, wrs, 10)
        eplan = wrs.Plan(int(sizes.max()), args.split_size, args.precision, tune=btune)
        te = wrs.empty(tot, wrs.ALIAS_DTYPE)
        se = wrs.empty(nseg * per, np.uint32)
        wrs.build_sample_each(eplan, bw, off, te, per, 42, se, stream=st)
        st.synchronize()
        each = []
        for _ in range(3):
            t0 = time.perf_counter()
            wrs.build_sample_each(eplan, bw, off, te, per, 42, se, stream=st)
            st.synchronize()
            each.append((time.perf_counter() - t0) * 1e3)
        each_ms = min(each)
        same = bool(np.array_equal(tb.numpy(), te.numpy()) and np.array_equal(sb.numpy(), se.numpy()))
        batched = dict(distributions=nseg, weights=tot, samples_per_distribution=per,
                       batched_graph_ms=round(batched_ms, 4), per_distribution_ms=round(each_ms, 3),
                       speedup=round(each_ms / batched_ms, 2), identical_tables_and_samples=same,
                       tuning=btune, tuning_candidates_ms=[
                           [e["partition"], e["merge_tile"], e["merge_per"], round(e["ms"], 4)]
                           for e in btune_log],
                       batched_build_gbs=round((32 * tot + 12 * int(np.sum(-(-sizes // args.split_size))))
                                               / (batched_ms * 1e-3) / 1e9, 1))
        log(f"batched: {batched}")
        del bw, tb, sb, te, se

    # ---- sampling throughput over N, as the reference plots it -------------
    # S = 1e7 samples from U(0,1) weights at N = 1e5 .. 1e8: alias build + sample
    # (S / (t_build + t_sample), its_sampling_throughput_with_construction), ITS
    # binary search and cutpoint sampling (S / t_sample, its_sampling_throughput)
    sweep = None
    if "sweep" in sections:
        sweep = []
        Ss = 10_000_000
        osw = wrs.empty(Ss, np.uint32)
        for nn in (100_000, 1_000_000, 10_000_000, 100_000_000):
            wsw = wrs.empty(nn, np.float32)
            wrs.generate(wrs.GEN_UNIFORM01, nn, 0, wsw, stream=st)
            st.synchronize()
            tsw_tune = None
            if not args.no_tune:
                tsw_tune, _ = wrs.autotune_psa(wsw, nn, precision=args.precision,
                                               split_size=args.split_size, stream=st)
            pl = wrs.Plan(nn, args.split_size, args.precision, tune=tsw_tune)
            tsw = wrs.empty(nn, wrs.ALIAS_DTYPE)
            pl.psa_build(wsw, nn, tsw, stream=st)
            b_ms = time_warm(lambda: pl.psa_build(wsw, nn, tsw, stream=st), st, wrs, 5)
            a_ms = time_warm(lambda: wrs.alias_sample(tsw, nn, Ss, 5, osw, stream=st), st, wrs, 5)
            hi_ = wrs.empty(nn, np.float32)
            lo_ = wrs.empty(nn, np.float32)
            pl.its_build(wsw, nn, hi_, lo_, stream=st)
            i_ms = time_warm(lambda: wrs.its_sample(hi_, lo_, nn, Ss, 5, osw, stream=st), st, wrs, 3)
            gd_ = wrs.empty(nn + 1, np.uint32)
            wrs.cutpoint_build(hi_, lo_, nn, gd_, nn, stream=st)
            c_ms = time_warm(lambda: wrs.cutpoint_sample(hi_, lo_, nn, gd_, nn, Ss, 5, osw, stream=st),
                           st, wrs, 3)
            # windowed samplers, window 128 (the reference's cooperative size)
            aw_ms = time_warm(lambda: wrs.alias_sample_window(tsw, nn, Ss, 5, 128, osw, stream=st),
                            st, wrs, 5)
            iw_ms = time_warm(lambda: wrs.its_sample_window(hi_, lo_, nn, Ss, 5, 128, osw, stream=st),
                            st, wrs, 3)
            # the same with the caches flushed before every call
            bc = time_cold(lambda: pl.psa_build(wsw, nn, tsw, stream=st), st, wrs, 5, flush)
            ac = time_cold(lambda: wrs.alias_sample(tsw, nn, Ss, 5, osw, stream=st), st, wrs, 5,
                           flush)
            ic = time_cold(lambda: wrs.its_sample(hi_, lo_, nn, Ss, 5, osw, stream=st), st, wrs,
                           3, flush)
            cc = time_cold(lambda: wrs.cutpoint_sample(hi_, lo_, nn, gd_, nn, Ss, 5, osw,
                                                       stream=st), st, wrs, 3, flush)
            row = dict(n=nn, psa_build_ms=round(b_ms, 4), alias_sample_ms=round(a_ms, 4),
                       alias_build_plus_sample_per_s=round(Ss / ((b_ms + a_ms) * 1e-3), 1),
                       its_sample_per_s=round(Ss / (i_ms * 1e-3), 1),
                       cutpoint_sample_per_s=round(Ss / (c_ms * 1e-3), 1),
                       alias_window128_sample_per_s=round(Ss / (aw_ms * 1e-3), 1),
                       its_window128_sample_per_s=round(Ss / (iw_ms * 1e-3), 1),
                       cold=dict(alias_build_plus_sample_per_s=round(Ss / ((bc + ac) * 1e-3), 1),
                                 its_sample_per_s=round(Ss / (ic * 1e-3), 1),
                                 cutpoint_sample_per_s=round(Ss / (cc * 1e-3), 1)))
            sweep.append(row)
            log(f"sweep: {row}")
            del wsw, tsw, hi_, lo_, gd_, pl
        del osw

    # ---- CPU baseline: the reference CPU path restated (oracle) ------------
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.psa_alias_table_f32(w_host[: 1 << 20], (1 << 20) // args.split_size, True)
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            O.psa_alias_table_f32(w_host, K, True)
            times.append(time.perf_counter() - t0)
        t_cpu = min(times)
        cpu = dict(value=round(B_build / t_cpu / 1e9, 3), unit="GB/s", cores=1, kind="port",
                   sample=f"full workload: PSA build of the same N=2^{args.log2n} Zipf weights by "
                          f"the restated reference CPU path (float, 1e-4 scale, reference split), "
                          f"best of 5 after a warm-up, {t_cpu * 1e3:.0f} ms; host {cpu_model()}, "
                          f"{os.cpu_count()} logical CPUs visible")
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"PSA alias build, N=2^{args.log2n} Zipf(1.1) FP32 weights over a seeded "
                        f"rank permutation (seed 2), split size {args.split_size}, "
                        f"{PREC_NAMES[args.precision]}",
            "n": n, "split_size": args.split_size, "B_build_bytes": B_build,
            "heavy_count": H,
            "parallelism": "single GPU",
            "tuning": tune,
            "tuning_candidates_ms": [[e["partition"], e["direct"], e["merge_tile"],
                                      e["merge_per"], round(e["ms"], 4)] for e in tune_log],
        },
        "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
        "cold": {"ms_per_step": round(cold_ms, 5), "gbs": round(cold_gbs, 2),
                 "frac_hbm_peak": round(cold_gbs / HBM_PEAK_GBS, 4),
                 "flush": "512 MiB copy before every build, median of the steps"},
        "lower_bound_12N_gbs": round(12 * n / (ms_step * 1e-3) / 1e9, 2),
        "roofline": {"bound": "hbm", "kernel": kname[dom], "achieved": round(dom_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes": alg[dom], "avg_ms": round(dom_ms, 5),
                     "traffic": traffic},
        "phases_ms": {k: round(float(v), 5) for k, v in phase.items()},
        "build_kernels": build_kernels,
        "copy_kernel_gbs": round(copy_gbs, 1),
        "samples_per_s": round(samples_per_s, 1),
        "sampling": {"S": S, "ms": round(sample_ms, 4),
                     "gbs_12B": round(12 * S / (sample_ms * 1e-3) / 1e9, 1),
                     "statistics": sstat},
        "max_abs_error_vs_fp64": None if err is None else err["max_abs"],
        "max_rel_error_vs_fp64": None if err is None else err["max_rel"],
        "precision_study": study,
        "c1": c1,
        "scan_2e26": scan26,
        "split_pack_baselines": split_pack,
        "its": its,
        "spike": spike,
        "batched": batched,
        "sampling_sweep_S1e7": sweep,
        "cpu_baseline": cpu,
        "device": info["name"] or info["arch"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
