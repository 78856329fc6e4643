#!/usr/bin/env python3
"""Benchmark of the weighted-random-sampling hot path on one MI355X.

One step = one PSA alias-table build (FP64 total -> heavy/light prefix-partition
-> block-boundary splits -> merge pack) of N = 2^26 Zipf(1.1) FP32 weights (ranks a seeded
permutation of 1..N, seed 2), FP64 tile sums with the exact fixed-point carry,
split size 32, inputs resident in HBM.  The value is the build throughput in
GB/s of algorithmic bytes B_build = 32N + 12K (K = N/32): mean read 4N,
partition read 4N + index 4N + prefix 4N, split 12K, pack index 4N + weight 4N
+ table 8N.

Also reported on the same line: the fraction of the 8 TB/s HBM peak, the
dominant kernel's roofline (algorithmic bytes per launch / its average
duration, from HIP events on the stream the kernels run on), a copy-kernel
calibration, alias-sampling samples/s (S = 2^26, 12 B/sample), the implied
distribution's max relative error vs FP64, the batched 4096-distribution build
+ sample against per-distribution launches, and the reference CPU path (the
oracle's restatement, single-threaded) on the same weights.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "merian-wrs_amd"))

HBM_PEAK_GBS = 8000.0
PREC_NAMES = ['fp32', 'fp64-carry', 'compensated-fp32', 'fp64-tiles']
METRIC = "PSA build GB/s (% HBM peak) + samples/s; batched vs per-launch at fixed error"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def implied_error(table: np.ndarray, w: np.ndarray) -> dict:
    n = w.size
    p = np.clip(table["p"].astype(np.float64), 0.0, 1.0)
    q = np.bincount(np.arange(n), weights=p, minlength=n)
    q += np.bincount(table["a"].astype(np.int64), weights=1.0 - p, minlength=n)
    q /= n
    w64 = w.astype(np.float64)
    t = w64 / np.sum(np.sort(w64))  # sorted pairwise sum: relative error ~1e-16
    d = np.abs(q - t)
    nz = t > 0
    return dict(max_abs=float(d.max()), max_rel=float((d[nz] / t[nz]).max()))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=26)
    ap.add_argument("--split-size", type=int, default=32)
    ap.add_argument("--precision", type=int, default=2)
    ap.add_argument("--log2s", type=int, default=26)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    args = ap.parse_args()

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
    plan = wrs.Plan(n, args.split_size, args.precision)
    K = plan.split_count(n)
    table = wrs.empty(n, wrs.ALIAS_DTYPE)
    st.synchronize()
    B_build = 32 * n + 12 * K

    # ---- warmup + timed region --------------------------------------------
    for _ in range(args.warmup):
        plan.psa_build(w, n, table, stream=st)
    st.synchronize()
    e0, e1 = wrs.Event(), wrs.Event()
    e0.record(st)
    for _ in range(args.steps):
        plan.psa_build(w, n, table, stream=st)
    e1.record(st)
    st.synchronize()
    ms_total = e0.elapsed_ms(e1)
    ms_step = ms_total / args.steps
    gbs = B_build / (ms_step * 1e-3) / 1e9
    log(f"PSA build N=2^{args.log2n}: {ms_step:.4f} ms/step, {gbs:.1f} GB/s "
        f"({100 * gbs / HBM_PEAK_GBS:.1f}% of peak)")

    # ---- per-kernel durations (events on the same stream) ------------------
    plan.set_timing(True)
    acc = np.zeros(5)
    for _ in range(args.steps):
        plan.psa_build(w, n, table, stream=st)
        acc += np.array(plan.phase_times())
    plan.set_timing(False)
    phase = acc / args.steps
    names = ["reduce", "finalize", "partition", "coarse_split", "merge_pack"]
    # bytes each kernel must move: reduce reads w; the partition reads w and
    # writes the indices (its FP64 prefix samples are 16N/64 more); the merge
    # pack reads the indices, gathers the weights and writes the 8-byte table
    alg = {"reduce": 4 * n, "finalize": 0, "partition": 8 * n + 16 * (n // 64),
           "coarse_split": 0, "merge_pack": 16 * n}
    log("phases (ms): " + ", ".join(f"{k} {v:.4f}" for k, v in zip(names, phase)))
    dom = max(("reduce", "partition", "merge_pack"), key=lambda k: phase[names.index(k)])
    dom_ms = float(phase[names.index(dom)])
    dom_gbs = alg[dom] / (dom_ms * 1e-3) / 1e9

    # ---- copy calibration ---------------------------------------------------
    cdst = wrs.empty(n, np.float32)
    wrs.copy_f32(cdst, w, n, stream=st)
    c0, c1 = wrs.Event(), wrs.Event()
    c0.record(st)
    for _ in range(10):
        wrs.copy_f32(cdst, w, n, stream=st)
    c1.record(st)
    st.synchronize()
    copy_gbs = 2 * 4 * n / (c0.elapsed_ms(c1) / 10 * 1e-3) / 1e9
    del cdst
    log(f"copy kernel: {copy_gbs:.1f} GB/s")

    # ---- alias sampling -----------------------------------------------------
    smp = wrs.empty(S, np.uint32)
    wrs.alias_sample(table, n, S, 12345, smp, stream=st)
    s0, s1 = wrs.Event(), wrs.Event()
    s0.record(st)
    reps = 5
    for r in range(reps):
        wrs.alias_sample(table, n, S, 12345, smp, counter_base=r * S, stream=st)
    s1.record(st)
    st.synchronize()
    sample_ms = s0.elapsed_ms(s1) / reps
    samples_per_s = S / (sample_ms * 1e-3)
    log(f"alias sampling S=2^{args.log2s}: {sample_ms:.4f} ms, {samples_per_s / 1e9:.2f} Gsamples/s")
    del smp

    # ---- accuracy of the benchmarked table ----------------------------------
    w_host = w.numpy()
    err = None
    if not args.no_error:
        err = implied_error(table.numpy(), w_host)
        log(f"implied distribution vs FP64: max_rel {err['max_rel']:.3e}")

    # ---- batched 4096-distribution build + sample --------------------------
    batched = None
    if not args.no_batched:
        nseg, per = 4096, 64
        sizes = wrs.segment_sizes(3, nseg)
        off = wrs.segment_offsets(sizes)
        tot = int(off[-1])
        bw = wrs.empty(tot, np.float32)
        wrs.generate(wrs.GEN_LOGNORMAL, tot, 3, bw, a=0.0, b=3.0, stream=st)
        doff = wrs.to_device(off)
        bp = wrs.BatchPlan(off, args.split_size, args.precision)
        tb = wrs.empty(tot, wrs.ALIAS_DTYPE)
        sb = wrs.empty(nseg * per, np.uint32)
        g = wrs.Graph(st)
        with g:
            bp.build(bw, tb, stream=st)
            wrs.alias_sample_batched(tb, doff, nseg, per, 42, sb, stream=st)
        g.launch()
        st.synchronize()
        b0, b1 = wrs.Event(), wrs.Event()
        b0.record(st)
        for _ in range(10):
            g.launch()
        b1.record(st)
        st.synchronize()
        batched_ms = b0.elapsed_ms(b1) / 10
        eplan = wrs.Plan(int(sizes.max()), args.split_size, args.precision)
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
                       batched_build_gbs=round((32 * tot + 12 * int(np.sum(-(-sizes // args.split_size))))
                                               / (batched_ms * 1e-3) / 1e9, 1))
        log(f"batched: {batched}")
        del bw, tb, sb, te, se

    # ---- CPU baseline: the reference CPU path restated (oracle) ------------
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.psa_alias_table_f32(w_host[: 1 << 20], (1 << 20) // args.split_size, True)
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            O.psa_alias_table_f32(w_host, K, True)
            times.append(time.perf_counter() - t0)
        t_cpu = min(times)
        cpu = dict(value=round(B_build / t_cpu / 1e9, 3), unit="GB/s", cores=1, kind="port",
                   sample=f"full workload: PSA build of the same N=2^{args.log2n} Zipf weights by "
                          f"the restated reference CPU path (float, 1e-4 scale, reference split), "
                          f"best of 2 after a warm-up, {t_cpu * 1e3:.0f} ms; host {cpu_model()}, "
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
            "parallelism": "single GPU",
        },
        "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
        "lower_bound_12N_gbs": round(12 * n / (ms_step * 1e-3) / 1e9, 2),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes": alg[dom], "avg_ms": round(dom_ms, 5), "traffic": None},
        "phases_ms": {k: round(float(v), 5) for k, v in zip(names, phase)},
        "copy_kernel_gbs": round(copy_gbs, 1),
        "samples_per_s": round(samples_per_s, 1),
        "sampling": {"S": S, "ms": round(sample_ms, 4), "gbs_12B": round(12 * S / (sample_ms * 1e-3) / 1e9, 1)},
        "max_rel_error_vs_fp64": None if err is None else err["max_rel"],
        "batched": batched,
        "cpu_baseline": cpu,
        "device": info["name"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
