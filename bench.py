"""Benchmark of the batched affine-gap alignment path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step"): the batched uneven DNA set -- 1,048,576 pairs, |H| ~
LogNormal(ln 300, 1.0) clipped to 100..10,000 bp, V = H with 8% substitutions
and 2% geometric indels, semi-global AlignConfig<true,false,false,true>,
Score<int,Simple>(2,-3,-1,-5) -- scored by one call of the library with the
residues already resident in HBM.  Prints one JSON line.

Also measured in the same run: the plain one-pair-per-launch GPU form on a
sample, and the reference SeqAn 2 CPU path (oracle/_ref, built from the
reference sources) on a bounded sample using all host cores; the GPU scores of
that sample are bit-checked against the reference's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import seqan_amd as sa  # noqa: E402

METRIC = "GCUPS exact affine DP, batched vs plain on 1 GPU; fwd+bwd GCUPS; % VALU roof"
SEED = 0x5EED0002
N_PAIRS = 1 << 20
MODE = sa.FREE_TOP | sa.FREE_BOTTOM
# Peak packed-int16 VALU element operations: 256 CU x 64 lanes x 2 halves x 2.4 GHz.
PEAK_TOPS = 256 * 64 * 2 * 2.4e9 / 1e12
OPS_PER_CELL = 8          # Hd+s, E+e, F+e, H+o, max E, max F, 2x max H


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(H, V, score, count, gpu_scores):
    """Reference SeqAn 2 ExecutionPolicy<Parallel,Vectorial> on the first
    `count` pairs (all host cores); returns (dict, parity_ok)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if po.ref_binary() is None:
        return None, None
    hs = sa.SeqSet(H.residues[:H.offsets[count]], H.offsets[:count + 1])
    vs = sa.SeqSet(V.residues[:V.offsets[count]], V.offsets[:count + 1])
    cores = host_cores()
    r = po.run_reference(po.OP_SCORE, hs.residues, hs.offsets, vs.residues, vs.offsets, MODE, alphabet="dna",
                         match=2, mismatch=-3, gap_extend=-1, gap_open=-5, threads=cores, repeats=2)
    cells = int((hs.lengths() * vs.lengths()).sum())
    ok = bool(np.array_equal(r["scores"], gpu_scores[:count]))
    return {"value": round(cells / r["best_seconds"] / 1e9, 3), "unit": "GCUPS", "cores": int(r["threads"]),
            "kind": "reference",
            "sample": f"first {count} pairs ({cells:.3e} cells) of the same workload; "
                      f"{r['binary']} ({r['simd_lanes']} int32 lanes), best of 2 timed calls",
            "parity_vs_gpu": ok}, ok


def summarize_rocpd(paths):
    """Prints per-kernel duration statistics and counter totals of rocprofv3
    result databases (the rocpd SQLite output)."""
    import sqlite3
    for path in paths:
        db = sqlite3.connect(path)
        print(f"# {path}")
        rows = db.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                          "from kernels group by name order by sum(duration) desc").fetchall()
        print(f"{'calls':>6} {'total_ms':>10} {'avg_ms':>10} {'min_ms':>10} {'max_ms':>10}  kernel")
        for name, n, tot, avg, mn, mx in rows:
            print(f"{n:6d} {tot / 1e6:10.3f} {avg / 1e6:10.3f} {mn / 1e6:10.3f} {mx / 1e6:10.3f}  {name[:110]}")
        try:
            pm = db.execute("select k.name, p.counter_name, count(*), sum(p.counter_value) from pmc_events p "
                            "join kernels k on k.dispatch_id = p.dispatch_id group by k.name, p.counter_name "
                            "order by k.name").fetchall()
        except sqlite3.Error:
            pm = []
        if pm:
            print(f"{'dispatches':>10} {'counter':>18} {'sum':>18} {'per_dispatch':>18}  kernel")
            for name, cname, n, tot in pm:
                print(f"{n:10d} {cname:>18} {tot:18.1f} {tot / n:18.1f}  {name[:90]}")
        print()


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--summarize-rocpd":
        summarize_rocpd(sys.argv[2:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=N_PAIRS)
    ap.add_argument("--cpu-sample", type=int, default=N_PAIRS // 8)
    ap.add_argument("--plain-sample", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also copy this many MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)

    t0 = time.time()
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, SEED, 0, args.pairs)
    cells = int((H.lengths() * V.lengths()).sum())
    log(f"generated {args.pairs} pairs, {cells:.4e} cells in {time.time() - t0:.1f}s")
    score = sa.simple_score(2, -3, -1, -5)

    ctx = sa.Context(0)
    dh = sa.DeviceBuffer.from_array(H.residues)
    dv = sa.DeviceBuffer.from_array(V.residues)
    out = sa.DeviceBuffer(4 * args.pairs)
    sa.synchronize()

    def step():
        ctx.align_score(H, V, score, MODE, h_dev=dh.address, v_dev=dv.address, scores_dev=out.address)

    for w in range(args.warmup):
        tw = time.perf_counter()
        step()
        log(f"warmup {w}: {time.perf_counter() - tw:.3f}s {ctx.stats()}")
    sa.synchronize()
    dev_ms, plan_ms = [], []
    t1 = time.perf_counter()
    for k in range(args.steps):
        step()
        st = ctx.stats()
        dev_ms.append(st["device_ms"])
        plan_ms.append(st["host_plan_ms"])
        log(f"step {k}: device {st['device_ms']:.1f} ms, plan {st['host_plan_ms']:.1f} ms")
    sa.synchronize()
    elapsed = time.perf_counter() - t1
    ms_per_step = elapsed * 1e3 / args.steps
    gcups = cells / (ms_per_step / 1e3) / 1e9
    scores = out.to_array(np.int32, args.pairs)
    st = ctx.stats()
    log(f"batched: {ms_per_step:.1f} ms/step, {gcups:.1f} GCUPS, device {np.mean(dev_ms):.1f} ms, "
        f"plan {np.mean(plan_ms):.1f} ms, stats {st}")

    kern_ms = float(np.mean(dev_ms))
    achieved = OPS_PER_CELL * cells / (kern_ms / 1e3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_TOPS, 2), "unit": "TOP/s",
                "frac": round(achieved / PEAK_TOPS, 4), "traffic": None,
                "note": "packed-int16 VALU element ops, 8 per algorithmic cell; kernel time from HIP events "
                        "on the library stream around the alignment kernels"}

    if args.copy_calib:
        nb = args.copy_calib << 20
        src, dst = sa.DeviceBuffer(nb), sa.DeviceBuffer(nb)
        sa.hip().hipMemcpy(dst.ptr, src.ptr, nb, 3)     # hipMemcpyDeviceToDevice
        sa.synchronize()
        log(f"copied {args.copy_calib} MiB device-to-device")

    # Plain form: one kernel launch per pair, same library, same GPU.
    ps = min(args.plain_sample, args.pairs)
    if ps == 0:
        print(json.dumps({"metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
                          "roofline": roofline, "note": "profiling run: plain and CPU legs skipped"}))
        return
    idx = np.arange(0, args.pairs, args.pairs // ps)[:ps]
    hp = sa.SeqSet.from_list([H[i] for i in idx])
    vp = sa.SeqSet.from_list([V[i] for i in idx])
    pcells = int((hp.lengths() * vp.lengths()).sum())
    warm = sa.SeqSet.from_list([H[i] for i in idx[:8]]), sa.SeqSet.from_list([V[i] for i in idx[:8]])
    ctx.align_score(warm[0], warm[1], score, MODE, flags=sa.PLAIN_PER_PAIR)
    plain_scores = np.empty(ps, np.int32)
    plain_s = 0.0
    chunk = 256
    for c0 in range(0, ps, chunk):
        sub_h = sa.SeqSet.from_list([H[i] for i in idx[c0:c0 + chunk]])
        sub_v = sa.SeqSet.from_list([V[i] for i in idx[c0:c0 + chunk]])
        tp = time.perf_counter()
        plain_scores[c0:c0 + chunk] = ctx.align_score(sub_h, sub_v, score, MODE, flags=sa.PLAIN_PER_PAIR)
        plain_s += time.perf_counter() - tp
        log(f"plain: {min(c0 + chunk, ps)}/{ps} pairs, {plain_s:.2f}s")
    plain_gcups = pcells / plain_s / 1e9
    plain_ok = bool(np.array_equal(plain_scores, scores[idx]))
    log(f"plain: {ps} pairs, {plain_gcups:.2f} GCUPS, parity {plain_ok}")

    cpu = None
    if not args.no_cpu:
        cpu, ok = cpu_baseline(H, V, score, min(args.cpu_sample, args.pairs), scores)
        log(f"cpu baseline: {cpu}")

    res = {
        "metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i16", "data": "synthetic",
        "config": {"workload": "batched uneven DNA: log-normal(300, 1.0) 100-10,000 bp, 8% subst, 2% geometric "
                               "indels, semi-global AlignConfig<1,0,0,1>, Score<int,Simple>(2,-3,-1,-5)",
                   "pairs": args.pairs, "cells": cells, "seed": SEED},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": {"value": round(plain_gcups, 2), "unit": "GCUPS", "sample_pairs": ps,
                  "batched_over_plain": round(gcups / plain_gcups, 1), "parity_vs_batched": plain_ok},
        "host_plan_ms": round(float(np.mean(plan_ms)), 2),
        "pairs_int32": int(st["pairs_int32"]),
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
