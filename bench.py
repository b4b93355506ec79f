"""Benchmark of the batched affine-gap alignment path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step"): the batched uneven DNA set -- 1,048,576 pairs, |H| ~
LogNormal(ln 300, 1.0) clipped to 100..10,000 bp, V = H with 8% substitutions
and 2% geometric indels, semi-global AlignConfig<true,false,false,true>,
Score<int,Simple>(2,-3,-1,-5) -- scored by one call of the library with the
residues already resident in HBM.  Prints one JSON line.

Also measured in the same run, each reported beside the headline figure:
  - the plain one-pair-per-launch GPU form on a sample of the same pairs;
  - config 1: 4,096 DNA pairs of 150 bp, global: batched (device and host
    buffers, with per-call host time), plain per-pair form, reference CPU;
  - fwd+bwd: score plus traceback counts (sqmi_align_counts) of 65,536 DNA
    pairs of 1 kbp, global, and the forward alone on the same pairs;
  - protein local: 262,144 Smith-Waterman pairs, Blosum62(-1, -11);
  - long fwd+bwd: score plus traceback counts of 64 DNA pairs of 100 kbp
    (tiled int32 kernels with checkpoints), and the forward alone;
  - banded: 1,048,576 DNA pairs of 150 bp (8% substitutions), global, band
    [-16, 16] (int32 lane-per-diagonal kernel), GCUPS over the band's cells;
  - the reference SeqAn 2 CPU path (oracle/_ref, built from the reference
    sources) on bounded samples of each workload, all host cores and one
    core; the GPU results of each sample are bit-checked against it.
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
SEED_DNA150, N_DNA150 = 0x5EED0001, 4096
SEED_PROTEIN, N_PROTEIN = 0x5EED0003, 1 << 18
SEED_DNA1K, N_DNA1K = 0x5EED0004, 1 << 16
SEED_DNA100K, N_DNA100K = 0x5EED0005, 64
SEED_BANDED, N_BANDED, L_BANDED, BAND = 0x5EED0006, 1 << 20, 150, (-16, 16)
# Peak int32 VALU operations (one element per lane): 256 CU x 64 lanes x 2.4 GHz.
PEAK_TOPS_INT32 = 256 * 64 * 2.4e9 / 1e12
# Peak packed-int16 VALU element operations: 256 CU x 64 lanes x 2 halves x 2.4 GHz.
PEAK_TOPS = 256 * 64 * 2 * 2.4e9 / 1e12
OPS_PER_CELL = 8          # Hd+s, E+e, F+e, H+o, max E, max F, 2x max H
OPS_PER_CELL_LOCAL = 9    # + max(H, 0)
PROFILE = os.path.join(ROOT, "profiles", "fwd16_dna_lognormal_semiglobal.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def head(S, count):
    return sa.SeqSet(S.residues[:S.offsets[count]], S.offsets[:count + 1])


def reference_run(op, H, V, count, mode, threads, repeats, gpu_scores, gpu_counts=None, **kw):
    """The reference SeqAn 2 CPU path (ExecutionPolicy<Parallel,Vectorial>)
    on the first `count` pairs; returns a baseline dict (GCUPS, parity with
    the GPU results of the same pairs) or None when oracle/_ref is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if po.ref_binary() is None:
        return None
    hs, vs = head(H, count), head(V, count)
    r = po.run_reference(op, hs.residues, hs.offsets, vs.residues, vs.offsets, mode, threads=threads,
                         repeats=repeats, **kw)
    cells = int((hs.lengths() * vs.lengths()).sum())
    ok = bool(np.array_equal(r["scores"], gpu_scores[:count]))
    if gpu_counts is not None:
        ok = ok and bool(np.array_equal(r["counts"], gpu_counts[:count]))
    return {"value": round(cells / r["best_seconds"] / 1e9, 3), "unit": "GCUPS", "cores": int(r["threads"]),
            "kind": "reference",
            "sample": f"first {count} pairs ({cells:.3e} cells) of the same workload; {r['binary']} "
                      f"({r['simd_lanes']} int32 lanes), best of {repeats} timed call(s); {cpu_model()}",
            "parity_vs_gpu": ok}


def shape_provenance():
    """Where the planner's shape table came from (its JSON's provenance)."""
    try:
        with open(os.path.join(ROOT, "seqan_amd", "tuning", "shapes.json")) as f:
            return json.load(f).get("provenance")
    except (OSError, ValueError):
        return None


def profiled_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    try:
        with open(PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def timed(fn, steps, warmup, label):
    for _ in range(warmup):
        fn()
    sa.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        fn()
    sa.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    log(f"{label}: {ms:.2f} ms/step")
    return ms


def config1_leg(ctx, steps, warmup, cpu):
    """4,096 DNA pairs of 150 bp, global, Score<int,Simple>(2,-3,-1,-5): the
    batched call with residues and scores resident in HBM, the same call
    from host buffers (PCIe copies inside), the plain form (one launch per
    pair) on the same pairs, and the reference CPU path on the same pairs.
    Per-call host time (planning and launch) is reported beside the device
    time: at 9.2e7 cells a call is dominated by it."""
    H, V = sa.workload(sa.WL_DNA150, SEED_DNA150, 0, N_DNA150)
    cells = int((H.lengths() * V.lengths()).sum())
    score = sa.simple_score(2, -3, -1, -5)
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    ds = sa.DeviceBuffer(4 * N_DNA150)
    dev_ms, plan_ms, call_ms = [], [], []

    def step():
        t = time.perf_counter()
        ctx.align_score(H, V, score, 0, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
        call_ms.append((time.perf_counter() - t) * 1e3)
        st = ctx.stats()
        dev_ms.append(st["device_ms"])
        plan_ms.append(st["host_plan_ms"])

    os.environ["SQMI_GRAPH"] = "0"      # every call plans afresh
    ms = timed(step, steps, warmup, "config1 batched")
    scores = ds.to_array(np.int32, N_DNA150)
    n = steps
    os.environ.pop("SQMI_GRAPH", None)
    rep_call = []

    def replay_step():
        t = time.perf_counter()
        ctx.align_score(H, V, score, 0, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
        rep_call.append((time.perf_counter() - t) * 1e3)

    rep_ms = timed(replay_step, steps, 2, "config1 batched, replayed graph")
    replay = {"ms_per_step": round(rep_ms, 4), "gcups": round(cells / (rep_ms / 1e3) / 1e9, 2),
              "call_ms": round(float(np.median(rep_call[-steps:])), 4), "replayed": int(ctx.stats()["replayed"]),
              "device_ms": round(ctx.stats()["device_ms"], 4)}
    kern = float(np.median(dev_ms[-n:]))
    achieved = OPS_PER_CELL * cells / (kern / 1e3) / 1e12
    host_ms = timed(lambda: ctx.align_score(H, V, score, 0), steps, warmup, "config1 batched, host buffers")
    host_scores = ctx.align_score(H, V, score, 0)
    # Plain form: one kernel launch per pair, same pairs, same library.
    t = time.perf_counter()
    plain_scores = ctx.align_score(H, V, score, 0, flags=sa.PLAIN_PER_PAIR)
    plain_s = time.perf_counter() - t
    res = {"value": round(cells / (ms / 1e3) / 1e9, 2), "unit": "GCUPS", "ms_per_step": round(ms, 4),
           "device_ms": round(kern, 4), "host_plan_ms": round(float(np.median(plan_ms[-n:])), 4),
           "call_ms": round(float(np.median(call_ms[-n:])), 4),
           "replay": replay,
           "host_buffers_ms_per_step": round(host_ms, 4),
           "host_buffers_gcups": round(cells / (host_ms / 1e3) / 1e9, 2),
           "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_TOPS, 2),
                        "unit": "TOP/s", "frac": round(achieved / PEAK_TOPS, 4),
                        "note": "packed-int16 VALU ops, 8 per cell, over the median HIP-event time of one call"},
           "plain": {"value": round(cells / plain_s / 1e9, 3), "unit": "GCUPS", "ms_per_call": round(plain_s * 1e3, 2),
                     "launches": N_DNA150, "batched_over_plain": round((cells / (ms / 1e3)) / (cells / plain_s), 1),
                     "parity_vs_batched": bool(np.array_equal(plain_scores, scores))},
           "parity_host_vs_device_buffers": bool(np.array_equal(host_scores, scores)),
           "workload": "4,096 DNA pairs of 150 bp, 5% subst, 1% single-base indels, global, "
                       "Score<int,Simple>(2,-3,-1,-5)", "pairs": N_DNA150, "cells": cells, "seed": SEED_DNA150}
    if cpu:
        res["cpu_baseline"] = reference_run(0, H, V, N_DNA150, 0, host_cores(), 3, scores, alphabet="dna", match=2,
                                            mismatch=-3, gap_extend=-1, gap_open=-5)
        if res["cpu_baseline"] is not None:
            one = reference_run(0, H, V, N_DNA150, 0, 1, 3, scores, alphabet="dna", match=2, mismatch=-3,
                                gap_extend=-1, gap_open=-5)
            res["cpu_baseline"]["one_core"] = {"value": one["value"], "parity_vs_gpu": one["parity_vs_gpu"]}
            res["cpu_baseline"]["sample"] = ("all 4,096 pairs (the whole workload); warm-up call, best of 3 timed "
                                             f"calls; {cpu_model()}")
        log(f"config1 cpu baseline: {res['cpu_baseline']}")
    for b in (dh, dv, ds):
        b.free()
    return res


def backward_leg(ctx, steps, warmup, cpu):
    """Score + traceback counts of the 1 kbp global workload, beside the
    forward alone on the same pairs; inputs and outputs resident in HBM."""
    H, V = sa.workload(sa.WL_DNA1K, SEED_DNA1K, 0, N_DNA1K)
    cells = int((H.lengths() * V.lengths()).sum())
    score = sa.simple_score(2, -3, -1, -5)
    S = score.alphabet_size ** 2 + 2
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    ds, dc = sa.DeviceBuffer(4 * N_DNA1K), sa.DeviceBuffer(4 * S * N_DNA1K)
    dev = dict(h_dev=dh.address, v_dev=dv.address)
    fwd_ms = timed(lambda: ctx.align_score(H, V, score, 0, scores_dev=ds.address, **dev), steps, warmup,
                   "dna1k forward")
    fwd_dev = ctx.stats()["device_ms"]
    bwd_ms = timed(lambda: ctx.align_counts(H, V, score, 0, scores_dev=ds.address, counts_dev=dc.address, **dev),
                   steps, warmup, "dna1k fwd+bwd")
    st = ctx.stats()
    scores = ds.to_array(np.int32, N_DNA1K)
    counts = dc.to_array(np.int32, S * N_DNA1K).reshape(N_DNA1K, S)
    par = np.concatenate([score.table.ravel(), [score.gap_open, score.gap_extend]]).astype(np.int64)
    ident = bool(np.array_equal(counts.astype(np.int64) @ par, scores.astype(np.int64)))
    res = {"value": round(cells / (bwd_ms / 1e3) / 1e9, 2), "unit": "GCUPS", "ms_per_step": round(bwd_ms, 3),
           "forward_ms_per_step": round(fwd_ms, 3), "backward_over_forward": round(bwd_ms / fwd_ms, 3),
           "device_ms": round(st["device_ms"], 3), "forward_device_ms": round(fwd_dev, 3),
           "workload": "65,536 DNA pairs of 1 kbp, 10% subst, 2% indels, global, Score<int,Simple>(2,-3,-1,-5)",
           "pairs": N_DNA1K, "cells": cells, "seed": SEED_DNA1K, "sum_count_param_eq_score": ident}
    if cpu:
        res["cpu_baseline"] = reference_run(1, H, V, 4096, 0, host_cores(), 1, scores, counts, alphabet="dna",
                                            match=2, mismatch=-3, gap_extend=-1, gap_open=-5)
        log(f"dna1k cpu baseline: {res['cpu_baseline']}")
    # The same pairs under AlignConfig<true,false,false,true> (free first
    # and last row): the trace kernel tracks the end cell on the bottom row.
    semi = sa.FREE_TOP | sa.FREE_BOTTOM
    sfwd = timed(lambda: ctx.align_score(H, V, score, semi, scores_dev=ds.address, **dev), steps, warmup,
                 "dna1k semi-global forward")
    sbwd = timed(lambda: ctx.align_counts(H, V, score, semi, scores_dev=ds.address, counts_dev=dc.address, **dev),
                 steps, warmup, "dna1k semi-global fwd+bwd")
    st = ctx.stats()
    res["semi_global"] = {"value": round(cells / (sbwd / 1e3) / 1e9, 2), "unit": "GCUPS",
                          "ms_per_step": round(sbwd, 3), "forward_ms_per_step": round(sfwd, 3),
                          "backward_over_forward": round(sbwd / sfwd, 3), "pairs_int16": int(st["pairs_int16"]),
                          "workload": "the same pairs, AlignConfig<true,false,false,true>"}
    if cpu:
        sscores = ds.to_array(np.int32, N_DNA1K)
        scounts = dc.to_array(np.int32, S * N_DNA1K).reshape(N_DNA1K, S)
        res["semi_global"]["cpu_baseline"] = reference_run(1, H, V, 4096, semi, host_cores(), 1, sscores, scounts,
                                                           alphabet="dna", match=2, mismatch=-3, gap_extend=-1,
                                                           gap_open=-5)
    # Local alignment of the same pairs: the trace kernel tracks the end cell
    # over every cell and the walk stops where H falls to 0.
    lfwd = timed(lambda: ctx.align_score(H, V, score, sa.LOCAL, scores_dev=ds.address, **dev), steps, warmup,
                 "dna1k local forward")
    lbwd = timed(lambda: ctx.align_counts(H, V, score, sa.LOCAL, scores_dev=ds.address, counts_dev=dc.address,
                                          **dev), steps, warmup, "dna1k local fwd+bwd")
    st = ctx.stats()
    res["local"] = {"value": round(cells / (lbwd / 1e3) / 1e9, 2), "unit": "GCUPS",
                    "ms_per_step": round(lbwd, 3), "forward_ms_per_step": round(lfwd, 3),
                    "backward_over_forward": round(lbwd / lfwd, 3), "pairs_int16": int(st["pairs_int16"]),
                    "workload": "the same pairs, localAlignment counts"}
    if cpu:
        lscores = ds.to_array(np.int32, N_DNA1K)
        lcounts = dc.to_array(np.int32, S * N_DNA1K).reshape(N_DNA1K, S)
        res["local"]["cpu_baseline"] = reference_run(1, H, V, 4096, sa.LOCAL, host_cores(), 1, lscores, lcounts,
                                                     alphabet="dna", match=2, mismatch=-3, gap_extend=-1,
                                                     gap_open=-5)
    for b in (dh, dv, ds, dc):
        b.free()
    return res


def long_leg(ctx, steps, warmup, cpu):
    """Score, and score + traceback counts, of 64 DNA pairs of 100 kbp
    (global): the tiled int32 long-pair kernels, inputs and outputs resident
    in HBM."""
    H, V = sa.workload(sa.WL_DNA100K, SEED_DNA100K, 0, N_DNA100K)
    cells = int((H.lengths() * V.lengths()).sum())
    score = sa.simple_score(2, -3, -1, -5)
    S = score.alphabet_size ** 2 + 2
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    ds, dc = sa.DeviceBuffer(4 * N_DNA100K), sa.DeviceBuffer(4 * S * N_DNA100K)
    dev = dict(h_dev=dh.address, v_dev=dv.address)
    fwd_ms = timed(lambda: ctx.align_score(H, V, score, 0, scores_dev=ds.address, **dev), steps, warmup,
                   "dna100k forward")
    st_f = ctx.stats()
    bwd_ms = timed(lambda: ctx.align_counts(H, V, score, 0, scores_dev=ds.address, counts_dev=dc.address, **dev),
                   steps, warmup, "dna100k fwd+bwd")
    st = ctx.stats()
    scores = ds.to_array(np.int32, N_DNA100K)
    counts = dc.to_array(np.int32, S * N_DNA100K).reshape(N_DNA100K, S).astype(np.int64)
    par = np.concatenate([score.table.ravel(), [score.gap_open, score.gap_extend]]).astype(np.int64)
    ident = bool(np.array_equal(counts @ par, scores.astype(np.int64)) and np.array_equal(
        2 * counts[:, :S - 2].sum(axis=1) + counts[:, S - 2] + counts[:, S - 1], H.lengths() + V.lengths()))
    achieved = OPS_PER_CELL * cells / (st_f["device_ms"] / 1e3) / 1e12
    packed = st_f["pairs_long_packed"] == N_DNA100K
    peak = PEAK_TOPS if packed else PEAK_TOPS_INT32
    res = {"value": round(cells / (bwd_ms / 1e3) / 1e9, 2), "unit": "GCUPS", "ms_per_step": round(bwd_ms, 3),
           "forward_ms_per_step": round(fwd_ms, 3), "forward_gcups": round(cells / (fwd_ms / 1e3) / 1e9, 2),
           "backward_over_forward": round(bwd_ms / fwd_ms, 3), "device_ms": round(st["device_ms"], 3),
           "forward_device_ms": round(st_f["device_ms"], 3),
           "forward_roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2),
                                "unit": "TOP/s", "frac": round(achieved / peak, 4),
                                "note": ("packed int16 forward (two pairs per lane, values relative to a "
                                         "per-tile base), 8 ops per cell" if packed else
                                         "int32 forward, 8 ops per cell, one cell per lane")},
           "workload": "64 DNA pairs of 100 kbp, 12% divergence, 3% geometric indels, global, "
                       "Score<int,Simple>(2,-3,-1,-5)",
           "pairs": N_DNA100K, "pairs_long": int(st["pairs_long"]),
           "pairs_long_packed": int(st_f["pairs_long_packed"]), "cells": cells, "seed": SEED_DNA100K,
           "counts_form_alignments": ident}
    # The same pairs' scores under a free-end AlignConfig and local alignment
    # (the long route's int32 forward and its result pass).
    modes = {}
    for name, mode in (("semi_global_1001", sa.FREE_TOP | sa.FREE_BOTTOM), ("local", sa.LOCAL)):
        ms = timed(lambda: ctx.align_score(H, V, score, mode, scores_dev=ds.address, **dev), 1, 1,
                   f"dna100k forward, {name}")
        stm = ctx.stats()
        ach = OPS_PER_CELL * cells / (stm["device_ms"] / 1e3) / 1e12
        modes[name] = {"ms_per_step": round(ms, 3), "gcups": round(cells / (ms / 1e3) / 1e9, 2),
                       "device_ms": round(stm["device_ms"], 3), "pairs_long": int(stm["pairs_long"]),
                       "frac_of_int32_roof": round(ach / PEAK_TOPS_INT32, 4)}
    res["forward_other_modes"] = modes
    if cpu:
        # The reference runs 16 pairs per SIMD batch on one thread; full-length
        # pairs would take minutes, so the sample is the first 16 pairs of the
        # same generator cut to 10 kbp.
        Hs, Vs = sa.workload(sa.WL_DNA100K, SEED_DNA100K, 0, 16, 10000)
        gs = ctx.align_score(Hs, Vs, score, 0)
        res["cpu_baseline"] = reference_run(0, Hs, Vs, 16, 0, host_cores(), 1, gs, alphabet="dna", match=2,
                                            mismatch=-3, gap_extend=-1, gap_open=-5)
        if res["cpu_baseline"] is not None:
            cs = int((Hs.lengths() * Vs.lengths()).sum())
            res["cpu_baseline"]["sample"] = (f"first 16 pairs of the 100 kbp generator (same seed) cut to 10 kbp "
                                             f"({cs:.3e} cells), one timed call; {cpu_model()}")
        log(f"dna100k cpu baseline: {res['cpu_baseline']}")
    for b in (dh, dv, ds, dc):
        b.free()
    return res


def protein_leg(ctx, steps, warmup, cpu):
    """Smith-Waterman scores of the protein workload (Blosum62(-1, -11))."""
    H, V = sa.workload(sa.WL_PROTEIN, SEED_PROTEIN, 0, N_PROTEIN)
    cells = int((H.lengths() * V.lengths()).sum())
    score = sa.blosum62(-1, -11)
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    ds = sa.DeviceBuffer(4 * N_PROTEIN)
    os.environ["SQMI_GRAPH"] = "0"
    ms = timed(lambda: ctx.align_score(H, V, score, sa.LOCAL, h_dev=dh.address, v_dev=dv.address,
                                       scores_dev=ds.address), steps, warmup, "protein local")
    st = ctx.stats()
    os.environ["SQMI_SHAPE_RULE"] = "fixed"
    ctx.align_score(H, V, score, sa.LOCAL, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
    fixed_dev = ctx.stats()["device_ms"]
    os.environ.pop("SQMI_SHAPE_RULE", None)
    os.environ.pop("SQMI_GRAPH", None)
    scores = ds.to_array(np.int32, N_PROTEIN)
    kern = st["device_ms"]
    achieved = OPS_PER_CELL_LOCAL * cells / (kern / 1e3) / 1e12
    res = {"value": round(cells / (ms / 1e3) / 1e9, 2), "unit": "GCUPS", "ms_per_step": round(ms, 3),
           "device_ms": round(kern, 3), "roofline_frac": round(achieved / PEAK_TOPS, 4),
           "shape_table": {"tuned_device_ms": round(kern, 3), "fixed_rule_device_ms": round(fixed_dev, 3)},
           "workload": "262,144 protein pairs, query U[50,1000] aa, target U[50,2000] aa, 30% homologs, "
                       "local, Blosum62(-1,-11)",
           "pairs": N_PROTEIN, "cells": cells, "seed": SEED_PROTEIN, "pairs_int32": int(st["pairs_int32"])}
    if cpu:
        res["cpu_baseline"] = reference_run(0, H, V, 16384, sa.LOCAL, host_cores(), 1, scores, alphabet="aa",
                                            matrix="blosum62", gap_extend=-1, gap_open=-11)
        log(f"protein cpu baseline: {res['cpu_baseline']}")
    for b in (dh, dv, ds):
        b.free()
    return res


def band_cells(n, m, lower, upper):
    """Cells (i, j), 1 <= i <= m, 1 <= j <= n, with lower <= j - i <= upper."""
    i = np.arange(1, m + 1)
    lo = np.maximum(1, i + lower)
    hi = np.minimum(n, i + upper)
    return int(np.clip(hi - lo + 1, 0, None).sum())


def banded_leg(ctx, steps, warmup, cpu):
    """Banded global scores: pairs of one length (the collections the
    reference's banded set path supports), substitutions only."""
    rng = np.random.default_rng(SEED_BANDED)
    h = rng.integers(0, 4, (N_BANDED, L_BANDED), dtype=np.uint8)
    v = h.copy()
    sub = rng.random(h.shape) < 0.08
    v[sub] = rng.integers(0, 4, int(sub.sum()), dtype=np.uint8)
    off = np.arange(N_BANDED + 1, dtype=np.int64) * L_BANDED
    H, V = sa.SeqSet(h.reshape(-1), off), sa.SeqSet(v.reshape(-1), off.copy())
    lo, up = BAND
    cells = band_cells(L_BANDED, L_BANDED, lo, up) * N_BANDED
    score = sa.simple_score(2, -3, -1, -5)
    scores = [None]
    ms = timed(lambda: scores.__setitem__(0, ctx.align_score_banded(H, V, score, lo, up)), steps, warmup, "banded")
    st = ctx.stats()
    kern = st["device_ms"]
    achieved = OPS_PER_CELL * cells / (kern / 1e3) / 1e12

    def band_roof(stats):
        # The packed int16 kernel's roof when every pair ran on it, else int32's.
        if stats["pairs_int16"] == stats["pairs"]:
            return PEAK_TOPS, "packed int16 VALU element ops (two pairs per lane), 8 per cell inside the band"
        return PEAK_TOPS_INT32, "int32 VALU ops, 8 per cell inside the band"

    peak, roof_note = band_roof(st)
    res = {"value": round(cells / (ms / 1e3) / 1e9, 2), "unit": "GCUPS (cells inside the band)",
           "ms_per_step": round(ms, 3), "device_ms": round(kern, 3),
           "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2),
                        "unit": "TOP/s", "frac": round(achieved / peak, 4),
                        "note": roof_note + "; kernel time from HIP events"},
           "pairs_packed": int(st["pairs_int16"]),
           "workload": f"{N_BANDED:,} DNA pairs of {L_BANDED} bp, 8% substitutions, global, band [{lo}, {up}], "
                       "Score<int,Simple>(2,-3,-1,-5); host residues (PCIe copies inside the step)",
           "pairs": N_BANDED, "cells": cells, "seed": SEED_BANDED}
    # Wide band (kBandedWide): 65,536 pairs of 1 kbp, band [-100, 100].
    nw, lw, (wlo, wup) = 1 << 16, 1000, (-100, 100)
    hw = rng.integers(0, 4, (nw, lw), dtype=np.uint8)
    vw = hw.copy()
    sub = rng.random(hw.shape) < 0.08
    vw[sub] = rng.integers(0, 4, int(sub.sum()), dtype=np.uint8)
    offw = np.arange(nw + 1, dtype=np.int64) * lw
    Hw, Vw = sa.SeqSet(hw.reshape(-1), offw), sa.SeqSet(vw.reshape(-1), offw.copy())
    wcells = band_cells(lw, lw, wlo, wup) * nw
    wms = timed(lambda: ctx.align_score_banded(Hw, Vw, score, wlo, wup), steps, warmup, "banded wide")
    wst = ctx.stats()
    wkern = wst["device_ms"]
    wach = OPS_PER_CELL * wcells / (wkern / 1e3) / 1e12
    wpeak, wnote = band_roof(wst)
    res["wide"] = {"value": round(wcells / (wms / 1e3) / 1e9, 2), "unit": "GCUPS (cells inside the band)",
                   "ms_per_step": round(wms, 3), "device_ms": round(wkern, 3),
                   "roofline_frac": round(wach / wpeak, 4), "roofline_note": wnote,
                   "pairs_packed": int(wst["pairs_int16"]),
                   "workload": f"{nw:,} DNA pairs of {lw} bp, 8% substitutions, global, band [{wlo}, {wup}]",
                   "cells": wcells}
    # Banded rows (sqmi_align_rows_banded, int32 traceback, one thread per
    # pair): the first 65,536 pairs of the narrow workload.
    nr = 1 << 16
    Hr, Vr = head(H, nr), head(V, nr)
    rcells = band_cells(L_BANDED, L_BANDED, lo, up) * nr
    rout = [None]
    rms = timed(lambda: rout.__setitem__(0, ctx.align_rows_banded(Hr, Vr, score, lo, up)), steps, warmup,
                "banded rows")
    res["rows"] = {"value": round(rcells / (rms / 1e3) / 1e9, 2), "unit": "GCUPS (cells inside the band)",
                   "ms_per_step": round(rms, 3), "device_ms": round(ctx.stats()["device_ms"], 3),
                   "scores_equal_banded_scores": bool(np.array_equal(rout[0][0], scores[0][:nr])),
                   "workload": f"first {nr:,} pairs of the narrow workload: scores, counts and rows "
                               "(host buffers, rows packed on the device)", "cells": rcells}
    # The same call with residues and outputs resident in device memory.
    A = score.alphabet_size
    seg_bytes = int((Hr.lengths() + Vr.lengths()).sum())
    bufs = [sa.DeviceBuffer.from_array(Hr.residues), sa.DeviceBuffer.from_array(Vr.residues),
            sa.DeviceBuffer(4 * nr), sa.DeviceBuffer(4 * (A * A + 2) * nr), sa.DeviceBuffer(seg_bytes),
            sa.DeviceBuffer(8 * nr)]
    try:
        out = tuple(b.address for b in bufs[2:])
        dms = timed(lambda: ctx.align_rows_banded(Hr, Vr, score, lo, up, h_dev=bufs[0].address,
                                                  v_dev=bufs[1].address, out_dev=out), steps, warmup,
                    "banded rows resident")
        dst = ctx.stats()
        dev_scores = bufs[2].to_array(np.int32, nr)
        dev_lens = bufs[5].to_array(np.int64, nr)
        dev_ops = bufs[4].to_array(np.uint8, int(dev_lens.sum()))
        res["rows"]["resident"] = {
            "ms_per_step": round(dms, 3), "device_ms": round(dst["device_ms"], 3),
            "host_plan_ms": round(dst["host_plan_ms"], 3),
            "value": round(rcells / (dms / 1e3) / 1e9, 2),
            "equal_to_host_buffers": bool(np.array_equal(dev_scores, rout[0][0]) and
                                          np.array_equal(dev_lens, np.diff(rout[0][2])) and
                                          np.array_equal(dev_ops, rout[0][3])),
            "note": "residues, scores, counts and packed rows in device memory"}
    finally:
        for b in bufs:
            b.free()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        if po.ref_binary() is not None:
            count = 1 << 17
            hs, vs = head(H, count), head(V, count)
            r = po.run_reference(po.OP_SCORE_BANDED, hs.residues, hs.offsets, vs.residues, vs.offsets, 0,
                                 threads=host_cores(), alphabet="dna", match=2, mismatch=-3, gap_extend=-1,
                                 gap_open=-5, lower_diag=lo, upper_diag=up)
            c = band_cells(L_BANDED, L_BANDED, lo, up) * count
            res["cpu_baseline"] = {
                "value": round(c / r["best_seconds"] / 1e9, 3), "unit": "GCUPS (cells inside the band)",
                "cores": int(r["threads"]), "kind": "reference",
                "sample": f"first {count} pairs ({c:.3e} band cells); {r['binary']} banded set path, one timed "
                          f"call; {cpu_model()}",
                "parity_vs_gpu": bool(np.array_equal(r["scores"], scores[0][:count]))}
            log(f"banded cpu baseline: {res['cpu_baseline']}")
    return res


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
    ap.add_argument("--no-extra", action="store_true", help="skip the fwd+bwd and protein legs")
    ap.add_argument("--leg", choices=["config1", "fwd_bwd", "protein_local", "long_fwd_bwd", "banded"], default=None,
                    help="run only this extra leg and print its JSON (for profiling)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also copy this many MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)

    if args.leg:
        legs = {"config1": config1_leg, "fwd_bwd": backward_leg, "protein_local": protein_leg,
                "long_fwd_bwd": long_leg, "banded": banded_leg}
        print(json.dumps({args.leg: legs[args.leg](sa.Context(0), args.steps, args.warmup, not args.no_cpu)}))
        return

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

    # Every timed step plans afresh (SQMI_GRAPH=0): a replayed call would
    # skip the host planning; its rate is reported separately below.
    os.environ["SQMI_GRAPH"] = "0"
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

    os.environ.pop("SQMI_GRAPH", None)
    # The same call replayed: plan kept, launches relaunched as one graph.
    replay_ms = timed(step, args.steps, 1, "batched, replayed graph")
    replay = {"ms_per_step": round(replay_ms, 3), "gcups": round(cells / (replay_ms / 1e3) / 1e9, 2),
              "host_key_ms": round(ctx.stats()["host_plan_ms"], 3), "replayed": int(ctx.stats()["replayed"]),
              "note": "identical calls: the plan is kept and the launches relaunch the captured HIP graph"}
    # The shape table against the fixed rule it replaced (device time).
    os.environ["SQMI_SHAPE_RULE"] = "fixed"
    os.environ["SQMI_GRAPH"] = "0"
    step()
    fixed_dev = ctx.stats()["device_ms"]
    os.environ.pop("SQMI_SHAPE_RULE", None)
    step()
    tuned_dev = ctx.stats()["device_ms"]
    os.environ.pop("SQMI_GRAPH", None)
    shape_table = {"tuned_device_ms": round(tuned_dev, 3), "fixed_rule_device_ms": round(fixed_dev, 3),
                   "table": "seqan_amd/tuning/shapes.json", "provenance": shape_provenance()}
    log(f"shape table: tuned {tuned_dev:.2f} ms, fixed rule {fixed_dev:.2f} ms")

    kern_ms = float(np.mean(dev_ms))
    achieved = OPS_PER_CELL * cells / (kern_ms / 1e3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(PEAK_TOPS, 2), "unit": "TOP/s",
                "frac": round(achieved / PEAK_TOPS, 4), "traffic": profiled_traffic(),
                "note": "packed-int16 VALU element ops, 8 per algorithmic cell; kernel time from HIP events "
                        "on the library stream around the alignment kernels; traffic (HBM bytes per launch) "
                        "from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/"}

    if args.copy_calib:
        nb = args.copy_calib << 20
        src, dst = sa.DeviceBuffer(nb), sa.DeviceBuffer(nb)
        sa.hip().hipMemcpy(dst.ptr, src.ptr, nb, 3)     # hipMemcpyDeviceToDevice
        sa.synchronize()
        log(f"copied {args.copy_calib} MiB device-to-device")

    # Plain form: one kernel launch per pair, same library, same GPU.
    ps = min(args.plain_sample, args.pairs)
    plain = None
    if ps > 0:
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
        plain = {"value": round(plain_gcups, 2), "unit": "GCUPS", "sample_pairs": ps,
                 "batched_over_plain": round(gcups / plain_gcups, 1), "parity_vs_batched": plain_ok}

    cpu = None
    if not args.no_cpu:
        count = min(args.cpu_sample, args.pairs)
        cpu = reference_run(0, H, V, count, MODE, host_cores(), 2, scores, alphabet="dna", match=2, mismatch=-3,
                            gap_extend=-1, gap_open=-5)
        log(f"cpu baseline: {cpu}")
        if cpu is not None:
            one = reference_run(0, H, V, count // 16, MODE, 1, 1, scores, alphabet="dna", match=2, mismatch=-3,
                                gap_extend=-1, gap_open=-5)
            cpu["one_core"] = {"value": one["value"], "sample": one["sample"], "parity_vs_gpu": one["parity_vs_gpu"]}
            log(f"cpu baseline, one core: {one}")
    extra = {}
    if not args.no_extra:
        extra["config1"] = config1_leg(ctx, 20, 3, not args.no_cpu)
        extra["fwd_bwd"] = backward_leg(ctx, 3, 1, not args.no_cpu)
        extra["protein_local"] = protein_leg(ctx, 3, 1, not args.no_cpu)
        extra["long_fwd_bwd"] = long_leg(ctx, 3, 1, not args.no_cpu)
        extra["banded"] = banded_leg(ctx, 3, 1, not args.no_cpu)

    res = {
        "metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i16", "data": "synthetic",
        "config": {"workload": "batched uneven DNA: log-normal(300, 1.0) 100-10,000 bp, 8% subst, 2% geometric "
                               "indels, semi-global AlignConfig<1,0,0,1>, Score<int,Simple>(2,-3,-1,-5)",
                   "pairs": args.pairs, "cells": cells, "seed": SEED},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": plain,
        "host_plan_ms": round(float(np.mean(plan_ms)), 2),
        "pairs_int32": int(st["pairs_int32"]),
        "replay": replay,
        "shape_table": shape_table,
    }
    res.update(extra)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
