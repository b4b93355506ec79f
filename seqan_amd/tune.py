"""Shape autotuner of the packed int16 forward (run on the GPU):

    python -m seqan_amd.tune [--out seqan_amd/tuning/shapes.json] [--cells 3e10]

For every category of the planner's shape table (DNA byte profile or
substitution table, global or local), every length bin b (2^(b-1) < |V| + 1
<= 2^b) and every count class (how many pairs of a call take the packed
route, over all bins: the launches of a call's shapes run side by side, so
the whole call is what fills the GPU), it times one call of the library on
synthetic pairs of that bin with each
compiled shape forced (G = 16 with R = 16 / 8: a 16-lane group per pair
couple; G = 1 with R = 16 / 32: one pair couple per lane; G = 64 with
R = 4 / 8: a wavefront per pair couple), checks that all
shapes give the same scores, and keeps the fastest.  A class is timed on
its representative pair count, cut to --cells cells per call but never
below FILL pairs (enough to fill the GPU with every group shape), so that
the large classes of long bins measure the shapes' full-GPU rates; a class
that would repeat a smaller class's count takes that class's choice.  The table it
writes replaces seqan_amd/tuning/shapes.json; rebuild the library
(python -m seqan_amd.build) to compile it into the planner.

The knob this mirrors is the reference's block size / number of parallel
alignments of its wavefront execution policy
(include/seqan/align_parallel/dp_parallel_execution_policies.h:98, 128-170),
which a SeqAn user sets by hand; here the planner takes the measured table.
"""
from __future__ import annotations

import argparse
import datetime
import json
import os
import platform
import time

import numpy as np

from . import (LOCAL, FREE_TOP, FREE_BOTTOM, WL_DNA1K, WL_DNA_LOGNORMAL, WL_PROTEIN, Context, DeviceBuffer,
               SeqSet, blosum62, force_shape, simple_score, synchronize, workload)

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "shapes.json")
SHAPES = [(16, 16), (16, 8), (1, 16), (1, 32), (64, 4), (64, 8)]
COUNT_CLASSES = [0, 16384, 131072, 524288]
CLASS_REP = [4096, 32768, 262144, 1048576]   # pairs timed for each class
FILL = 24576   # pairs that fill the GPU with G = 16 (256 CU x 12 waves x 8 pairs)
CATEGORIES = ["dna", "dna_local", "table", "table_local"]
# Measured bins per category: up to the largest bin whose pairs all stay in
# the packed kernels' int16 window (2 x 16k DNA pairs, or 11 x 4k Blosum62
# pairs, would not); smaller bins copy bin 4, larger ones the last measured.
BINS = {"dna": list(range(4, 14)), "dna_local": list(range(4, 14)), "table": list(range(4, 12)),
        "table_local": list(range(4, 12))}


def pairs_for(rng, count, bin_, alpha):
    """count pairs with |V| + 1 uniform in bin_ and |H| within 10% of |V|;
    V is H with 10% substitutions (both residues uniform)."""
    lo, hi = (1 << (bin_ - 1)), (1 << bin_) - 1          # |V| in [lo, hi]
    m = rng.integers(lo, hi + 1, count)
    n = np.maximum(1, (m * rng.uniform(0.9, 1.1, count)).astype(np.int64))
    H = rng.integers(0, alpha, int(n.sum()), dtype=np.uint8)
    hoff = np.zeros(count + 1, np.int64)
    hoff[1:] = np.cumsum(n)
    V = np.empty(int(m.sum()), np.uint8)
    voff = np.zeros(count + 1, np.int64)
    voff[1:] = np.cumsum(m)
    # V_i: the first m_i residues of H_i (cycled when m_i > n_i), mutated.
    idx = np.arange(int(m.sum())) - np.repeat(voff[:-1], m)
    src = np.repeat(hoff[:-1], m) + idx % np.repeat(n, m)
    V[:] = H[src]
    sub = rng.random(V.size) < 0.1
    V[sub] = rng.integers(0, alpha, int(sub.sum()), dtype=np.uint8)
    return SeqSet(H, hoff), SeqSet(V, voff)


def time_call(ctx, H, V, score, mode, flags, dh, dv, ds, reps):
    ctx.align_score(H, V, score, mode, flags=flags, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
    best = float("inf")
    for _ in range(reps):
        synchronize()
        t = time.perf_counter()
        ctx.align_score(H, V, score, mode, flags=flags, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
        best = min(best, time.perf_counter() - t)
    return best * 1e3


def smooth(choice, byClass, bins, log, margin=0.9):
    """A call's bins run as one launch per distinct shape, so a mixture of
    shapes costs more than the bins' separate timings add up to (each launch
    has its own tail, and launches of different register budgets share the
    CUs badly).  Per count class, the single shape with the smallest summed
    slowdown over the long bins (b >= 8) is preferred in every long bin where
    the bin's own best beats it by less than 10% (margin)."""
    long_bins = [b for b in bins if b >= 8]
    for k in range(len(COUNT_CLASSES)):
        timed = [(b, byClass[(b, k)]) for b in long_bins if (b, k) in byClass]
        if not timed:
            continue
        names = set.intersection(*(set(t) for _, t in timed))
        if not names:
            continue
        dom = min(names, key=lambda nm: sum(t[nm] / min(t.values()) for _, t in timed))
        di = next(i for i, sh in enumerate(SHAPES) if f"{sh[0]}x{sh[1]}" == dom)
        for b, t in timed:
            if min(t.values()) >= margin * t[dom] and choice[b][k] != di:
                log(f"  class {k} bin {b}: {SHAPES[choice[b][k]]} -> {SHAPES[di]} (one shape for the long bins)")
                choice[b][k] = di


# Whole-call refinement: per (category, count class), a representative
# workload of that class (the benchmark workloads), timed as one call while
# one bin at a time takes each other shape (SQMI_BIN_SHAPES).
REFINE = [("dna", "1,048,576 uneven DNA pairs, semi-global", WL_DNA_LOGNORMAL, 0x5EED0002, 1 << 20,
           FREE_TOP | FREE_BOTTOM),
          ("dna", "65,536 DNA pairs of 1 kbp, global", WL_DNA1K, 0x5EED0004, 1 << 16, 0),
          ("dna_local", "1,048,576 uneven DNA pairs, local", WL_DNA_LOGNORMAL, 0x5EED0002, 1 << 20, LOCAL),
          ("table_local", "262,144 protein pairs, local", WL_PROTEIN, 0x5EED0003, 1 << 18, LOCAL),
          ("table", "262,144 protein pairs, global", WL_PROTEIN, 0x5EED0003, 1 << 18, 0)]


def count_class(pairs):
    k = 0
    while k + 1 < len(COUNT_CLASSES) and pairs >= COUNT_CLASSES[k + 1]:
        k += 1
    return k


def refine(ctx, setups, choice, allClass, log, reps=2):
    """Per representative workload: first every group shape for all the bins
    holding >= 1% of its cells at once, then coordinate descent over those
    bins, heaviest first, each trying the shapes its own timing put within
    2x of its best; a change is kept when the whole call gets >= 2% faster."""
    notes = {}
    for cat, what, wl, seed, count, mode in REFINE:
        score = setups[cat][0]
        byClass = allClass[cat]
        H, V = workload(wl, seed, 0, count)
        k = count_class(count)
        m = np.diff(V.offsets)
        cells = m * np.diff(H.offsets)
        bins = np.ceil(np.log2(m + 1)).astype(int)
        share = {int(b): float(cells[bins == b].sum() / cells.sum()) for b in np.unique(bins)}
        pshare = {int(b): float((bins == b).mean()) for b in np.unique(bins)}
        order = sorted((b for b, f in share.items() if f >= 0.01), key=lambda b: -share[b])
        # Light bins (few cells, but their pairs may be wide: a shape with a
        # long per-pair latency there can set the whole call's time).
        light = sorted((b for b in share if share[b] < 0.01 and pshare[b] >= 0.001), key=lambda b: -share[b])
        cur = {b: SHAPES[choice[cat][b][k]] for b in share}
        dh, dv = DeviceBuffer.from_array(H.residues), DeviceBuffer.from_array(V.residues)
        ds = DeviceBuffer(4 * count)

        def timed(shapes):
            os.environ["SQMI_BIN_SHAPES"] = ",".join(f"{b}:{g}:{r}" for b, (g, r) in sorted(shapes.items()))
            try:
                ctx.align_score(H, V, score, mode, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
                best = float("inf")
                for _ in range(reps):
                    ctx.align_score(H, V, score, mode, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
                    best = min(best, ctx.stats()["device_ms"])
                return best
            finally:
                os.environ.pop("SQMI_BIN_SHAPES", None)

        t0 = timed(cur)
        best_t = t0
        # One shape for all the heavy bins first (mixtures of shapes pay for
        # sharing the GPU between launches, so single-bin moves alone can
        # stall short of a uniform choice), then one bin at a time.
        for sh in SHAPES:
            if sh[0] == 1 or all(cur[b] == sh for b in order):
                continue
            t = timed({**cur, **{b: sh for b in order}})
            if t < 0.98 * best_t:
                log(f"  refine {cat} class {k} ({what}): heavy bins -> {sh}: {best_t:.2f} -> {t:.2f} ms")
                cur.update({b: sh for b in order})
                best_t = t
        if light and order:
            # The light bins on the heavy bins' leading shape.
            sh = cur[order[0]]
            if any(cur[b] != sh for b in light):
                t = timed({**cur, **{b: sh for b in light}})
                if t < 0.98 * best_t:
                    log(f"  refine {cat} class {k} ({what}): light bins -> {sh}: {best_t:.2f} -> {t:.2f} ms")
                    cur.update({b: sh for b in light})
                    best_t = t
        for b in order + light:
            own = byClass.get((min(max(b, 4), 15), k)) or {}
            fast = min(own.values()) if own else None
            for sh in SHAPES:
                if sh == cur[b] or (own and (f"{sh[0]}x{sh[1]}" not in own or own[f"{sh[0]}x{sh[1]}"] > 2 * fast)):
                    continue
                t = timed({**cur, b: sh})
                if t < 0.98 * best_t:
                    log(f"  refine {cat} class {k} ({what}): bin {b} {cur[b]} -> {sh}: {best_t:.2f} -> {t:.2f} ms")
                    cur[b], best_t = sh, t
        for b, sh in cur.items():
            if 0 <= b < 16:
                choice[cat][b][k] = SHAPES.index(sh)
        for x in (dh, dv, ds):
            x.free()
        notes[f"{cat}_class{k}"] = {"workload": what, "device_ms_table": round(t0, 3), "device_ms_refined": round(best_t, 3)}
        log(f"refine {cat} class {k} ({what}): {t0:.2f} -> {best_t:.2f} ms")
    return notes


def tune(cells_budget=2e11, reps=2, log=print):
    ctx = Context(0)
    rng = np.random.default_rng(20261203)
    setups = {"dna": (simple_score(2, -3, -1, -5), FREE_TOP | FREE_BOTTOM, 4),
              "dna_local": (simple_score(2, -3, -1, -5), LOCAL, 4),
              "table": (blosum62(-1, -11), 0, 24),
              "table_local": (blosum62(-1, -11), LOCAL, 24)}
    choice = {c: [[0] * len(COUNT_CLASSES) for _ in range(16)] for c in CATEGORIES}
    meas = {c: {} for c in CATEGORIES}
    allClass = {}
    for cat in CATEGORIES:
        score, mode, alpha = setups[cat]
        byClass = {}
        for b in BINS[cat]:
            mean_cells = ((1 << (b - 1)) + (1 << b)) ** 2 / 4
            measured = []
            last = 0
            contenders = list(SHAPES)
            for k, rep in enumerate(CLASS_REP):
                count = rep
                if count * mean_cells > cells_budget:
                    floor = 256 if k == 0 else min(rep, FILL)
                    count = max(floor, int(cells_budget / mean_cells))
                if count <= last:
                    break
                last = count
                H, V = pairs_for(rng, count, b, alpha)
                dh, dv = DeviceBuffer.from_array(H.residues), DeviceBuffer.from_array(V.residues)
                ds = DeviceBuffer(4 * count)
                times, ref = {}, None
                for G, R in SHAPES:
                    if (G, R) not in contenders:
                        continue
                    ms = time_call(ctx, H, V, score, mode, force_shape(G, R), dh, dv, ds, reps)
                    got = ds.to_array(np.int32, count)
                    if ref is None:
                        ref = got
                    elif not np.array_equal(got, ref):
                        raise AssertionError(f"{cat} bin {b} count {count}: shape {G}x{R} disagrees")
                    times[f"{G}x{R}"] = round(ms, 3)
                for x in (dh, dv, ds):
                    x.free()
                best = min((i for i in range(len(SHAPES)) if SHAPES[i] in contenders),
                           key=lambda i: times[f"{SHAPES[i][0]}x{SHAPES[i][1]}"])
                # At a count that fills the GPU, shapes 2.5x slower than the
                # best drop out of this bin's larger classes (they only fall
                # further behind; below FILL the group shapes may just lack
                # pairs to fill the GPU).
                fastest = times[f"{SHAPES[best][0]}x{SHAPES[best][1]}"]
                if count >= FILL:
                    contenders = [sh for sh in contenders if times[f"{sh[0]}x{sh[1]}"] <= 2.5 * fastest]
                choice[cat][b][k] = best
                measured.append(k)
                meas[cat][f"bin{b}_count{count}"] = times
                byClass[(b, k)] = times
                log(f"{cat} bin {b} count {count}: {times} -> {SHAPES[best]}")
            for k in range(len(COUNT_CLASSES)):
                if k not in measured:
                    choice[cat][b][k] = choice[cat][b][max(measured)]
                    byClass[(b, k)] = byClass[(b, max(measured))]
        smooth(choice[cat], byClass, BINS[cat], log)
        for b in range(0, BINS[cat][0]):
            choice[cat][b] = list(choice[cat][BINS[cat][0]])
        for b in range(BINS[cat][-1] + 1, 16):
            choice[cat][b] = list(choice[cat][BINS[cat][-1]])
        allClass[cat] = byClass
    meas["refine"] = refine(ctx, setups, choice, allClass, log)
    return choice, meas


def gpu_name():
    """The first GPU agent's marketing name and gfx target, from rocminfo."""
    import shutil
    import subprocess
    exe = shutil.which("rocminfo") or "/opt/rocm/bin/rocminfo"
    try:
        out = subprocess.run([exe], capture_output=True, text=True, timeout=60).stdout
    except Exception:
        return "unknown"
    name, arch = None, None
    for line in out.splitlines():
        key, _, val = line.strip().partition(":")
        val = val.strip()
        if key == "Name" and val.startswith("gfx") and arch is None:
            arch = val
        if key == "Marketing Name" and arch is not None and name is None:
            name = val
    if not name:
        try:   # rocminfo leaves the marketing name blank on some hosts
            import torch
            name = torch.cuda.get_device_name(0)
        except Exception:
            name = None
    return f"{name or 'unknown'} ({arch or 'unknown'})"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--cells", type=float, default=2e11)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()
    t0 = time.time()
    choice, meas = tune(args.cells, args.reps, log=lambda *a: print(*a, flush=True))
    with open(TABLE) as f:
        about = json.load(f)["about"]
    table = {
        "about": about,
        "provenance": {"kind": "measured", "measured": True, "tool": "python -m seqan_amd.tune",
                       "gpu": gpu_name(), "host": platform.node(),
                       "date": datetime.datetime.utcnow().strftime("%Y-%m-%d %H:%M UTC"),
                       "cells_per_call": args.cells, "reps": args.reps, "seconds": round(time.time() - t0, 1)},
        "shapes": [list(s) for s in SHAPES],
        "count_classes": COUNT_CLASSES,
        "categories": choice,
        "coop_min_blocks": 10,
        "measurements_ms": meas,
    }
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
