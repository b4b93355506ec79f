"""Shape autotuner of the packed int16 forward (run on the GPU):

    python -m seqan_amd.tune [--out seqan_amd/tuning/shapes.json] [--cells 3e10]

For every category of the planner's shape table (DNA byte profile or
substitution table, global or local), every length bin b (2^(b-1) < |V| + 1
<= 2^b) and every count class (how many of a call's pairs fall in the bin),
it times one call of the library on synthetic pairs of that bin with each
compiled shape forced (G = 16 with R = 16 / 8: a 16-lane group per pair
couple; G = 1 with R = 16 / 32: one pair couple per lane; G = 64 with
R = 4 / 8: a wavefront per pair couple), checks that all
shapes give the same scores, and keeps the fastest.  Measurements bounded
by --cells (cells per call) cover the classes they can; a class needing more
cells takes the choice of the nearest measured class below it.  The table it
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

from . import (LOCAL, FREE_TOP, FREE_BOTTOM, Context, DeviceBuffer, SeqSet, blosum62, force_shape,
               simple_score, synchronize)

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "shapes.json")
SHAPES = [(16, 16), (16, 8), (1, 16), (1, 32), (64, 4), (64, 8)]
COUNT_CLASSES = [0, 16384, 131072, 524288]
CLASS_REP = [4096, 32768, 262144, 1048576]   # pairs timed for each class
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


def tune(cells_budget=3e10, reps=2, log=print):
    ctx = Context(0)
    rng = np.random.default_rng(20261203)
    setups = {"dna": (simple_score(2, -3, -1, -5), FREE_TOP | FREE_BOTTOM, 4),
              "dna_local": (simple_score(2, -3, -1, -5), LOCAL, 4),
              "table": (blosum62(-1, -11), 0, 24),
              "table_local": (blosum62(-1, -11), LOCAL, 24)}
    choice = {c: [[0] * len(COUNT_CLASSES) for _ in range(16)] for c in CATEGORIES}
    meas = {c: {} for c in CATEGORIES}
    for cat in CATEGORIES:
        score, mode, alpha = setups[cat]
        for b in BINS[cat]:
            mean_cells = ((1 << (b - 1)) + (1 << b)) ** 2 / 4
            measured = []
            for k, rep in enumerate(CLASS_REP):
                count = rep
                if count * mean_cells > cells_budget:
                    if k > 0:
                        break
                    count = max(256, int(cells_budget / mean_cells))
                H, V = pairs_for(rng, count, b, alpha)
                dh, dv = DeviceBuffer.from_array(H.residues), DeviceBuffer.from_array(V.residues)
                ds = DeviceBuffer(4 * count)
                times, ref = {}, None
                for G, R in SHAPES:
                    ms = time_call(ctx, H, V, score, mode, force_shape(G, R), dh, dv, ds, reps)
                    got = ds.to_array(np.int32, count)
                    if ref is None:
                        ref = got
                    elif not np.array_equal(got, ref):
                        raise AssertionError(f"{cat} bin {b} count {count}: shape {G}x{R} disagrees")
                    times[f"{G}x{R}"] = round(ms, 3)
                for x in (dh, dv, ds):
                    x.free()
                best = min(range(len(SHAPES)), key=lambda i: times[f"{SHAPES[i][0]}x{SHAPES[i][1]}"])
                choice[cat][b][k] = best
                measured.append(k)
                meas[cat][f"bin{b}_count{count}"] = times
                log(f"{cat} bin {b} count {count}: {times} -> {SHAPES[best]}")
            for k in range(len(COUNT_CLASSES)):
                if k not in measured:
                    choice[cat][b][k] = choice[cat][b][max(measured)]
        for b in range(0, BINS[cat][0]):
            choice[cat][b] = list(choice[cat][BINS[cat][0]])
        for b in range(BINS[cat][-1] + 1, 16):
            choice[cat][b] = list(choice[cat][BINS[cat][-1]])
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
    ap.add_argument("--cells", type=float, default=3e10)
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
