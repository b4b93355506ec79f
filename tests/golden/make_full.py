"""Writes the full-size parity fixtures under tests/golden/ (run in a
container that holds the reference tree and the oracle/_ref build):

    make -C oracle && python tests/golden/make_full.py [c2] [c3] [c4] [c5]

Each fixture holds the outputs for EVERY pair of one benchmark workload, not
the inputs: the inputs are regenerated from (config, seed) by the library's
own generator (sqmi_workload_*, the same bytes here and on the GPU box), and
the fixture stores a SHA-256 of the generated residues and offsets so that a
generator change fails loudly instead of comparing against stale outputs.

* ``full_dna_lognormal_semiglobal.npz`` -- 1,048,576 uneven DNA pairs
  (log-normal 100..10,000 bp), AlignConfig<true,false,false,true>,
  Score<int,Simple>(2,-3,-1,-5): the scores of the reference's
  globalAlignmentScore(ExecutionPolicy<Parallel,Vectorial>, ...).
* ``full_protein_local.npz`` -- 262,144 protein pairs, Blosum62(-1,-11):
  the scores of the reference's localAlignmentScore.
* ``full_dna1k_counts.npz`` -- 65,536 DNA pairs of 1 kbp, global: the
  reference's globalAlignment scores and the counts of its rows
  (computeAlignmentStats, include/seqan/align/evaluate_alignment.h:215-300).
* ``full_dna100k_counts.npz`` -- the 64 DNA pairs of 100 kbp, global: the
  reference's scores; counts from the oracle's traceback, because the
  reference's own traceback truncates trace positions past 2^32
  (dp_traceback_impl.h:559, 607) and cannot count pairs of this size.
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402
import seqan_amd as sa  # noqa: E402

SIMPLE = dict(match=2, mismatch=-3, gap_extend=-1, gap_open=-5)
THREADS = int(os.environ.get("SQMI_FIXTURE_THREADS", "8"))


def digest(H, V):
    """SHA-256 of a generated workload (residues and offsets of both sets)."""
    h = hashlib.sha256()
    for a in (H.residues, np.ascontiguousarray(H.offsets, np.int64), V.residues,
              np.ascontiguousarray(V.offsets, np.int64)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print("wrote", name, {k: getattr(v, "shape", v) for k, v in arrays.items()}, flush=True)


def scores_in_chunks(H, V, mode, chunk, **kw):
    """Reference scores of every pair, in windows of `chunk` pairs (bounds the
    driver's input file and memory); windows are contiguous, so the
    reference's own chunking of each call sees the pairs in their order."""
    n = len(H.offsets) - 1
    out = np.empty(n, np.int32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        hs = sa.SeqSet(H.residues[H.offsets[a]:H.offsets[b]], H.offsets[a:b + 1] - H.offsets[a])
        vs = sa.SeqSet(V.residues[V.offsets[a]:V.offsets[b]], V.offsets[a:b + 1] - V.offsets[a])
        t = time.time()
        r = po.run_reference(po.OP_SCORE, hs.residues, hs.offsets, vs.residues, vs.offsets, mode,
                             threads=THREADS, timeout=24 * 3600, **kw)
        out[a:b] = r["scores"]
        print(f"  pairs {a}..{b}: {time.time() - t:.1f} s", flush=True)
    return out


def c2():
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 0, 1 << 20)
    s = scores_in_chunks(H, V, sa.FREE_TOP | sa.FREE_BOTTOM, 1 << 17, alphabet="dna", **SIMPLE)
    save("full_dna_lognormal_semiglobal.npz", scores=s, sha256=digest(H, V), config=2, seed=0x5EED0002,
         pairs=1 << 20, mode=sa.FREE_TOP | sa.FREE_BOTTOM)


def c3():
    H, V = sa.workload(sa.WL_PROTEIN, 0x5EED0003, 0, 1 << 18)
    s = scores_in_chunks(H, V, sa.LOCAL, 1 << 16, alphabet="aa", matrix="blosum62", gap_extend=-1,
                         gap_open=-11)
    save("full_protein_local.npz", scores=s, sha256=digest(H, V), config=3, seed=0x5EED0003, pairs=1 << 18,
         mode=sa.LOCAL)


def c4():
    H, V = sa.workload(sa.WL_DNA1K, 0x5EED0004, 0, 1 << 16)
    n = 1 << 16
    scores = np.empty(n, np.int32)
    counts = np.empty((n, 18), np.int32)
    chunk = 1 << 13
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        hs = sa.SeqSet(H.residues[H.offsets[a]:H.offsets[b]], H.offsets[a:b + 1] - H.offsets[a])
        vs = sa.SeqSet(V.residues[V.offsets[a]:V.offsets[b]], V.offsets[a:b + 1] - V.offsets[a])
        t = time.time()
        r = po.run_reference(po.OP_COUNTS_SIMD, hs.residues, hs.offsets, vs.residues, vs.offsets, 0,
                             threads=THREADS, alphabet="dna", timeout=24 * 3600, **SIMPLE)
        scores[a:b] = r["scores"]
        counts[a:b] = r["counts"]
        print(f"  pairs {a}..{b}: {time.time() - t:.1f} s", flush=True)
    par = np.concatenate([sa.simple_score(2, -3, -1, -5).table.ravel(), [-5, -1]]).astype(np.int64)
    assert np.array_equal(counts.astype(np.int64) @ par, scores.astype(np.int64))
    save("full_dna1k_counts.npz", scores=scores, counts=counts, sha256=digest(H, V), config=4, seed=0x5EED0004,
         pairs=n, mode=0)


def c5():
    H, V = sa.workload(sa.WL_DNA100K, 0x5EED0005, 0, 64)
    t = time.time()
    ref = scores_in_chunks(H, V, 0, 16, alphabet="dna", **SIMPLE)
    print(f"  reference scores: {time.time() - t:.1f} s", flush=True)
    sc = po.Scoring(sa.simple_score(2, -3, -1, -5).table, -5, -1)
    counts = np.empty((64, 18), np.int32)
    step = 2   # two pairs at a time: a 10 GB trace matrix each
    for a in range(0, 64, step):
        hs = sa.SeqSet(H.residues[H.offsets[a]:H.offsets[a + step]], H.offsets[a:a + step + 1] - H.offsets[a])
        vs = sa.SeqSet(V.residues[V.offsets[a]:V.offsets[a + step]], V.offsets[a:a + step + 1] - V.offsets[a])
        t = time.time()
        ws, wc = po.traceback_batch(hs.residues, hs.offsets, vs.residues, vs.offsets, sc, 0, threads=step)
        assert np.array_equal(ws, ref[a:a + step]), (a, ws, ref[a:a + step])
        counts[a:a + step] = wc
        print(f"  oracle counts of pairs {a}..{a + step}: {time.time() - t:.1f} s", flush=True)
    save("full_dna100k_counts.npz", scores=ref, counts=counts, sha256=digest(H, V), config=5, seed=0x5EED0005,
         pairs=64, mode=0)


if __name__ == "__main__":
    todo = [a for a in sys.argv[1:] if a in ("c2", "c3", "c4", "c5")] or ["c2", "c3", "c4", "c5"]
    for name in todo:
        print("fixture", name, flush=True)
        t0 = time.time()
        globals()[name]()
        print(f"fixture {name}: {time.time() - t0:.1f} s", flush=True)
