"""Writes the golden fixtures under tests/golden/ (run in a container that
holds the reference tree and the oracle/_ref build):

    make -C oracle && python tests/golden/make_golden.py
    python tests/golden/make_golden.py --long   # the 100 kbp fixture (~25 GB of RAM)
    python tests/golden/make_golden.py --modes  # only the AlignConfig rows fixtures
    python tests/golden/make_golden.py --local  # only the local rows fixtures
    python tests/golden/make_golden.py --linear # only the linear-gap rows fixtures
    python tests/golden/make_golden.py --banded # only the banded score fixtures

* ``ref_*.npz``   -- inputs (residue ordinals + offsets) and the outputs of
  the reference library itself (oracle/_ref: SeqAn 2 globalAlignmentScore /
  localAlignmentScore / globalAlignment under ExecutionPolicy<Parallel,
  Vectorial>, plus computeAlignmentStats on the rows, and the rows as one
  byte per column: M, V (gap in the H row), H (gap in the V row)).
* ``pair_align.json`` -- sequences and expected scores / rows copied from the
  reference's app tests (apps/pair_align/tests/*.fa, *.stdout, *_out.fa).
* ``BLOSUM62.txt`` -- the reference's tests/score/BLOSUM62 data file.
"""
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402
import seqan_amd as sa  # noqa: E402

REF = "/root/reference"
PA = os.path.join(REF, "apps", "pair_align", "tests")

SIMPLE = dict(match=2, mismatch=-3, gap_extend=-1, gap_open=-5)


def save(name, H, V, **extra):
    np.savez_compressed(os.path.join(HERE, name), h_res=H.residues, h_off=H.offsets, v_res=V.residues,
                        v_off=V.offsets, **extra)
    print("wrote", name, {k: getattr(v, "shape", v) for k, v in extra.items()})


def ref_scores(H, V, mode, **kw):
    r = po.run_reference(po.OP_SCORE, H.residues, H.offsets, V.residues, V.offsets, mode, **kw)
    return r["scores"]


def random_pairs(rng, count, lo, hi, alpha, related=True):
    hs, vs = [], []
    for _ in range(count):
        n = int(rng.integers(lo, hi + 1))
        h = rng.integers(0, alpha, n).astype(np.uint8)
        if related and rng.random() < 0.7:
            v = h.copy()
            mut = rng.random(n) < 0.2
            v[mut] = rng.integers(0, alpha, int(mut.sum()))
            if n > 4 and rng.random() < 0.5:
                a = int(rng.integers(0, n - 2))
                b = int(rng.integers(a + 1, min(n, a + 6)))
                v = np.concatenate([v[:a], v[b:]]) if rng.random() < 0.5 else np.concatenate(
                    [v[:a], rng.integers(0, alpha, b - a).astype(np.uint8), v[a:]])
        else:
            v = rng.integers(0, alpha, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        if len(v) == 0:
            v = np.array([0], np.uint8)
        hs.append(h)
        vs.append(v)
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


def main():
    os.makedirs(HERE, exist_ok=True)
    rng = np.random.default_rng(20261202)

    # --- reference data files -------------------------------------------
    shutil.copyfile(os.path.join(REF, "tests", "score", "BLOSUM62"), os.path.join(HERE, "BLOSUM62.txt"))
    pa = {"source": "apps/pair_align/tests (reference app tests)", "pairs": {}}
    for name in ["1aab", "1ad2", "2trx", "adeno1", "adeno2", "adeno3"]:
        txt = open(os.path.join(PA, name + ".fa")).read()
        recs = [r for r in re.split(r"^>.*$", txt, flags=re.M) if r.strip()]
        seqs = ["".join(r.split()) for r in recs]
        entry = {"h": seqs[0], "v": seqs[1], "scores": {}, "rows": {}}
        for f in sorted(os.listdir(PA)):
            m = re.match(re.escape(name) + r"(\.[A-Za-z0-9-]+)?\.stdout$", f)
            if not m:
                continue
            tag = (m.group(1) or ".default")[1:]
            sm = re.search(r"score:\s*(-?\d+)", open(os.path.join(PA, f)).read())
            if sm:
                entry["scores"][tag] = int(sm.group(1))
            rows_file = os.path.join(PA, f.replace(".stdout", "_out.fa"))
            if os.path.exists(rows_file):
                rt = open(rows_file).read()
                rr = [r for r in re.split(r"^>.*$", rt, flags=re.M) if r.strip()]
                entry["rows"][tag] = ["".join(r.split()) for r in rr]
        pa["pairs"][name] = entry
    with open(os.path.join(HERE, "pair_align.json"), "w") as f:
        json.dump(pa, f, indent=1, sort_keys=True)
    print("wrote pair_align.json")

    # --- 4,096 DNA pairs of 150 bp, global, every score -----------------
    H, V = sa.workload(sa.WL_DNA150, 0x5EED0001, 0, 4096)
    save("ref_dna150_global.npz", H, V, scores=ref_scores(H, V, 0, alphabet="dna", **SIMPLE),
         seed=0x5EED0001, mode=0)

    # --- uneven log-normal DNA, first 1,500 pairs, AlignConfig<1,0,0,1> --
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 0, 1500)
    save("ref_dna_lognormal_semiglobal.npz", H, V,
         scores=ref_scores(H, V, 9, alphabet="dna", **SIMPLE), seed=0x5EED0002, mode=9)

    # --- protein local, first 300 pairs, Blosum62(-1, -11) --------------
    H, V = sa.workload(sa.WL_PROTEIN, 0x5EED0003, 0, 300)
    save("ref_protein_local.npz", H, V,
         scores=ref_scores(H, V, 16, alphabet="aa", matrix="blosum62", gap_extend=-1, gap_open=-11),
         seed=0x5EED0003, mode=16)

    # --- every AlignConfig and local on short random DNA and protein -----
    H, V = random_pairs(rng, 400, 1, 90, 4)
    modes = list(range(16)) + [16]
    allm = np.stack([ref_scores(H, V, m, alphabet="dna", match=2, mismatch=-1, gap_extend=-1, gap_open=-3)
                     for m in modes])
    save("ref_dna_all_modes.npz", H, V, scores=allm, modes=np.array(modes), match=2, mismatch=-1,
         gap_extend=-1, gap_open=-3)
    H, V = random_pairs(rng, 200, 1, 120, 24)
    allm = np.stack([ref_scores(H, V, m, alphabet="aa", matrix="blosum62", gap_extend=-2, gap_open=-4)
                     for m in modes])
    save("ref_protein_all_modes.npz", H, V, scores=allm, modes=np.array(modes), gap_extend=-2, gap_open=-4)

    # --- 1 kbp DNA, first 96 pairs: counts (SIMD and scalar reference) --
    H, V = sa.workload(sa.WL_DNA1K, 0x5EED0004, 0, 96)
    rs = po.run_reference(po.OP_COUNTS_SIMD, H.residues, H.offsets, V.residues, V.offsets, 0, alphabet="dna",
                          **SIMPLE)
    rc = po.run_reference(po.OP_COUNTS_SCALAR, H.residues, H.offsets, V.residues, V.offsets, 0, alphabet="dna",
                          **SIMPLE)
    assert np.array_equal(rs["counts"], rc["counts"]) and np.array_equal(rs["scores"], rc["scores"])
    assert np.array_equal(rs["ops"], rc["ops"]) and np.array_equal(rs["op_off"], rc["op_off"])
    save("ref_dna1k_counts.npz", H, V, scores=rs["scores"], counts=rs["counts"], op_off=rs["op_off"], ops=rs["ops"],
         seed=0x5EED0004, mode=0)

    # --- short random DNA traceback counts with many ties ----------------
    H, V = random_pairs(rng, 300, 1, 60, 4)
    rs = po.run_reference(po.OP_COUNTS_SIMD, H.residues, H.offsets, V.residues, V.offsets, 0, alphabet="dna",
                          match=1, mismatch=-1, gap_extend=-1, gap_open=-2)
    save("ref_dna_short_counts.npz", H, V, scores=rs["scores"], counts=rs["counts"], op_off=rs["op_off"],
         ops=rs["ops"], match=1, mismatch=-1, gap_extend=-1, gap_open=-2)

    # --- protein traceback counts and rows, Blosum62(-1, -11) --------------
    H, V = random_pairs(rng, 200, 20, 400, 24)
    rs = po.run_reference(po.OP_COUNTS_SIMD, H.residues, H.offsets, V.residues, V.offsets, 0, alphabet="aa",
                          matrix="blosum62", gap_extend=-1, gap_open=-11)
    save("ref_protein_counts.npz", H, V, scores=rs["scores"], counts=rs["counts"], op_off=rs["op_off"],
         ops=rs["ops"], gap_extend=-1, gap_open=-11)


def trimmed_pairs(rng, count, lo, hi):
    """Related DNA pairs whose V often lacks a prefix or a suffix of H, so
    that free end gaps matter."""
    hs, vs = [], []
    for _ in range(count):
        h = rng.integers(0, 4, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        v = h.copy()
        f = rng.random(len(v)) < 0.2
        v[f] = rng.integers(0, 4, int(f.sum()))
        if rng.random() < 0.5:
            v = v[int(rng.integers(0, len(v) // 3 + 1)):]
        if rng.random() < 0.5:
            v = v[:max(1, len(v) - int(rng.integers(0, len(v) // 3 + 1)))]
        if rng.random() < 0.3:
            h, v = v, h
        hs.append(h)
        vs.append(v)
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


def mode_rows():
    """Counts and rows of global alignment under each of the 16 AlignConfigs
    (the reference's globalAlignment(ExecutionPolicy<Parallel, Vectorial>,
    gapsH, gapsV, score, AlignConfig<T, L, R, B>))."""
    rng = np.random.default_rng(0x5EED0100)
    for name, count, lo, hi, par in (("ref_dna_modes_rows.npz", 128, 5, 300, (2, -3, -1, -5)),
                                    ("ref_dna_modes_short_rows.npz", 200, 1, 40, (1, -1, -1, -2))):
        H, V = trimmed_pairs(rng, count, lo, hi)
        kw = dict(match=par[0], mismatch=par[1], gap_extend=par[2], gap_open=par[3])
        res = [po.run_reference(po.OP_COUNTS_SIMD, H.residues, H.offsets, V.residues, V.offsets, mode,
                                alphabet="dna", **kw) for mode in range(16)]
        mode_off = np.cumsum([0] + [len(r["ops"]) for r in res]).astype(np.int64)
        save(name, H, V, scores=np.stack([r["scores"] for r in res]), counts=np.stack([r["counts"] for r in res]),
             op_off=np.stack([r["op_off"] for r in res]), ops=np.concatenate([r["ops"] for r in res]),
             mode_off=mode_off, **kw)


def local_pairs(rng, count, lo, hi, alpha):
    """Pairs whose V embeds a mutated segment of H between random flanks
    (a third of them unrelated), so that local windows start and end inside
    both sequences."""
    hs, vs = [], []
    for _ in range(count):
        h = rng.integers(0, alpha, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        if rng.random() < 0.33:
            v = rng.integers(0, alpha, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        else:
            a = int(rng.integers(0, len(h)))
            b = int(rng.integers(a, len(h))) + 1
            seg = h[a:b].copy()
            f = rng.random(len(seg)) < 0.15
            seg[f] = rng.integers(0, alpha, int(f.sum()))
            cut = np.nonzero(rng.random(len(seg)) < 0.03)[0]
            seg = np.delete(seg, cut)
            left = rng.integers(0, alpha, int(rng.integers(0, hi // 4 + 1))).astype(np.uint8)
            right = rng.integers(0, alpha, int(rng.integers(0, hi // 4 + 1))).astype(np.uint8)
            v = np.concatenate([left, seg, right]).astype(np.uint8)
            if len(v) == 0:
                v = np.array([0], np.uint8)
        if rng.random() < 0.3:
            h, v = v, h
        hs.append(h)
        vs.append(v)
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


def local_rows():
    """Counts, rows and window begins of local alignment (the reference's
    localAlignment(ExecutionPolicy<Parallel, Vectorial>, gapsH, gapsV,
    score)); the scalar localAlignment must agree."""
    rng = np.random.default_rng(0x5EED0200)
    for name, count, lo, hi, alpha, kw in (
            ("ref_dna_local_rows.npz", 160, 1, 400, 4, dict(alphabet="dna", **SIMPLE)),
            ("ref_dna_local_short_rows.npz", 256, 1, 30, 4,
             dict(alphabet="dna", match=1, mismatch=-1, gap_extend=-1, gap_open=-2)),
            ("ref_protein_local_rows.npz", 128, 10, 400, 27,
             dict(alphabet="aa", matrix="blosum62", gap_extend=-1, gap_open=-11))):
        H, V = local_pairs(rng, count, lo, hi, 24 if alpha == 27 else alpha)
        r = po.run_reference(po.OP_COUNTS_SIMD, H.residues, H.offsets, V.residues, V.offsets, 16, **kw)
        q = po.run_reference(po.OP_COUNTS_SCALAR, H.residues, H.offsets, V.residues, V.offsets, 16, **kw)
        for key in ("scores", "counts", "op_off", "ops", "begins"):
            assert np.array_equal(r[key], q[key]), (name, key)
        extra = {k: v for k, v in kw.items() if k not in ("alphabet", "matrix")}
        save(name, H, V, scores=r["scores"], counts=r["counts"], op_off=r["op_off"], ops=r["ops"],
             begins=r["begins"], **extra)


def linear_rows():
    """Rows under linear gaps (gap open == gap extend), where the reference
    switches to its linear-gap algorithm (_usesAffineGaps, dp_setup.h:244-260)
    with its own tie-breaks: global under each AlignConfig (modes 0-15) and
    local (mode 16)."""
    rng = np.random.default_rng(0x5EED0300)
    for name, count, lo, hi, alpha, kw, modes in (
            ("ref_dna_linear_rows.npz", 160, 1, 200, 4,
             dict(alphabet="dna", match=1, mismatch=-1, gap_extend=-1, gap_open=-1), range(17)),
            ("ref_dna_linear2_rows.npz", 96, 20, 400, 4,
             dict(alphabet="dna", match=2, mismatch=-3, gap_extend=-2, gap_open=-2), (0, 9, 15, 16)),
            ("ref_protein_linear_rows.npz", 96, 10, 300, 27,
             dict(alphabet="aa", matrix="blosum62", gap_extend=-4, gap_open=-4), (0, 9, 16))):
        H, V = trimmed_pairs(rng, count, lo, hi) if alpha == 4 else local_pairs(rng, count, lo, hi, 24)
        res = [po.run_reference(po.OP_COUNTS_SIMD, H.residues, H.offsets, V.residues, V.offsets, mode, **kw)
               for mode in modes]
        extra = {k: v for k, v in kw.items() if k not in ("alphabet", "matrix")}
        mode_off = np.cumsum([0] + [len(r["ops"]) for r in res]).astype(np.int64)
        save(name, H, V, modes=np.array(list(modes), np.int32), scores=np.stack([r["scores"] for r in res]),
             counts=np.stack([r["counts"] for r in res]), op_off=np.stack([r["op_off"] for r in res]),
             ops=np.concatenate([r["ops"] for r in res]), mode_off=mode_off,
             begins=res[-1]["begins"] if modes[-1] == 16 else np.zeros((0, 2), np.int64), **extra)


def banded_scores():
    """Banded globalAlignmentScore(seqH, seqV, score, AlignConfig, lowerDiag,
    upperDiag) for one pair (global_alignment_banded.h:394-470, driven with
    repeats = -1), and localAlignmentScore(..., lowerDiag, upperDiag) for
    mode 16 (local_alignment_banded.h:317-342), under several (mode, band)
    cases: `scores`.  Beside them,
    `simd_scores` from the set path the ExecutionPolicy wrapper reaches
    (:555-583) on the same collections, whose H and V each have one length as
    its documentation requires (on these cases the two agree).  All bands
    hold both corner diagonals with a margin and every H has two residues or
    more: the domain where the reference's band code computes the banded
    recurrence."""
    rng = np.random.default_rng(0x5EED0400)

    def equal_length(count, n, m, alpha):
        hs, vs = [], []
        for _ in range(count):
            h = rng.integers(0, alpha, n).astype(np.uint8)
            v = h[:m].copy() if m <= n else np.concatenate([h, rng.integers(0, alpha, m - n).astype(np.uint8)])
            f = rng.random(m) < 0.12
            v[f] = rng.integers(0, alpha, int(f.sum()))
            hs.append(h)
            vs.append(v.astype(np.uint8))
        return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)

    def run(H, V, cases, kw, scalar=True):
        out = []
        for mode, lo, up in cases:
            r = po.run_reference(po.OP_SCORE_BANDED, H.residues, H.offsets, V.residues, V.offsets, mode,
                                 lower_diag=lo, upper_diag=up, repeats=-1 if scalar else 1, **kw)
            out.append(r["scores"])
        return np.stack(out)

    dna = dict(alphabet="dna", **SIMPLE)
    for name, count, n, m, alpha, kw, cases in (
            ("ref_dna_banded_150.npz", 256, 150, 150, 4, dna,
             [(0, -5, 5), (0, -1, 1), (0, -20, 7), (9, -10, 10), (15, -3, 12), (6, -31, 32), (16, -5, 5),
              (16, -20, 7)]),
            ("ref_dna_banded_300x280.npz", 96, 300, 280, 4, dna,
             [(0, -60, 80), (0, -10, 100), (9, -150, 150), (15, -5, 25), (0, -500, 500), (16, -60, 80)]),
            ("ref_dna_banded_linear.npz", 128, 120, 126, 4,
             dict(alphabet="dna", match=1, mismatch=-1, gap_extend=-1, gap_open=-1),
             [(0, -8, 8), (9, -40, 30), (15, -7, 2)]),
            ("ref_protein_banded.npz", 96, 120, 100, 27,
             dict(alphabet="aa", matrix="blosum62", gap_extend=-1, gap_open=-11),
             [(0, -30, 40), (0, -1, 21), (9, -5, 70), (16, -30, 40), (16, 3, 90)])):
        H, V = equal_length(count, n, m, 24 if alpha == 27 else alpha)
        extra = {k: v for k, v in kw.items() if k not in ("alphabet", "matrix")}
        save(name, H, V, cases=np.array(cases, np.int64), scores=run(H, V, cases, kw),
             simd_scores=run(H, V, cases, kw, scalar=False), **extra)
    # Varying lengths: the per-pair function.
    hs, vs = [], []
    for _ in range(200):
        n = int(rng.integers(2, 260))
        h = rng.integers(0, 4, n).astype(np.uint8)
        v = np.delete(h, np.nonzero(rng.random(n) < 0.05)[0])
        if len(v) == 0:
            v = h.copy()
        f = rng.random(len(v)) < 0.1
        v[f] = rng.integers(0, 4, int(f.sum()))
        hs.append(h)
        vs.append(v)
    H, V = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    lim = int(max(abs(len(a) - len(b)) for a, b in zip(hs, vs)))
    cases = [(0, -lim - 3, lim + 3), (9, -lim - 40, lim + 5), (15, -lim - 1, lim + 70), (16, -lim - 3, lim + 3)]
    save("ref_dna_banded_varying.npz", H, V, cases=np.array(cases, np.int64), scores=run(H, V, cases, dna),
         **SIMPLE)


def banded_rows():
    """Banded rows, pair by pair: globalAlignment(gapsH, gapsV, score,
    AlignConfig, lowerDiag, upperDiag) (global_alignment_banded.h:152-181)
    and, for mode 16, localAlignment(gapsH, gapsV, score, lowerDiag,
    upperDiag) (local_alignment_banded.h:242-259), under several
    (mode, lower, upper) cases (`cases`): per case the scores, counts and
    rows (ops of case k in ops[case_off[k]:case_off[k + 1]]), and the
    window begins of local cases (`begins`, one [count, 2] block per case,
    zeros for global ones).  Varying lengths; every band is in the domain
    the library takes (the one of the banded score fixtures)."""
    rng = np.random.default_rng(0x5EED0500)

    def pairs(count, lo_len, hi_len, alpha, dele=0.05, sub=0.12):
        hs, vs = [], []
        for _ in range(count):
            n = int(rng.integers(lo_len, hi_len + 1))
            h = rng.integers(0, alpha, n).astype(np.uint8)
            v = np.delete(h, np.nonzero(rng.random(n) < dele)[0])
            if len(v) == 0:
                v = h.copy()
            f = rng.random(len(v)) < sub
            v[f] = rng.integers(0, alpha, int(f.sum()))
            hs.append(h)
            vs.append(v.astype(np.uint8))
        return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)

    def limit(H, V):
        return int(np.max(np.abs(H.lengths() - V.lengths())))

    for name, count, lo_len, hi_len, alpha, kw, spans in (
            ("ref_dna_banded_rows.npz", 160, 2, 220, 4, dict(alphabet="dna", **SIMPLE),
             [(0, 3, 3), (8, 2, 10), (9, 40, 5), (15, 1, 70), (4, 6, 1), (6, 2, 2), (16, 3, 3), (16, 12, 30)]),
            ("ref_dna_banded_ties_rows.npz", 300, 2, 30, 4,
             dict(alphabet="dna", match=1, mismatch=-1, gap_extend=-1, gap_open=-2),
             [(0, 1, 1), (8, 1, 1), (8, 2, 4), (9, 1, 2), (15, 1, 1), (2, 3, 1), (16, 1, 1), (16, 4, 6)]),
            ("ref_dna_banded_linear_rows.npz", 160, 2, 60, 4,
             dict(alphabet="dna", match=1, mismatch=-1, gap_extend=-1, gap_open=-1),
             [(0, 2, 2), (8, 1, 3), (15, 1, 1), (16, 2, 2)]),
            ("ref_protein_banded_rows.npz", 96, 2, 150, 24,
             dict(alphabet="aa", matrix="blosum62", gap_extend=-1, gap_open=-11),
             [(0, 4, 4), (9, 20, 10), (16, 5, 5)])):
        H, V = pairs(count, lo_len, hi_len, alpha)
        lim = limit(H, V)
        # (mode, a, b) -> band [-lim - a, lim + b]: both corner diagonals inside.
        cases = [(mode, -lim - a, lim + b) for mode, a, b in spans]
        res = [po.run_reference(po.OP_ROWS_BANDED, H.residues, H.offsets, V.residues, V.offsets, mode,
                                lower_diag=lo, upper_diag=up, **kw) for mode, lo, up in cases]
        extra = {k: v for k, v in kw.items() if k not in ("alphabet", "matrix")}
        case_off = np.cumsum([0] + [len(r["ops"]) for r in res]).astype(np.int64)
        begins = np.stack([r["begins"] if mode == 16 else np.zeros((count, 2), np.int64)
                           for r, (mode, _, _) in zip(res, cases)])
        save(name, H, V, cases=np.array(cases, np.int64), scores=np.stack([r["scores"] for r in res]),
             counts=np.stack([r["counts"] for r in res]), op_off=np.stack([r["op_off"] for r in res]),
             ops=np.concatenate([r["ops"] for r in res]), case_off=case_off, begins=begins, **extra)


def banded_end_rows():
    """Banded rows whose end cell lies on the band's lower diagonal in the
    last row (free last row), where the reference's traceback takes the
    cell's own direction rather than preferring the vertical gap: short DNA
    pairs under three scorings, the band [-3, 6] (|H| - |V| in [-2, 5]), the
    reference's rows pair by pair, kept when its end cell (m, n - trailing
    gap columns) has j - i == -3.  Per scoring s: cases[s] = (mode, lower,
    upper, match, mismatch, gap_extend, gap_open), pairs sel[s]."""
    rng = np.random.default_rng(0x5EED0600)
    lo, up = -3, 6
    hs_all, vs_all, rows = [], [], []
    for s, (ma, mi, ge, go) in enumerate(((1, -1, -1, -2), (2, -3, -1, -5), (3, -2, -3, -2))):
        for mode in (8, 9):
            hs, vs = [], []
            for _ in range(3000):
                m = int(rng.integers(4, 24))
                n = m + int(rng.integers(-2, 6))
                n = max(n, 2)
                hs.append(rng.integers(0, 4, n).astype(np.uint8))
                vs.append(rng.integers(0, 4, m).astype(np.uint8))
            H, V = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
            r = po.run_reference(po.OP_ROWS_BANDED, H.residues, H.offsets, V.residues, V.offsets, mode,
                                 lower_diag=lo, upper_diag=up, alphabet="dna", match=ma, mismatch=mi,
                                 gap_extend=ge, gap_open=go)
            kept = 0
            for k in range(len(hs)):
                ops = bytes(r["ops"][r["op_off"][k]:r["op_off"][k + 1]])
                tail = len(ops) - len(ops.rstrip(b"H"))
                if ops.endswith(b"V") or len(hs[k]) - tail - len(vs[k]) != lo:
                    continue
                hs_all.append(hs[k])
                vs_all.append(vs[k])
                rows.append((mode, lo, up, ma, mi, ge, go, int(r["scores"][k]), r["counts"][k], ops))
                kept += 1
                if kept == 60:
                    break
    H, V = sa.SeqSet.from_list(hs_all), sa.SeqSet.from_list(vs_all)
    lens = [len(x[9]) for x in rows]
    save("ref_dna_banded_end_rows.npz", H, V, cases=np.array([x[:7] for x in rows], np.int64),
         scores=np.array([x[7] for x in rows], np.int32), counts=np.stack([x[8] for x in rows]),
         op_off=np.cumsum([0] + lens).astype(np.int64),
         ops=np.frombuffer(b"".join(x[9] for x in rows), np.uint8))


def long_pairs():
    """100 kbp fixture: the first two pairs of the 64 x 100 kbp workload at
    full length (run on its own: --long).  Scores come from the reference's
    globalAlignmentScore.  Counts come from the oracle's traceback (a 10 GB
    trace matrix per pair): the reference's own traceback cannot give them
    at this size, because it passes the end cell's trace-matrix position as
    `unsigned` (dp_traceback_impl.h:559, 607), which truncates positions past
    2^32 (each pair has ~1e10 cells); its counts for these pairs violate
    aligned pairs + gap characters == residues."""
    H, V = sa.workload(sa.WL_DNA100K, 0x5EED0005, 0, 2)
    ref = ref_scores(H, V, 0, alphabet="dna", **SIMPLE)
    sc = po.Scoring(sa.simple_score(2, -3, -1, -5).table, -5, -1)
    ws, wc = po.traceback_batch(H.residues, H.offsets, V.residues, V.offsets, sc, 0, threads=2)
    assert np.array_equal(ws, ref), (ws, ref)
    save("ref_dna100k_counts.npz", H, V, scores=ref, counts=wc, seed=0x5EED0005, mode=0)


if __name__ == "__main__":
    if "--long" in sys.argv:
        long_pairs()
    elif "--modes" in sys.argv:
        mode_rows()
    elif "--local" in sys.argv:
        local_rows()
    elif "--linear" in sys.argv:
        linear_rows()
    elif "--banded-rows" in sys.argv:
        banded_rows()
        banded_end_rows()
    elif "--banded" in sys.argv:
        banded_scores()
    else:
        main()
        mode_rows()
        local_rows()
        linear_rows()
        banded_scores()
        banded_rows()
        banded_end_rows()
