"""Backward (sqmi_align_counts): score plus substitution / gap-open /
gap-extension counts along the path the reference's traceback picks
(dp_traceback_impl.h:115-546, counted like evaluate_alignment.h:215-300).

Checked against the reference's own counts (fixtures written by the compiled
reference), the oracle's traceback, and properties every alignment satisfies:
sum(count * parameter) == score, and aligned pairs + gap characters account
for every residue of both sequences."""
import os

import numpy as np
import pytest

import pyoracle as po
import seqan_amd as sa

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _params(score: sa.Score):
    return np.concatenate([score.table.ravel(), [score.gap_open, score.gap_extend]]).astype(np.int64)


def _check_invariants(H, V, score, scores, counts):
    A = score.alphabet_size
    ok = scores != np.iinfo(np.int32).min
    assert np.array_equal((counts[ok].astype(np.int64) @ _params(score)), scores[ok].astype(np.int64))
    matched = counts[:, :A * A].sum(axis=1).astype(np.int64)
    gaps = counts[:, A * A].astype(np.int64) + counts[:, A * A + 1]
    n, m = H.lengths(), V.lengths()
    assert np.array_equal(gaps[ok], (n + m - 2 * matched)[ok])
    assert (counts >= 0).all()
    assert not counts[~ok].any()


def _oracle(H, V, score):
    return po.traceback_batch(H.residues, H.offsets, V.residues, V.offsets,
                              po.Scoring(score.table, score.gap_open, score.gap_extend), 0)


@pytest.mark.parametrize("fixture", ["ref_dna1k_counts.npz", "ref_dna_short_counts.npz"])
def test_counts_match_reference_fixtures(ctx, fixture):
    z = _load(fixture)
    if "match" in z:
        score = sa.simple_score(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    else:
        score = sa.simple_score(2, -3, -1, -5)
    H = sa.SeqSet(z["h_res"], z["h_off"])
    V = sa.SeqSet(z["v_res"], z["v_off"])
    print("fixture", fixture, len(z["scores"]), "pairs", flush=True)
    s, c = ctx.align_counts(H, V, score)
    assert np.array_equal(s, z["scores"])
    assert np.array_equal(c, z["counts"])
    s32, c32 = ctx.align_counts(H, V, score, flags=sa.FORCE_INT32)
    assert np.array_equal(s32, z["scores"])
    assert np.array_equal(c32, z["counts"])


def _random_pairs(rng, count, lo, hi, alpha, sub=0.1, indel=0.03):
    hs, vs = [], []
    for _ in range(count):
        n = int(rng.integers(lo, hi + 1))
        h = rng.integers(0, alpha, n).astype(np.uint8)
        v = []
        for x in h:
            r = rng.random()
            if r < indel / 2:
                continue
            if r < indel:
                v.append(int(rng.integers(0, alpha)))
            v.append(int(rng.integers(0, alpha)) if rng.random() < sub else int(x))
        if rng.random() < 0.15:
            v = list(rng.integers(0, alpha, int(rng.integers(lo, hi + 1))))
        hs.append(h)
        vs.append(np.array(v, dtype=np.uint8))
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


@pytest.mark.parametrize("lo,hi", [(1, 40), (200, 300), (500, 1300)])
def test_counts_match_oracle_dna(ctx, lo, hi):
    rng = np.random.default_rng(lo * 7 + hi)
    H, V = _random_pairs(rng, 200, lo, hi, 4)
    for score in (sa.simple_score(2, -3, -1, -5), sa.simple_score(1, -1, -1, -2), sa.simple_score(5, -4, -1, -11)):
        print("dna", lo, hi, score.table[0, :2], score.gap_open, flush=True)
        s, c = ctx.align_counts(H, V, score)
        ws, wc = _oracle(H, V, score)
        assert np.array_equal(s, ws)
        bad = np.nonzero((c != wc).any(axis=1))[0]
        assert len(bad) == 0, (bad[:5], c[bad[:1]], wc[bad[:1]])
        _check_invariants(H, V, score, s, c)


def test_counts_match_oracle_protein(ctx):
    rng = np.random.default_rng(5)
    H, V = _random_pairs(rng, 120, 20, 700, 24, sub=0.4)
    score = sa.blosum62(-1, -11)
    print("protein blosum62", flush=True)
    s, c = ctx.align_counts(H, V, score)
    ws, wc = _oracle(H, V, score)
    assert np.array_equal(s, ws)
    assert np.array_equal(c, wc)
    _check_invariants(H, V, score, s, c)


def test_counts_empty_and_mixed_routes(ctx):
    # Empty sequences give INT32_MIN and zero counts; a pair outside the
    # int16 window (match 200 over 300 bp) takes the int32 kernel.
    H = sa.SeqSet.from_list([[0, 1, 2], [], [1] * 300, [2, 2], [3]])
    V = sa.SeqSet.from_list([[0, 2], [1, 1], [1] * 300, [], [3]])
    score = sa.simple_score(200, -3, -1, -5)
    s, c = ctx.align_counts(H, V, score)
    ws, wc = _oracle(H, V, score)
    assert np.array_equal(s, ws)
    assert np.array_equal(c, wc)
    st = ctx.stats()
    assert st["pairs_int32"] >= 1 and st["pairs_int16"] >= 1


def test_counts_chunked_trace_budget(ctx, monkeypatch):
    rng = np.random.default_rng(9)
    H, V = _random_pairs(rng, 300, 100, 900, 4)
    score = sa.simple_score(2, -3, -1, -5)
    whole = ctx.align_counts(H, V, score)
    monkeypatch.setenv("SQMI_TRACE_BUDGET_MB", "2")
    print("chunked: 2 MiB trace budget", flush=True)
    chunked = ctx.align_counts(H, V, score)
    chunked32 = ctx.align_counts(H, V, score, flags=sa.FORCE_INT32)
    assert np.array_equal(whole[0], chunked[0]) and np.array_equal(whole[1], chunked[1])
    assert np.array_equal(whole[0], chunked32[0]) and np.array_equal(whole[1], chunked32[1])


def test_counts_scores_equal_forward_scores(ctx):
    H, V = sa.workload(sa.WL_DNA1K, 0x5EED0004, 0, 4096)
    score = sa.simple_score(2, -3, -1, -5)
    s, c = ctx.align_counts(H, V, score)
    assert np.array_equal(s, ctx.align_score(H, V, score, 0))
    _check_invariants(H, V, score, s, c)
    ws, wc = _oracle(sa.SeqSet(H.residues[:H.offsets[512]], H.offsets[:513]),
                     sa.SeqSet(V.residues[:V.offsets[512]], V.offsets[:513]), score)
    assert np.array_equal(s[:512], ws) and np.array_equal(c[:512], wc)


def test_counts_full_dna1k_workload_invariants(ctx):
    # The backward workload at full size: 65,536 pairs of 1 kbp.
    H, V = sa.workload(sa.WL_DNA1K, 0x5EED0004, 0, 65536)
    score = sa.simple_score(2, -3, -1, -5)
    print("dna1k full: 65536 pairs", flush=True)
    s, c = ctx.align_counts(H, V, score)
    print("dna1k full stats", ctx.stats(), flush=True)
    assert np.array_equal(s, ctx.align_score(H, V, score, 0))
    _check_invariants(H, V, score, s, c)
