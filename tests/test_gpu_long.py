"""Long pairs: the tiled int32 forward that keeps every tile border (256 rows
x 256 columns) as a checkpoint, and the checkpointed traceback that walks
back through windows of recomputed tiles.

Global DNA pairs outside the int16 window with both sides >=
SQMI_LONG_MIN_SIDE (2048 by default) take this route; its forward runs in
packed int16 relative to a per-tile base when the scoring allows, else in
int32.  The parity tests lower the threshold to 256 and run every case on
both forwards: int32 (FORCE_INT32) and packed (SQMI_LONG_ALL=1, which sends
pairs the packed kernels could score down the long route too), so that
pairs of a few thousand bases, which the oracle finishes in seconds, cross
many row blocks and column tiles; the full 64 x 100 kbp workload is checked through
properties every alignment satisfies (sum of count x parameter == score,
aligned pairs + gap characters == residues), against the packed kernels on
its first pairs cut short, and against a fixture of its first two pairs."""
import os

import numpy as np
import pytest

import pyoracle as po
import seqan_amd as sa

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

F32 = sa.FORCE_INT32
# The last two have gap open == gap extend: the reference's linear-gap walk.
# (The long route takes DNA scorings with mismatch >= gap open.)
SCORES = [sa.simple_score(2, -3, -1, -5), sa.simple_score(1, -1, -1, -2), sa.simple_score(5, -4, -2, -11),
          sa.simple_score(3, -2, -3, -2), sa.simple_score(3, -2, -2, -2), sa.simple_score(1, -1, -1, -1)]


def _params(score):
    return np.concatenate([score.table.ravel(), [score.gap_open, score.gap_extend]]).astype(np.int64)


def _check_invariants(H, V, score, scores, counts):
    # sum(count x parameter) == score needs gapExtend >= gapOpen: otherwise
    # the DP may score two adjacent gap characters as two opens, which the
    # run-length counts (evaluate_alignment.h:215-300) report as open + extend.
    A = score.alphabet_size
    if score.gap_extend >= score.gap_open:
        assert np.array_equal(counts.astype(np.int64) @ _params(score), scores.astype(np.int64))
    matched = counts[:, :A * A].sum(axis=1).astype(np.int64)
    gaps = counts[:, A * A].astype(np.int64) + counts[:, A * A + 1]
    assert np.array_equal(gaps, H.lengths() + V.lengths() - 2 * matched)
    assert (counts >= 0).all()


def _mutated_pairs(rng, count, lo, hi, sub=0.08, indel=0.02, unrelated=0.2):
    hs, vs = [], []
    for _ in range(count):
        n = int(rng.integers(lo, hi + 1))
        h = rng.integers(0, 4, n).astype(np.uint8)
        if rng.random() < unrelated:
            v = rng.integers(0, 4, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        else:
            keep = rng.random(n) >= indel / 2
            v = h[keep].copy()
            flip = rng.random(len(v)) < sub
            v[flip] = rng.integers(0, 4, int(flip.sum()))
            ins = np.nonzero(rng.random(len(v)) < indel / 2)[0]
            v = np.insert(v, ins, rng.integers(0, 4, len(ins)).astype(np.uint8))
        hs.append(h)
        vs.append(v.astype(np.uint8))
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


def _oracle_scores(H, V, score):
    return po.score_batch(H.residues, H.offsets, V.residues, V.offsets,
                          po.Scoring(score.table, score.gap_open, score.gap_extend), 0)


def _oracle_counts(H, V, score):
    return po.traceback_batch(H.residues, H.offsets, V.residues, V.offsets,
                              po.Scoring(score.table, score.gap_open, score.gap_extend), 0)


@pytest.fixture(params=["int32", "packed"])
def route(request, monkeypatch):
    """Flags for the long route with threshold 256 and the given forward."""
    monkeypatch.setenv("SQMI_LONG_MIN_SIDE", "256")
    if request.param == "packed":
        monkeypatch.setenv("SQMI_LONG_ALL", "1")
        return 0
    return F32


def _check_route(ctx, count, flags):
    st = ctx.stats()
    assert st["pairs_long"] == count
    assert st["pairs_long_packed"] == (0 if flags & F32 else count)


@pytest.mark.parametrize("lo,hi", [(256, 700), (900, 2600)])
def test_long_scores_match_oracle(ctx, route, lo, hi):
    rng = np.random.default_rng(lo + hi)
    H, V = _mutated_pairs(rng, 40, lo, hi)
    for score in SCORES:
        print("long scores", lo, hi, score.table[0, :2], score.gap_extend, score.gap_open, flush=True)
        s = ctx.align_score(H, V, score, 0, flags=route)
        _check_route(ctx, len(H), route)
        assert np.array_equal(s, _oracle_scores(H, V, score))


@pytest.mark.parametrize("lo,hi", [(256, 700), (900, 2600)])
def test_long_counts_match_oracle(ctx, route, lo, hi):
    rng = np.random.default_rng(3 * lo + hi)
    H, V = _mutated_pairs(rng, 40, lo, hi)
    for score in SCORES:
        print("long counts", lo, hi, score.table[0, :2], score.gap_extend, score.gap_open, flush=True)
        s, c = ctx.align_counts(H, V, score, flags=route)
        _check_route(ctx, len(H), route)
        ws, wc = _oracle_counts(H, V, score)
        assert np.array_equal(s, ws)
        bad = np.nonzero((c != wc).any(axis=1))[0]
        assert len(bad) == 0, (bad[:5], c[bad[:1]], wc[bad[:1]])
        _check_invariants(H, V, score, s, c)


def test_long_edges(ctx, route):
    # Sides at and around the block (256 rows) and tile (256 columns) edges,
    # very unequal sides, and identical / unrelated sequences.
    rng = np.random.default_rng(11)
    shapes = [(256, 256), (511, 257), (512, 255 + 256), (513, 767), (1024, 1023), (3000, 260), (260, 3000),
              (2049, 2048), (700, 1800)]
    hs, vs = [], []
    for n, m in shapes:
        hs.append(rng.integers(0, 4, n).astype(np.uint8))
        vs.append(rng.integers(0, 4, m).astype(np.uint8))
    same = rng.integers(0, 4, 1500).astype(np.uint8)
    hs.append(same)
    vs.append(same.copy())
    H, V = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    for score in SCORES[:2]:
        print("long edges", score.table[0, :2], flush=True)
        s, c = ctx.align_counts(H, V, score, flags=route)
        _check_route(ctx, len(H), route)
        ws, wc = _oracle_counts(H, V, score)
        assert np.array_equal(s, ws)
        assert np.array_equal(c, wc)
        assert np.array_equal(ctx.align_score(H, V, score, 0, flags=route), ws)


def test_long_mixed_with_packed_routes(ctx, route):
    # Long pairs beside short ones, an empty pair, and score calls under
    # free end gaps and local alignment (the long route's int32 forward; its
    # packed forward is global only; the traceback is global only).
    rng = np.random.default_rng(12)
    Hl, Vl = _mutated_pairs(rng, 6, 600, 1200)
    hs = [Hl[i] for i in range(6)] + [rng.integers(0, 4, 100).astype(np.uint8), np.zeros(0, np.uint8)]
    vs = [Vl[i] for i in range(6)] + [rng.integers(0, 4, 90).astype(np.uint8), rng.integers(0, 4, 5).astype(np.uint8)]
    H, V = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    score = SCORES[0]
    s, c = ctx.align_counts(H, V, score, flags=route)
    _check_route(ctx, 6, route)
    ws, wc = _oracle_counts(H, V, score)
    assert np.array_equal(s, ws) and np.array_equal(c, wc)
    for mode in (0, 9, 15, sa.LOCAL):
        print("mixed mode", mode, flush=True)
        want = po.score_batch(H.residues, H.offsets, V.residues, V.offsets,
                              po.Scoring(score.table, score.gap_open, score.gap_extend), mode)
        assert np.array_equal(ctx.align_score(H, V, score, mode, flags=route), want)
        st = ctx.stats()
        assert st["pairs_long"] == 6
        assert st["pairs_long_packed"] == (6 if mode == 0 and not route & F32 else 0)


def test_long_chunked_checkpoints(ctx, route, monkeypatch):
    # A 1 MiB budget puts every pair in a chunk of its own (and an odd
    # pair count leaves the last couple of the packed forward half empty).
    rng = np.random.default_rng(13)
    H, V = _mutated_pairs(rng, 11, 800, 1600)
    score = SCORES[0]
    whole = ctx.align_counts(H, V, score, flags=route)
    _check_route(ctx, len(H), route)
    monkeypatch.setenv("SQMI_TRACE_BUDGET_MB", "1")
    print("long chunked", flush=True)
    chunked = ctx.align_counts(H, V, score, flags=route)
    assert np.array_equal(whole[0], chunked[0]) and np.array_equal(whole[1], chunked[1])
    assert np.array_equal(ctx.align_score(H, V, score, 0, flags=route), whole[0])
    ws, wc = _oracle_counts(H, V, score)
    assert np.array_equal(whole[0], ws) and np.array_equal(whole[1], wc)


def test_long_workload_prefix_matches_packed(ctx, monkeypatch):
    # The 100 kbp workload's first 16 pairs cut to 4 kbp: long route (int32
    # forced) against the packed int16 kernels.
    H, V = sa.workload(sa.WL_DNA100K, 0x5EED0005, 0, 16, 4000)
    score = SCORES[0]
    s_long, c_long = ctx.align_counts(H, V, score, flags=F32)
    assert ctx.stats()["pairs_long"] == 16
    s_pk, c_pk = ctx.align_counts(H, V, score)
    assert ctx.stats()["pairs_long"] == 0 and ctx.stats()["pairs_int16"] == 16
    assert np.array_equal(s_long, s_pk) and np.array_equal(c_long, c_pk)
    monkeypatch.setenv("SQMI_LONG_ALL", "1")
    s_l16, c_l16 = ctx.align_counts(H, V, score)
    assert ctx.stats()["pairs_long_packed"] == 16
    assert np.array_equal(s_l16, s_pk) and np.array_equal(c_l16, c_pk)


@pytest.mark.parametrize("flags", [0, F32], ids=["packed", "int32"])
def test_long_100kbp_fixture(ctx, flags):
    # The 100 kbp workload's first two pairs at full length: scores from the
    # reference's globalAlignmentScore, counts from the oracle's traceback.
    z = np.load(os.path.join(GOLDEN, "ref_dna100k_counts.npz"), allow_pickle=False)
    H, V = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    score = SCORES[0]
    print("100 kbp fixture: 2 pairs, flags", flags, flush=True)
    assert np.array_equal(ctx.align_score(H, V, score, 0, flags=flags), z["scores"])
    _check_route(ctx, 2, flags)
    s, c = ctx.align_counts(H, V, score, flags=flags)
    assert np.array_equal(s, z["scores"]) and np.array_equal(c, z["counts"])


def test_long_full_100kbp_workload(ctx):
    # 64 pairs of 100 kbp: scores and counts by the long route, checked by
    # the alignment invariants and score == forward score.
    H, V = sa.workload(sa.WL_DNA100K, 0x5EED0005, 0, 64)
    score = SCORES[0]
    print("100 kbp: 64 pairs, forward", flush=True)
    s = ctx.align_score(H, V, score, 0)
    st = ctx.stats()
    print("100 kbp forward stats", st, flush=True)
    _check_route(ctx, 64, 0)
    assert np.array_equal(ctx.align_score(H, V, score, 0, flags=F32), s)
    _check_route(ctx, 64, F32)
    print("100 kbp: backward", flush=True)
    s2, c = ctx.align_counts(H, V, score)
    print("100 kbp backward stats", ctx.stats(), flush=True)
    assert np.array_equal(s, s2)
    _check_invariants(H, V, score, s2, c)


@pytest.mark.parametrize("lo,hi", [(256, 700), (900, 2600)])
def test_long_scores_every_alignconfig_and_local(ctx, monkeypatch, lo, hi):
    # Scores of long pairs under every AlignConfig and local alignment run
    # the tiled int32 forward (free left / top borders in the sweep; the free
    # last row / column reduced from the checkpoints; local with every H
    # floored at 0 and the running maximum of every tile).
    monkeypatch.setenv("SQMI_LONG_MIN_SIDE", "256")
    rng = np.random.default_rng(5 * lo + hi)
    H, V = _mutated_pairs(rng, 40, lo, hi)
    for score in SCORES[:3]:
        sc = po.Scoring(score.table, score.gap_open, score.gap_extend)
        for mode in list(range(16)) + [sa.LOCAL]:
            print("long mode", mode, lo, hi, score.table[0, :2], score.gap_extend, score.gap_open, flush=True)
            s = ctx.align_score(H, V, score, mode, flags=F32)
            assert ctx.stats()["pairs_long"] == len(H), mode
            want = po.score_batch(H.residues, H.offsets, V.residues, V.offsets, sc, mode)
            bad = np.nonzero(s != want)[0]
            assert len(bad) == 0, (mode, bad[:5], s[bad[:3]], want[bad[:3]])


def test_long_local_and_free_ends_edges(ctx, monkeypatch):
    # Unrelated pairs (small local scores), identical pairs, very unequal
    # sides, sides at the block / tile edges, under local and the free-end
    # configs that reduce the last row / column.
    monkeypatch.setenv("SQMI_LONG_MIN_SIDE", "256")
    rng = np.random.default_rng(23)
    shapes = [(256, 256), (511, 257), (513, 767), (3000, 260), (260, 3000), (1024, 1023)]
    hs, vs = [], []
    for n, m in shapes:
        hs.append(rng.integers(0, 4, n).astype(np.uint8))
        vs.append(rng.integers(0, 4, m).astype(np.uint8))
    same = rng.integers(0, 4, 1500).astype(np.uint8)
    hs.append(same)
    vs.append(same.copy())
    H, V = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    for score in SCORES[:2]:
        sc = po.Scoring(score.table, score.gap_open, score.gap_extend)
        for mode in (sa.LOCAL, 2, 4, 6, 9, 15):
            print("long edges mode", mode, flush=True)
            s = ctx.align_score(H, V, score, mode, flags=F32)
            assert ctx.stats()["pairs_long"] == len(H)
            want = po.score_batch(H.residues, H.offsets, V.residues, V.offsets, sc, mode)
            assert np.array_equal(s, want), mode


def test_long_streaming_forward_matches_diagonal_launches(ctx, monkeypatch):
    """The packed long-pair forward as one persistent launch (SQMI_LONG_STREAM=1:
    tiles claimed in anti-diagonal order, each waiting for its two producer
    tiles) against the same forward launched once per tile anti-diagonal
    (the default) and the oracle: an odd number of pairs (the last couple
    holds one), sides from 256 to 3,000 and couples of unequal shapes; then
    the counts after the persistent forward."""
    monkeypatch.setenv("SQMI_LONG_MIN_SIDE", "256")
    monkeypatch.setenv("SQMI_LONG_ALL", "1")
    rng = np.random.default_rng(77)
    H, V = _mutated_pairs(rng, 37, 256, 3000)
    score = SCORES[0]
    print("long streaming: 37 pairs", flush=True)
    per_diag = ctx.align_score(H, V, score, 0)
    monkeypatch.setenv("SQMI_LONG_STREAM", "1")
    got = ctx.align_score(H, V, score, 0)
    _check_route(ctx, 37, 0)
    assert np.array_equal(got, per_diag)
    assert np.array_equal(got, _oracle_scores(H, V, score))
    s, c = ctx.align_counts(H, V, score)
    ws, wc = _oracle_counts(H, V, score)
    assert np.array_equal(s, got) and np.array_equal(s, ws)
    assert np.array_equal(c, wc)
