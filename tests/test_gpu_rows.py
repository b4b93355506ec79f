"""Alignment rows (sqmi_align_rows): the rows the reference's globalAlignment
writes into its Gaps objects (traceback dp_traceback_impl.h:115-546), one
byte per alignment column: M (residues of H and V), V (a gap in the H row),
H (a gap in the V row).

Checked against the reference's own rows (fixtures written by the compiled
reference), the oracle's rows (pinned to those fixtures by the CPU tests),
and what any alignment's rows must satisfy: they hold every residue of both
sequences in order, and the counts read from them are the counts
sqmi_align_counts returns.  Local rows (sqmi_align_local_rows, the
reference's localAlignment on Gaps sets) are checked the same way, with the
window begins."""
import os

import numpy as np
import pytest

import pyoracle as po
import seqan_amd as sa

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M, V, H = ord("M"), ord("V"), ord("H")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _score_of(z, fixture):
    if "match" in z:
        return sa.simple_score(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    if "protein" in fixture:
        return sa.blosum62(int(z["gap_extend"]), int(z["gap_open"]))
    return sa.simple_score(2, -3, -1, -5)


def _counts_from_rows(Hs, Vs, op_off, ops, A):
    """Counts read from the rows the way computeAlignmentStats reads them
    (evaluate_alignment.h:215-300), plus a check that the rows hold both
    sequences."""
    out = np.zeros((len(Hs), A * A + 2), np.int64)
    for k in range(len(Hs)):
        row = ops[op_off[k]:op_off[k + 1]]
        h, v = Hs[k], Vs[k]
        assert (row == M).sum() + (row == H).sum() == len(h)
        assert (row == M).sum() + (row == V).sum() == len(v)
        hi = vi = 0
        prev = 0
        for op in row:
            if op == M:
                out[k, int(h[hi]) * A + int(v[vi])] += 1
                hi += 1
                vi += 1
            else:
                out[k, A * A + (1 if prev == op else 0)] += 1
                if op == V:
                    vi += 1
                else:
                    hi += 1
            prev = op if op != M else 0
    return out


def _oracle_rows(Hs, Vs, score):
    return po.rows_batch(Hs.residues, Hs.offsets, Vs.residues, Vs.offsets,
                         po.Scoring(score.table, score.gap_open, score.gap_extend), 0)


@pytest.mark.parametrize("fixture", ["ref_dna1k_counts.npz", "ref_dna_short_counts.npz", "ref_protein_counts.npz"])
@pytest.mark.parametrize("flags", [0, sa.FORCE_INT32], ids=["packed", "int32"])
def test_rows_match_reference_fixtures(ctx, fixture, flags):
    z = _load(fixture)
    score = _score_of(z, fixture)
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    print("rows fixture", fixture, "flags", flags, flush=True)
    s, c, op_off, ops = ctx.align_rows(Hs, Vs, score, flags=flags)
    assert np.array_equal(s, z["scores"])
    assert np.array_equal(c, z["counts"])
    assert np.array_equal(op_off, z["op_off"])
    assert np.array_equal(ops, z["ops"])


def test_rows_match_oracle_and_counts(ctx):
    rng = np.random.default_rng(21)
    hs, vs = [], []
    for _ in range(300):
        n = int(rng.integers(1, 900))
        h = rng.integers(0, 4, n).astype(np.uint8)
        v = h.copy()
        flip = rng.random(n) < 0.15
        v[flip] = rng.integers(0, 4, int(flip.sum()))
        cut = np.nonzero(rng.random(n) < 0.03)[0]
        v = np.delete(v, cut)
        if rng.random() < 0.2:
            v = rng.integers(0, 4, int(rng.integers(1, 900))).astype(np.uint8)
        hs.append(h)
        vs.append(v)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    for score in (sa.simple_score(2, -3, -1, -5), sa.simple_score(1, -1, -1, -2), sa.simple_score(3, -2, -3, -2)):
        print("rows vs oracle", score.table[0, :2], score.gap_extend, score.gap_open, flush=True)
        s, c, op_off, ops = ctx.align_rows(Hs, Vs, score)
        ws, wc, woff, wops = _oracle_rows(Hs, Vs, score)
        assert np.array_equal(s, ws) and np.array_equal(c, wc)
        assert np.array_equal(op_off, woff) and np.array_equal(ops, wops)
        s2, c2 = ctx.align_counts(Hs, Vs, score)
        assert np.array_equal(s2, s) and np.array_equal(c2, c)
        assert np.array_equal(_counts_from_rows(hs, vs, op_off, ops, 4), c)


@pytest.mark.parametrize("route", ["int32", "packed"])
def test_rows_long_pairs(ctx, monkeypatch, route):
    # The long-pair route (threshold lowered to 256) walks windows of
    # recomputed tiles; its rows are written across traceback rounds.
    monkeypatch.setenv("SQMI_LONG_MIN_SIDE", "256")
    flags = sa.FORCE_INT32
    if route == "packed":
        monkeypatch.setenv("SQMI_LONG_ALL", "1")
        flags = 0
    rng = np.random.default_rng(22)
    hs, vs = [], []
    for _ in range(20):
        n = int(rng.integers(256, 2500))
        h = rng.integers(0, 4, n).astype(np.uint8)
        v = np.delete(h, np.nonzero(rng.random(n) < 0.02)[0])
        flip = rng.random(len(v)) < 0.1
        v[flip] = rng.integers(0, 4, int(flip.sum()))
        hs.append(h)
        vs.append(v if len(v) >= 256 else h.copy())
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    score = sa.simple_score(2, -3, -1, -5)
    s, c, op_off, ops = ctx.align_rows(Hs, Vs, score, flags=flags)
    assert ctx.stats()["pairs_long"] == len(hs)
    ws, wc, woff, wops = _oracle_rows(Hs, Vs, score)
    assert np.array_equal(s, ws) and np.array_equal(c, wc)
    assert np.array_equal(op_off, woff) and np.array_equal(ops, wops)


def test_rows_100kbp_fixture(ctx):
    # Two 100 kbp pairs: the rows give the fixture's counts (reference scores,
    # oracle counts) and hold both sequences.
    z = _load("ref_dna100k_counts.npz")
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    score = sa.simple_score(2, -3, -1, -5)
    print("rows 100 kbp fixture", flush=True)
    s, c, op_off, ops = ctx.align_rows(Hs, Vs, score)
    assert np.array_equal(s, z["scores"]) and np.array_equal(c, z["counts"])
    assert np.array_equal(_counts_from_rows([Hs[0], Hs[1]], [Vs[0], Vs[1]], op_off, ops, 4), z["counts"])


def test_rows_empty_pairs_and_errors(ctx):
    Hs = sa.SeqSet.from_list([[0, 1, 2], [], [3, 3], [1]])
    Vs = sa.SeqSet.from_list([[0, 2], [1, 1], [], [1]])
    score = sa.simple_score(2, -3, -1, -5)
    s, c, op_off, ops = ctx.align_rows(Hs, Vs, score)
    assert s[1] == np.iinfo(np.int32).min and s[2] == np.iinfo(np.int32).min
    assert list(np.diff(op_off)) == [3, 0, 0, 1]
    assert bytes(ops[op_off[3]:op_off[4]]) == b"M"
    with pytest.raises(sa.SqmiError):
        ctx.align_rows(Hs, Vs, score, mode=sa.LOCAL)   # local rows: align_local_rows


@pytest.mark.parametrize("fixture", ["ref_dna_modes_rows.npz", "ref_dna_modes_short_rows.npz"])
@pytest.mark.parametrize("flags", [0, sa.FORCE_INT32], ids=["packed", "int32"])
def test_rows_every_align_config_match_reference(ctx, fixture, flags):
    # Free end gaps: the end cell moves to the best tracked cell and the rows
    # carry the free leading / trailing gaps (counted like any gap).  Packed:
    # the trace kernel tracks the end cell (kFwdTrace16 with bottom / right
    # tracking); int32: kCounts32.
    z = _load(fixture)
    score = _score_of(z, fixture)
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    nonempty = int(((Hs.lengths() > 0) & (Vs.lengths() > 0)).sum())
    for mode in range(16):
        print("rows, AlignConfig mode", mode, "flags", flags, flush=True)
        s, c, op_off, ops = ctx.align_rows(Hs, Vs, score, mode=mode, flags=flags)
        assert ctx.stats()["pairs_int16"] == (0 if flags else nonempty), mode
        assert np.array_equal(s, z["scores"][mode]), mode
        assert np.array_equal(c, z["counts"][mode]), mode
        assert np.array_equal(op_off, z["op_off"][mode]), mode
        assert np.array_equal(ops, z["ops"][z["mode_off"][mode]:z["mode_off"][mode + 1]]), mode
        s2, c2 = ctx.align_counts(Hs, Vs, score, mode=mode, flags=flags)
        assert np.array_equal(s2, s) and np.array_equal(c2, c), mode
        assert np.array_equal(ctx.align_score(Hs, Vs, score, mode), s), mode


def test_rows_align_configs_match_oracle(ctx):
    rng = np.random.default_rng(23)
    hs, vs = [], []
    for _ in range(150):
        h = rng.integers(0, 4, int(rng.integers(1, 600))).astype(np.uint8)
        v = h[int(rng.integers(0, len(h) // 2 + 1)):].copy()
        flip = rng.random(len(v)) < 0.12
        v[flip] = rng.integers(0, 4, int(flip.sum()))
        if len(v) == 0 or rng.random() < 0.2:
            v = rng.integers(0, 4, int(rng.integers(1, 600))).astype(np.uint8)
        if rng.random() < 0.5:
            h, v = v, h
        hs.append(h)
        vs.append(v)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    score = sa.simple_score(2, -3, -1, -5)
    for mode in (1, 2, 4, 8, 9, 6, 15):
        print("rows vs oracle, mode", mode, flush=True)
        s, c, op_off, ops = ctx.align_rows(Hs, Vs, score, mode=mode)
        ws, wc, woff, wops = po.rows_batch(Hs.residues, Hs.offsets, Vs.residues, Vs.offsets,
                                           po.Scoring(score.table, score.gap_open, score.gap_extend), mode)
        assert np.array_equal(s, ws) and np.array_equal(c, wc)
        assert np.array_equal(op_off, woff) and np.array_equal(ops, wops)
        assert np.array_equal(_counts_from_rows(hs, vs, op_off, ops, 4), c)


def _oracle_local(Hs, Vs, score):
    return po.rows_batch(Hs.residues, Hs.offsets, Vs.residues, Vs.offsets,
                         po.Scoring(score.table, score.gap_open, score.gap_extend), sa.LOCAL, with_begins=True)


@pytest.mark.parametrize("fixture", ["ref_dna_local_rows.npz", "ref_dna_local_short_rows.npz",
                                     "ref_protein_local_rows.npz"])
@pytest.mark.parametrize("flags", [0, sa.FORCE_INT32], ids=["packed", "int32"])
def test_local_rows_match_reference_fixtures(ctx, fixture, flags):
    # localAlignment(ExecutionPolicy<Parallel, Vectorial>, gapsH, gapsV,
    # score): scores, counts, clipped rows and window begins, from the packed
    # trace kernel (end cell tracked over every cell, walk stopped where H
    # falls to 0) and from the int32 kernel.
    z = _load(fixture)
    score = _score_of(z, fixture)
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    print("local rows fixture", fixture, "flags", flags, flush=True)
    s, c, op_off, ops, begins = ctx.align_local_rows(Hs, Vs, score, flags=flags)
    nonempty = int(((np.diff(z["h_off"]) > 0) & (np.diff(z["v_off"]) > 0)).sum())
    assert ctx.stats()["pairs_int16"] == (0 if flags else nonempty)
    assert np.array_equal(s, z["scores"])
    assert np.array_equal(c, z["counts"])
    assert np.array_equal(op_off, z["op_off"])
    assert np.array_equal(ops, z["ops"])
    assert np.array_equal(begins, z["begins"])
    s2, c2 = ctx.align_counts(Hs, Vs, score, mode=sa.LOCAL)
    assert np.array_equal(s2, s) and np.array_equal(c2, c)
    assert np.array_equal(ctx.align_score(Hs, Vs, score, sa.LOCAL), s)


@pytest.mark.parametrize("flags", [0, sa.FORCE_INT32], ids=["packed", "int32"])
def test_local_rows_match_oracle(ctx, flags):
    rng = np.random.default_rng(24)
    for alpha, score in ((4, sa.simple_score(2, -3, -1, -5)), (4, sa.simple_score(1, -1, -1, -2)),
                         (27, sa.blosum62(-1, -11))):
        hs, vs = [], []
        for _ in range(200):
            h = rng.integers(0, min(alpha, 24), int(rng.integers(1, 700))).astype(np.uint8)
            a = int(rng.integers(0, len(h)))
            seg = h[a:a + int(rng.integers(1, 300))].copy()
            flip = rng.random(len(seg)) < 0.15
            seg[flip] = rng.integers(0, min(alpha, 24), int(flip.sum()))
            flank = lambda: rng.integers(0, min(alpha, 24), int(rng.integers(0, 200))).astype(np.uint8)
            v = np.concatenate([flank(), seg, flank()]).astype(np.uint8)
            if rng.random() < 0.25:
                v = rng.integers(0, min(alpha, 24), int(rng.integers(1, 700))).astype(np.uint8)
            hs.append(h)
            vs.append(v)
        Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
        print("local rows vs oracle, alphabet", alpha, "flags", flags, flush=True)
        s, c, op_off, ops, begins = ctx.align_local_rows(Hs, Vs, score, flags=flags)
        assert ctx.stats()["pairs_int16"] == (0 if flags else len(hs))
        ws, wc, woff, wops, wbeg = _oracle_local(Hs, Vs, score)
        assert np.array_equal(s, ws) and np.array_equal(c, wc)
        assert np.array_equal(op_off, woff) and np.array_equal(ops, wops)
        assert np.array_equal(begins, wbeg)


@pytest.mark.parametrize("flags", [0, sa.FORCE_INT32], ids=["packed", "int32"])
def test_local_rows_empty_zero_and_errors(ctx, flags):
    # Empty pairs: INT32_MIN and no window.  A pair with no positive cell:
    # score 0, empty window at (0, 0).  Local rows come only from
    # sqmi_align_local_rows; free-end bits beside LOCAL are ignored, as in
    # sqmi_align_score (localAlignment has no AlignConfig).
    Hs = sa.SeqSet.from_list([[0, 1, 2], [], [0, 0, 0], [1, 2, 3, 0]])
    Vs = sa.SeqSet.from_list([[0, 2], [1, 1], [1, 2, 3], []])
    score = sa.simple_score(2, -3, -1, -5)
    s, c, op_off, ops, begins = ctx.align_local_rows(Hs, Vs, score, flags=flags)
    assert s[1] == np.iinfo(np.int32).min and s[3] == np.iinfo(np.int32).min
    assert s[2] == 0 and s[0] == 2
    assert list(np.diff(op_off)) == [1, 0, 0, 0]
    assert bytes(ops) == b"M"
    assert begins.tolist() == [[0, 0], [0, 0], [0, 0], [0, 0]]
    assert (c[1:] == 0).all()
    with pytest.raises(sa.SqmiError):
        ctx.align_rows(Hs, Vs, score, mode=sa.LOCAL)
    s2, c2 = ctx.align_counts(Hs, Vs, score, mode=sa.LOCAL | sa.FREE_TOP, flags=flags)
    assert np.array_equal(s2, s) and np.array_equal(c2, c)


@pytest.mark.parametrize("fixture", ["ref_dna_linear_rows.npz", "ref_dna_linear2_rows.npz",
                                     "ref_protein_linear_rows.npz"])
@pytest.mark.parametrize("flags", [0, sa.FORCE_INT32], ids=["packed", "int32"])
def test_linear_gap_rows_match_reference(ctx, fixture, flags):
    # gap open == gap extend: the reference's linear-gap traceback (its own
    # tie-breaks) for global modes under AlignConfigs and for local windows.
    z = _load(fixture)
    score = _score_of(z, fixture)
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    for k, mode in enumerate(z["modes"]):
        mode = int(mode)
        print("linear-gap rows", fixture, "mode", mode, "flags", flags, flush=True)
        if mode == sa.LOCAL:
            s, c, op_off, ops, begins = ctx.align_local_rows(Hs, Vs, score, flags=flags)
            assert np.array_equal(begins, z["begins"])
        else:
            s, c, op_off, ops = ctx.align_rows(Hs, Vs, score, mode=mode, flags=flags)
        assert np.array_equal(s, z["scores"][k]), mode
        assert np.array_equal(c, z["counts"][k]), mode
        assert np.array_equal(op_off, z["op_off"][k]), mode
        assert np.array_equal(ops, z["ops"][z["mode_off"][k]:z["mode_off"][k + 1]]), mode


@pytest.mark.parametrize("scoring", [(2, -3, -1, -5), (1, -1, -1, -2), (1, -1, -1, -1), (3, -2, -3, -2)],
                         ids=["affine", "ties", "linear", "extend-below-open"])
def test_rows_free_ends_packed_match_oracle(ctx, scoring):
    # The packed free-end traceback over items of several 256-row blocks,
    # where the last column's candidates come from several blocks and lanes
    # and the two pairs of a couple end at different columns.
    rng = np.random.default_rng(sum(abs(x) for x in scoring) * 7 + scoring[0])
    hs, vs = [], []
    for _ in range(160):
        n = int(rng.integers(1, 1400))
        h = rng.integers(0, 4, n).astype(np.uint8)
        lo = int(rng.integers(0, n // 3 + 1))
        hi = n - int(rng.integers(0, n // 3 + 1))
        v = h[lo:max(hi, lo + 1)].copy()
        flip = rng.random(len(v)) < 0.1
        v[flip] = rng.integers(0, 4, int(flip.sum()))
        if rng.random() < 0.15:
            v = rng.integers(0, 4, int(rng.integers(1, 1400))).astype(np.uint8)
        if rng.random() < 0.5:
            h, v = v, h
        hs.append(h)
        vs.append(v)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    score = sa.simple_score(*scoring)
    sc = po.Scoring(score.table, score.gap_open, score.gap_extend)
    for mode in (1, 2, 4, 8, 9, 12, 6, 15):
        print("packed free-end rows vs oracle, scoring", scoring, "mode", mode, flush=True)
        s, c, op_off, ops = ctx.align_rows(Hs, Vs, score, mode=mode)
        assert ctx.stats()["pairs_int16"] == len(hs)
        ws, wc, woff, wops = po.rows_batch(Hs.residues, Hs.offsets, Vs.residues, Vs.offsets, sc, mode)
        bad = np.nonzero(s != ws)[0]
        assert len(bad) == 0, (mode, bad[:5])
        bad = np.nonzero((c != wc).any(axis=1))[0]
        assert len(bad) == 0, (mode, bad[:5])
        assert np.array_equal(op_off, woff) and np.array_equal(ops, wops)


@pytest.mark.parametrize("scoring", [(1, -1, -1, -2), (1, -1, -1, -1), (2, -1, -2, -2), (3, -2, -3, -2)],
                         ids=["ties", "linear", "cheap-gaps", "extend-below-open"])
def test_local_rows_packed_ties_multi_block(ctx, scoring):
    # Local windows on the packed trace kernel with many equal maxima: short
    # repeats of a small alphabet, V up to 1,400 residues (several 256-row
    # blocks, so equal maxima in one column come from different blocks and
    # lanes), compared with the oracle; the packed route is asserted.
    rng = np.random.default_rng(sum(scoring) + 50)
    score = sa.simple_score(*scoring)
    hs, vs = [], []
    for _ in range(160):
        unit = rng.integers(0, 4, int(rng.integers(1, 5))).astype(np.uint8)
        n, m = int(rng.integers(1, 500)), int(rng.integers(1, 1400))
        h = np.resize(unit, n).copy()
        v = np.resize(unit, m).copy()
        for s in (h, v):
            f = rng.random(len(s)) < 0.05
            s[f] = rng.integers(0, 4, int(f.sum()))
        hs.append(h)
        vs.append(v)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    print("packed local ties, scoring", scoring, flush=True)
    s, c, off, ops, begins = ctx.align_local_rows(Hs, Vs, score)
    assert ctx.stats()["pairs_int16"] == len(hs)
    ws, wc, woff, wops, wbeg = _oracle_local(Hs, Vs, score)
    assert np.array_equal(s, ws) and np.array_equal(c, wc)
    assert np.array_equal(off, woff) and np.array_equal(ops, wops)
    assert np.array_equal(begins, wbeg)


@pytest.mark.parametrize("where", ["host", "device"])
def test_rows_segment_and_packed_layouts(ctx, where):
    """The two layouts of sqmi_align_rows: each row at the end of the
    caller's segment (the default, which the SeqAn Gaps adapter reads) and
    back to back from op_offsets[0] (SQMI_ROWS_PACKED, packed on the device),
    with the outputs in host memory or in device memory."""
    import ctypes
    fx = "ref_dna_short_counts.npz"
    z = _load(fx)
    score = _score_of(z, fx)
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    n, A = len(Hs.offsets) - 1, score.alphabet_size
    ch, _kh = sa._seqs(Hs)
    cv, _kv = sa._seqs(Vs)
    sc = score._c()
    seg = np.zeros(n + 1, np.int64)
    seg[1:] = np.cumsum(Hs.lengths() + Vs.lengths() + 3)
    seg += 5   # op_offsets[0] > 0 and slack in every segment
    total = int(seg[-1])
    got = {}
    for flag in (0, sa.ROWS_PACKED):
        print("rows layout", where, hex(flag), flush=True)
        scores = np.empty(n, np.int32)
        counts = np.empty((n, A * A + 2), np.int32)
        if where == "host":
            buf = np.zeros(total, np.uint8)
            lens = np.zeros(n, np.int64)
            st = ctx._L.sqmi_align_rows(ctx._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), 0,
                                        scores.ctypes.data, counts.ctypes.data, buf.ctypes.data, seg.ctypes.data,
                                        lens.ctypes.data, flag)
        else:
            dev = [sa.DeviceBuffer(4 * n), sa.DeviceBuffer(counts.nbytes), sa.DeviceBuffer(total),
                   sa.DeviceBuffer(8 * n)]
            st = ctx._L.sqmi_align_rows(ctx._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), 0,
                                        dev[0].address, dev[1].address, dev[2].address, seg.ctypes.data,
                                        dev[3].address, flag | sa.SCORES_ON_DEVICE)
            scores = dev[0].to_array(np.int32, n)
            buf = dev[2].to_array(np.uint8, total)
            lens = dev[3].to_array(np.int64, n)
            for b in dev:
                b.free()
        assert st == 0
        assert np.array_equal(scores, z["scores"])
        got[flag] = (buf, lens)
    (buf0, lens0), (buf1, lens1) = got[0], got[sa.ROWS_PACKED]
    assert np.array_equal(lens0, lens1) and np.array_equal(lens0, np.diff(z["op_off"]))
    at_ends = np.concatenate([buf0[seg[i + 1] - lens0[i]:seg[i + 1]] for i in range(n)])
    packed = buf1[seg[0]:seg[0] + int(lens1.sum())]
    assert np.array_equal(at_ends, z["ops"]) and np.array_equal(packed, z["ops"])
    if where == "host":
        # Only the packed rows come back: the bytes around them keep their values.
        assert not buf1[:seg[0]].any() and not buf1[seg[0] + int(lens1.sum()):].any()
