"""Pins the oracle (oracle/oracle.c) before it is trusted as the GPU checker.

Known answers come from the reference's own tests (file:line cited), from its
pair_align app tests, and from fixtures produced by the compiled reference
(tests/golden/make_golden.py).  Brute-force enumeration and closed forms
cover what those leave open.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DNA5 = "ACGTN"
AA = "ABCDEFGHIJKLMNOPQRSTUVWYZX*"


def enc(s, alphabet=DNA5):
    return np.array([alphabet.index(c) for c in s], dtype=np.uint8)


def rows_from_ops(h, v, ops, alphabet=DNA5):
    rh, rv, i, j = [], [], 0, 0
    for op in ops:
        if op == "M":
            rh.append(alphabet[h[j]]); rv.append(alphabet[v[i]]); i += 1; j += 1
        elif op == "H":
            rh.append(alphabet[h[j]]); rv.append("-"); j += 1
        else:
            rh.append("-"); rv.append(alphabet[v[i]]); i += 1
    return "".join(rh), "".join(rv)


SC_2_1_1_3 = po.Scoring.simple(2, -1, -1, -3, alpha=5)   # Score<int,Simple>(2,-1,-1,-3)


# tests/align/test_align_global_alignment_score.h:133-175
def test_kat_global_alignment_score():
    h, v = enc("ATGT"), enc("ATAGAT")
    assert po.score(h, v, SC_2_1_1_3, 0) == 2
    assert po.score(h, v, SC_2_1_1_3, 15) == 4


# tests/align/test_alignment_algorithms_global.h:420-498 (scores and rows)
@pytest.mark.parametrize("sh,sv,score,rh,rv", [
    ("ATGT", "ATAGAT", 2, "AT-G-T", "ATAGAT"),
    ("AAAAAGGGGTTTT", "AAAGTT", 1, "AAAAAGGGGTTTT", "AAA-----GTT--"),
    ("AAAAAATTTTTGGG", "TTTTTTTTGGGGGGGG", 1, "AAAAAATTTTTGGG-----", "---TTTTTTTTGGGGGGGG"),
])
def test_kat_global_affine_rows(sh, sv, score, rh, rv):
    h, v = enc(sh), enc(sv)
    assert po.score(h, v, SC_2_1_1_3, 0) == score
    s, counts, ops = po.traceback(h, v, SC_2_1_1_3, 0)
    assert s == score
    assert rows_from_ops(h, v, ops) == (rh, rv)


# tests/align/test_alignment_algorithms_global.h:2007-2083 and :2318-2384
# (AlignConfig<true,false,false,true>)
@pytest.mark.parametrize("sh,sv,score", [
    ("ATGT", "ATAGAT", 2), ("AAAAAGGGGTTTT", "AAAGTT", 7), ("AAAAAATTTTTGGG", "TTTTTTTTGGGGGGGG", 6)])
def test_kat_semiglobal_1001(sh, sv, score):
    assert po.score(enc(sh), enc(sv), SC_2_1_1_3, po.FREE_TOP | po.FREE_BOTTOM) == score


# tests/align/test_alignment_algorithms_local.h:250-300
def test_kat_local_affine():
    assert po.score(enc("CACACTTAACTTCACAA"), enc("GGGGCTTGAGAGCTTGGGG"), SC_2_1_1_3, po.LOCAL) == 8


# tests/align/test_alignment_algorithms_local.h:250-299: the clipped rows of
# localAlignment on Align and on Gaps objects.
def test_kat_local_affine_rows():
    h, v = enc("CACACTTAACTTCACAA"), enc("GGGGCTTGAGAGCTTGGGG")
    s, c, ops, (bh, bv) = po.traceback(h, v, SC_2_1_1_3, po.LOCAL, with_begins=True)
    assert s == 8
    assert rows_from_ops(h[bh:], v[bv:], ops) == ("CTT---AACTT", "CTTGAGAGCTT")


def _pair_align():
    with open(os.path.join(GOLDEN, "pair_align.json")) as f:
        return json.load(f)["pairs"]


def _blosum():
    return po.Scoring(_blosum62_fixture(), -11, -1)


def _blosum62_fixture():
    lines = [l for l in open(os.path.join(GOLDEN, "BLOSUM62.txt")) if l.strip() and not l.startswith("#")]
    cols = lines[0].split()
    t = np.zeros((27, 27), np.int32)
    for line in lines[1:]:
        parts = line.split()
        a = AA.index(parts[0])
        for c, val in zip(cols, parts[1:]):
            t[a, AA.index(c)] = int(val)
    return t


# apps/pair_align/tests/*.stdout: -c <TOP><LEFT><RIGHT><BOTTOM>, -m gotoh / sw,
# -g / -e overrides; Blosum62 with gop -11, gex -1 (pair_align.cpp:59-67).
@pytest.mark.parametrize("name", ["1aab", "1ad2", "2trx"])
def test_kat_pair_align_protein(name):
    p = _pair_align()[name]
    h, v = enc(p["h"], AA), enc(p["v"], AA)
    t = _blosum62_fixture()
    sc = po.Scoring(t, -11, -1)
    for tag, want in p["scores"].items():
        if tag in ("default", "protein", "mgotoh", "lo5", "ms10", "mm-8"):
            assert po.score(h, v, sc, 0) == want, tag
        elif tag == "msw":
            assert po.score(h, v, sc, po.LOCAL) == want, tag
        elif tag.startswith("c") and len(tag) == 5:
            bits = [c == "t" for c in tag[1:]]
            mode = (po.FREE_TOP if bits[0] else 0) | (po.FREE_LEFT if bits[1] else 0) | (
                po.FREE_RIGHT if bits[2] else 0) | (po.FREE_BOTTOM if bits[3] else 0)
            assert po.score(h, v, sc, mode) == want, tag
        elif tag == "g-20":
            assert po.score(h, v, po.Scoring(t, -20, -1), 0) == want, tag
        elif tag == "e-5":
            assert po.score(h, v, po.Scoring(t, -11, -5), 0) == want, tag
        elif tag == "mnw":   # linear gaps: every gap character costs gapExtend
            assert po.score(h, v, po.Scoring(t, -1, -1), 0) == want, tag


@pytest.mark.parametrize("name", ["adeno1", "adeno2", "adeno3"])
def test_kat_pair_align_dna(name):
    # Dna5, Score<int>(5, -4, -1, -11) (pair_align.cpp:59-67)
    p = _pair_align()[name]
    h, v = enc(p["h"]), enc(p["v"])
    assert po.score(h, v, po.Scoring.simple(5, -4, -1, -11, alpha=5), 0) == p["scores"]["default"]


def test_pair_align_rows_give_counts():
    """The gotoh rows of pair_align, scored by per-row gap runs, equal the score;
    our traceback walks to the same rows."""
    t = _blosum62_fixture()
    sc = po.Scoring(t, -11, -1)
    for name in ["1aab", "1ad2", "2trx"]:
        p = _pair_align()[name]
        rh, rv = p["rows"]["mgotoh"]
        h, v = enc(p["h"], AA), enc(p["v"], AA)
        s, counts, ops = po.traceback(h, v, sc, 0)
        assert rows_from_ops(h, v, ops, AA) == (rh, rv)
        assert s == p["scores"]["mgotoh"]
        assert int((counts[:729] * t.ravel()).sum() + counts[729] * -11 + counts[730] * -1) == s


def test_blosum62_fixture_matches_library_preset():
    import seqan_amd as sa
    assert np.array_equal(sa.blosum62().table, _blosum62_fixture())


def _load(name):
    return np.load(os.path.join(GOLDEN, name))


def test_oracle_matches_reference_dna150():
    z = _load("ref_dna150_global.npz")
    sc = po.Scoring.simple(2, -3, -1, -5)
    got = po.score_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, 0)
    assert np.array_equal(got, z["scores"])


def test_oracle_matches_reference_lognormal_semiglobal():
    z = _load("ref_dna_lognormal_semiglobal.npz")
    sc = po.Scoring.simple(2, -3, -1, -5)
    got = po.score_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, 9)
    assert np.array_equal(got, z["scores"])


def test_oracle_matches_reference_protein_local():
    z = _load("ref_protein_local.npz")
    sc = po.Scoring(_blosum62_fixture(), -11, -1)
    got = po.score_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, 16)
    assert np.array_equal(got, z["scores"])


@pytest.mark.parametrize("fixture,make", [
    ("ref_dna_all_modes.npz", lambda z: po.Scoring.simple(int(z["match"]), int(z["mismatch"]),
                                                          int(z["gap_extend"]), int(z["gap_open"]))),
    ("ref_protein_all_modes.npz", lambda z: po.Scoring(_blosum62_fixture(), int(z["gap_open"]),
                                                       int(z["gap_extend"]))),
])
def test_oracle_matches_reference_all_modes(fixture, make):
    z = _load(fixture)
    sc = make(z)
    for k, mode in enumerate(z["modes"]):
        got = po.score_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, int(mode))
        assert np.array_equal(got, z["scores"][k]), int(mode)


@pytest.mark.parametrize("fixture", ["ref_dna1k_counts.npz", "ref_dna_short_counts.npz"])
def test_oracle_traceback_counts_match_reference(fixture):
    z = _load(fixture)
    if "match" in z:
        sc = po.Scoring.simple(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    else:
        sc = po.Scoring.simple(2, -3, -1, -5)
    s, c = po.traceback_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, 0)
    assert np.array_equal(s, z["scores"])
    assert np.array_equal(c, z["counts"])
    # sum(count * parameter) == score for every pair
    par = np.concatenate([sc.table.ravel(), [sc.gap_open, sc.gap_extend]])
    assert np.array_equal(c.astype(np.int64) @ par, s.astype(np.int64))


@pytest.mark.parametrize("fixture", ["ref_dna1k_counts.npz", "ref_dna_short_counts.npz", "ref_protein_counts.npz"])
def test_oracle_rows_match_reference(fixture):
    # The rows the reference's globalAlignment writes into its Gaps objects,
    # one byte per column (M, V: gap in the H row, H: gap in the V row).
    z = _load(fixture)
    if "match" in z:
        sc = po.Scoring.simple(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    elif "ref_protein" in fixture:
        sc = po.Scoring(_blosum62_fixture(), int(z["gap_open"]), int(z["gap_extend"]))
    else:
        sc = po.Scoring.simple(2, -3, -1, -5)
    s, c, op_off, ops = po.rows_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, 0)
    assert np.array_equal(s, z["scores"])
    assert np.array_equal(c, z["counts"])
    assert np.array_equal(op_off, z["op_off"])
    assert np.array_equal(ops, z["ops"])


@pytest.mark.parametrize("fixture", ["ref_dna_modes_rows.npz", "ref_dna_modes_short_rows.npz"])
def test_oracle_rows_match_reference_every_align_config(fixture):
    # globalAlignment(..., AlignConfig<T, L, R, B>) for all 16 configs: free
    # end gaps move the end cell and add leading / trailing gaps to the rows.
    z = _load(fixture)
    sc = po.Scoring.simple(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    for mode in range(16):
        s, c, op_off, ops = po.rows_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, mode)
        assert np.array_equal(s, z["scores"][mode]), mode
        assert np.array_equal(c, z["counts"][mode]), mode
        assert np.array_equal(op_off, z["op_off"][mode]), mode
        assert np.array_equal(ops, z["ops"][z["mode_off"][mode]:z["mode_off"][mode + 1]]), mode


def _local_scoring(z, fixture):
    if "match" in z:
        return po.Scoring.simple(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    return po.Scoring(_blosum62_fixture(), int(z["gap_open"]), int(z["gap_extend"]))


@pytest.mark.parametrize("fixture", ["ref_dna_local_rows.npz", "ref_dna_local_short_rows.npz",
                                     "ref_protein_local_rows.npz"])
def test_oracle_local_rows_match_reference(fixture):
    # localAlignment(ExecutionPolicy<Parallel, Vectorial>, gapsH, gapsV,
    # score): the clipped rows of each local window and where it begins.
    z = _load(fixture)
    s, c, op_off, ops, begins = po.rows_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"],
                                              _local_scoring(z, fixture), po.LOCAL, with_begins=True)
    assert np.array_equal(s, z["scores"])
    assert np.array_equal(c, z["counts"])
    assert np.array_equal(op_off, z["op_off"])
    assert np.array_equal(ops, z["ops"])
    assert np.array_equal(begins, z["begins"])


@pytest.mark.parametrize("fixture", ["ref_dna_linear_rows.npz", "ref_dna_linear2_rows.npz",
                                     "ref_protein_linear_rows.npz"])
def test_oracle_linear_gap_rows_match_reference(fixture):
    # gap open == gap extend: the reference's linear-gap algorithm and its
    # tie-breaks (one step per cell, no end-cell correction), global under
    # AlignConfigs and local.
    z = _load(fixture)
    sc = _local_scoring(z, fixture)
    for k, mode in enumerate(z["modes"]):
        s, c, op_off, ops, begins = po.rows_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, int(mode),
                                                  with_begins=True)
        assert np.array_equal(s, z["scores"][k]), mode
        assert np.array_equal(c, z["counts"][k]), mode
        assert np.array_equal(op_off, z["op_off"][k]), mode
        assert np.array_equal(ops, z["ops"][z["mode_off"][k]:z["mode_off"][k + 1]]), mode
        if mode == 16:
            assert np.array_equal(begins, z["begins"])


def _window_score(h, v, ops, bh, bv, sc):
    """Value of the local window ops starting at (bh, bv), gap runs per row."""
    t = sc.table
    total, i, j, prev = 0, bv, bh, ""
    for op in ops:
        if op == "M":
            total += int(t[h[j], v[i]]); i += 1; j += 1
        else:
            total += sc.gap_extend if prev == op else sc.gap_open
            if op == "V":
                i += 1
            else:
                j += 1
        prev = op if op != "M" else ""
    return total, i, j


def test_oracle_local_traceback_vs_brute_force():
    # The window is inside both sequences, its value is the best local score
    # (exhaustive search), it is empty exactly when that score is 0, and the
    # counts read from it give the score back.
    rng = np.random.default_rng(77)
    for trial in range(400):
        alpha = int(rng.integers(2, 5))
        o = int(rng.integers(-6, 0))
        e = int(rng.integers(o, 1))
        t = rng.integers(-4, 5, size=(alpha, alpha)).astype(np.int32)
        sc = po.Scoring(t, o, e)
        n, m = int(rng.integers(1, 7)), int(rng.integers(1, 7))
        h = rng.integers(0, alpha, n).astype(np.uint8)
        v = rng.integers(0, alpha, m).astype(np.uint8)
        s, c, ops, (bh, bv) = po.traceback(h, v, sc, po.LOCAL, with_begins=True)
        assert s == po.brute_force(h, v, sc, po.LOCAL), (h, v, t, o, e)
        val, iv, jh = _window_score(h, v, ops, bh, bv, sc)
        assert val == s and jh <= n and iv <= m
        assert (ops == "") == (s == 0) and (ops != "" or (bh, bv) == (0, 0))
        par = np.concatenate([t.reshape(-1), [o, e]]).astype(np.int64)
        assert int(c.astype(np.int64) @ par) == s


BANDED = ["ref_dna_banded_150.npz", "ref_dna_banded_300x280.npz", "ref_dna_banded_linear.npz",
          "ref_protein_banded.npz", "ref_dna_banded_varying.npz"]


@pytest.mark.parametrize("fixture", BANDED)
def test_oracle_banded_matches_reference(fixture):
    # Banded globalAlignmentScore: the reference's set path on collections of
    # one H length and one V length, its per-pair function on varying ones.
    z = _load(fixture)
    sc = _local_scoring(z, fixture)
    hs = [z["h_res"][z["h_off"][k]:z["h_off"][k + 1]] for k in range(len(z["h_off"]) - 1)]
    vs = [z["v_res"][z["v_off"][k]:z["v_off"][k + 1]] for k in range(len(z["v_off"]) - 1)]
    for (mode, lo, up), want in zip(z["cases"], z["scores"]):
        got = np.array([po.score_banded(h, v, sc, int(mode), int(lo), int(up)) for h, v in zip(hs, vs)], np.int32)
        assert np.array_equal(got, want), (fixture, mode, lo, up)


def test_oracle_banded_full_band_is_unbanded():
    # A band holding every diagonal changes nothing.
    rng = np.random.default_rng(31)
    sc = po.Scoring.simple(2, -3, -1, -5)
    for _ in range(200):
        h = rng.integers(0, 4, int(rng.integers(1, 40))).astype(np.uint8)
        v = rng.integers(0, 4, int(rng.integers(1, 40))).astype(np.uint8)
        mode = int(rng.integers(0, 16))
        assert po.score_banded(h, v, sc, mode, -len(v) - 3, len(h) + 2) == po.score(h, v, sc, mode)


BANDED_ROWS = ["ref_dna_banded_rows.npz", "ref_dna_banded_ties_rows.npz", "ref_dna_banded_linear_rows.npz",
               "ref_protein_banded_rows.npz"]


@pytest.mark.parametrize("fixture", BANDED_ROWS)
def test_oracle_banded_rows_match_reference(fixture):
    # Banded globalAlignment / localAlignment on Gaps, pair by pair: scores,
    # counts, rows and local begins, under every (mode, band) case of the
    # fixture.  The ties fixture (short pairs, match 1) holds end cells on the
    # band's lower diagonal in the last row, where the reference does not
    # prefer the vertical gap; the oracle without that rule fails it.
    z = _load(fixture)
    sc = _local_scoring(z, fixture)
    for k, (mode, lo, up) in enumerate(z["cases"]):
        s, c, off, ops, beg = po.rows_batch(z["h_res"], z["h_off"], z["v_res"], z["v_off"], sc, int(mode),
                                            with_begins=True, band=(int(lo), int(up)))
        assert np.array_equal(s, z["scores"][k]), (fixture, mode, lo, up)
        assert np.array_equal(c, z["counts"][k]), (fixture, mode, lo, up)
        assert np.array_equal(off, z["op_off"][k]), (fixture, mode, lo, up)
        assert np.array_equal(ops, z["ops"][z["case_off"][k]:z["case_off"][k + 1]]), (fixture, mode, lo, up)
        if mode == 16:
            assert np.array_equal(beg, z["begins"][k]), (fixture, lo, up)


def test_oracle_banded_end_cell_rows_match_reference():
    # Pairs whose reference end cell lies on the band's lower diagonal in the
    # last row (free last row), each with its own scoring: the reference takes
    # that cell's own direction there; an oracle that prefers the vertical gap
    # at the end cell, as it does elsewhere, gets 48 of these 360 rows wrong.
    z = _load("ref_dna_banded_end_rows.npz")
    for k, (mode, lo, up, ma, mi, ge, go) in enumerate(z["cases"]):
        h = z["h_res"][z["h_off"][k]:z["h_off"][k + 1]]
        v = z["v_res"][z["v_off"][k]:z["v_off"][k + 1]]
        sc = po.Scoring.simple(int(ma), int(mi), int(ge), int(go))
        s, c, off, ops = po.rows_batch(h, np.array([0, len(h)]), v, np.array([0, len(v)]), sc, int(mode),
                                       band=(int(lo), int(up)))
        assert s[0] == z["scores"][k] and np.array_equal(c[0], z["counts"][k]), k
        assert np.array_equal(ops, z["ops"][z["op_off"][k]:z["op_off"][k + 1]]), k


def test_oracle_banded_rows_full_band_is_unbanded():
    # A band holding every diagonal gives the unbanded rows, in every mode.
    rng = np.random.default_rng(32)
    sc = po.Scoring.simple(1, -1, -1, -2)
    for _ in range(150):
        h = rng.integers(0, 4, int(rng.integers(2, 30))).astype(np.uint8)
        v = rng.integers(0, 4, int(rng.integers(2, 30))).astype(np.uint8)
        ho, vo = np.array([0, len(h)]), np.array([0, len(v)])
        for mode in (0, 8, 9, 15, 16):
            full = po.rows_batch(h, ho, v, vo, sc, mode, with_begins=True)
            band = po.rows_batch(h, ho, v, vo, sc, mode, with_begins=True, band=(-len(v) - 2, len(h) + 2))
            assert all(np.array_equal(a, b) for a, b in zip(full, band)), mode


def test_100kbp_fixture_is_a_valid_alignment():
    # Two 100 kbp pairs: reference scores, oracle counts (the reference's own
    # traceback truncates trace-matrix positions past 2^32 at this size).
    # The counts form an alignment of both sequences whose value is the score.
    z = _load("ref_dna100k_counts.npz")
    c = z["counts"].astype(np.int64)
    par = np.array([2 if a == b else -3 for a in range(4) for b in range(4)] + [-5, -1])
    assert np.array_equal(c @ par, z["scores"].astype(np.int64))
    n, m = np.diff(z["h_off"]), np.diff(z["v_off"])
    assert np.array_equal(2 * c[:, :16].sum(axis=1) + c[:, 16] + c[:, 17], n + m)
    assert (n >= 99000).all() and (m >= 99000).all()


@pytest.mark.parametrize("mode", list(range(16)) + [16])
def test_oracle_vs_brute_force(mode):
    rng = np.random.default_rng(mode)
    for trial in range(60):
        alpha = int(rng.integers(2, 5))
        o = int(rng.integers(-6, 0)); e = int(rng.integers(o, 1)) if rng.random() < 0.7 else int(rng.integers(-6, 1))
        t = rng.integers(-4, 5, size=(alpha, alpha)).astype(np.int32)
        sc = po.Scoring(t, o, e)
        n, m = int(rng.integers(1, 7)), int(rng.integers(1, 7))
        h = rng.integers(0, alpha, n).astype(np.uint8)
        v = rng.integers(0, alpha, m).astype(np.uint8)
        want = po.brute_force(h, v, sc, mode)
        assert po.score(h, v, sc, mode) == want, (mode, h, v, t, o, e)


def test_oracle_closed_forms():
    rng = np.random.default_rng(7)
    sc = po.Scoring.simple(2, -3, -1, -5)
    for L in (1, 5, 150, 1000):
        h = rng.integers(0, 4, L).astype(np.uint8)
        assert po.score(h, h, sc, 0) == 2 * L                       # identical: all matches
        assert po.score(h, h[:0], sc, 0) == np.iinfo(np.int32).min  # empty -> INT_MIN
    # a single gap run of k costs open + (k-1)*extend
    h = np.zeros(10, np.uint8)
    for k in range(1, 6):
        assert po.score(h, h[:10 - k], sc, 0) == 2 * (10 - k) - 5 - (k - 1)
    # local >= 0, symmetric under swapping H and V with a symmetric table
    t = _blosum62_fixture()
    bs = po.Scoring(t, -11, -1)
    for _ in range(20):
        a = rng.integers(0, 24, int(rng.integers(1, 80))).astype(np.uint8)
        b = rng.integers(0, 24, int(rng.integers(1, 80))).astype(np.uint8)
        s1, s2 = po.score(a, b, bs, po.LOCAL), po.score(b, a, bs, po.LOCAL)
        assert s1 == s2 >= 0
        assert po.score(a, b, bs, 15) >= po.score(a, b, bs, 9) >= po.score(a, b, bs, 0)
