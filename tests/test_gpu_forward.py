"""GPU parity of the forward (score) path through the C ABI.

Bit-exact against the oracle on the same seeded inputs, against the golden
fixtures written by the compiled reference, and, at full sizes, through
size-independent properties.
"""
import os

import numpy as np
import pytest

import pyoracle as po
import seqan_amd as sa

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SIMPLE = sa.simple_score(2, -3, -1, -5)           # Score<int,Simple>(2,-3,-1,-5)


def _oracle_scoring(s: sa.Score):
    return po.Scoring(s.table, s.gap_open, s.gap_extend)


def _oracle(H, V, score, mode):
    return po.score_batch(H.residues, H.offsets, V.residues, V.offsets, _oracle_scoring(score), mode)


def _load(name):
    z = np.load(os.path.join(GOLDEN, name))
    return sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"]), z


def _say(*a):
    print(*a, flush=True)


def test_kats(ctx):
    s = sa.simple_score(2, -1, -1, -3, alphabet_size=5)
    H = sa.SeqSet.from_list(["ATGT", "AAAAAGGGGTTTT", "AAAAAATTTTTGGG", "CACACTTAACTTCACAA"], "ACGTN")
    V = sa.SeqSet.from_list(["ATAGAT", "AAAGTT", "TTTTTTTTGGGGGGGG", "GGGGCTTGAGAGCTTGGGG"], "ACGTN")
    g = ctx.align_score(H, V, s, 0)
    assert list(g[:3]) == [2, 1, 1]
    assert list(ctx.align_score(H, V, s, sa.FREE_TOP | sa.FREE_BOTTOM)[:3]) == [2, 7, 6]
    assert ctx.align_score(H, V, s, 15)[0] == 4
    assert ctx.align_score(H, V, s, sa.LOCAL)[3] == 8


def test_golden_dna150_global(ctx):
    H, V, z = _load("ref_dna150_global.npz")
    _say("golden dna150 global", len(H))
    assert np.array_equal(ctx.align_score(H, V, SIMPLE, 0), z["scores"])
    assert np.array_equal(ctx.align_score(H, V, SIMPLE, 0, flags=sa.FORCE_INT32), z["scores"])


def test_golden_lognormal_semiglobal(ctx):
    H, V, z = _load("ref_dna_lognormal_semiglobal.npz")
    _say("golden lognormal semiglobal", len(H))
    assert np.array_equal(ctx.align_score(H, V, SIMPLE, 9), z["scores"])


def test_golden_protein_local(ctx):
    H, V, z = _load("ref_protein_local.npz")
    _say("golden protein local", len(H))
    got = ctx.align_score(H, V, sa.blosum62(-1, -11), sa.LOCAL)
    assert np.array_equal(got, z["scores"])


@pytest.mark.parametrize("fixture", ["ref_dna_all_modes.npz", "ref_protein_all_modes.npz"])
def test_golden_all_modes(ctx, fixture):
    H, V, z = _load(fixture)
    if "match" in z:
        sc = sa.simple_score(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    else:
        sc = sa.blosum62(int(z["gap_extend"]), int(z["gap_open"]))
    for k, mode in enumerate(z["modes"]):
        _say("all modes", fixture, int(mode))
        got = ctx.align_score(H, V, sc, int(mode))
        assert np.array_equal(got, z["scores"][k]), int(mode)
        got32 = ctx.align_score(H, V, sc, int(mode), flags=sa.FORCE_INT32)
        assert np.array_equal(got32, z["scores"][k]), int(mode)


@pytest.mark.parametrize("mode", [0, 9, 15, 6, 4, 8, 1, 2, 16])
def test_random_vs_oracle(ctx, mode):
    rng = np.random.default_rng(100 + mode)
    for trial in range(3):
        alpha = int(rng.choice([4, 5, 24]))
        o = int(rng.integers(-12, -1)); e = int(rng.integers(-4, 1))
        if alpha == 24 and rng.random() < 0.5:
            sc = sa.blosum62(e, o)
        else:
            t = rng.integers(-5, 6, size=(alpha, alpha)).astype(np.int32)
            sc = sa.Score(t, e, o)
        n = int(rng.integers(50, 400))
        hs = [rng.integers(0, alpha, int(rng.integers(1, 700))).astype(np.uint8) for _ in range(n)]
        vs = []
        for h in hs:
            if rng.random() < 0.5:
                v = h.copy()
                mut = rng.random(len(v)) < 0.15
                v[mut] = rng.integers(0, alpha, int(mut.sum()))
            else:
                v = rng.integers(0, alpha, int(rng.integers(1, 700))).astype(np.uint8)
            vs.append(v)
        H, V = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
        sct = sc if sc.alphabet_size == alpha else sa.Score(sc.table, sc.gap_extend, sc.gap_open)
        _say("random", mode, trial, alpha, o, e, n)
        want = _oracle(H, V, sct, mode)
        assert np.array_equal(ctx.align_score(H, V, sct, mode), want)


def test_empty_and_tiny(ctx):
    H = sa.SeqSet.from_list(["", "A", "ACGT", "A", ""])
    V = sa.SeqSet.from_list(["ACG", "", "T", "A", ""])
    for mode in (0, 9, 15, 16):
        got = ctx.align_score(H, V, SIMPLE, mode)
        want = _oracle(H, V, SIMPLE, mode)
        assert np.array_equal(got, want), mode
        assert got[0] == got[1] == got[4] == np.iinfo(np.int32).min


def test_length_mismatch_error(ctx):
    H = sa.SeqSet.from_list(["ACGT", "AC"])
    V = sa.SeqSet.from_list(["ACGT"])
    with pytest.raises(sa.SqmiError) as ei:
        ctx.align_score(H, V, SIMPLE, 0)
    assert ei.value.status == 2


def test_int32_route_for_long_identical(ctx):
    """Identical 20 kbp sequences score 40,000 > int16: routed to int32, exact."""
    rng = np.random.default_rng(3)
    h = rng.integers(0, 4, 20000).astype(np.uint8)
    v = h.copy()
    v[::97] = (v[::97] + 1) % 4
    H, V = sa.SeqSet.from_list([h, h]), sa.SeqSet.from_list([h, v])
    got = ctx.align_score(H, V, SIMPLE, 0)
    st = ctx.stats()
    assert st["pairs_int32"] == 2
    assert got[0] == 40000
    assert np.array_equal(got, _oracle(H, V, SIMPLE, 0))


def test_dna150_4096_pairs_vs_oracle(ctx):
    H, V = sa.workload(sa.WL_DNA150, 0x5EED0001, 0, 4096)
    got = ctx.align_score(H, V, SIMPLE, 0)
    assert np.array_equal(got, _oracle(H, V, SIMPLE, 0))


def test_lognormal_semiglobal_window_vs_oracle(ctx):
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 5000, 6000)
    _say("log-normal semi-global window", len(H), int((H.lengths() * V.lengths()).sum()))
    got = ctx.align_score(H, V, SIMPLE, 9)
    assert np.array_equal(got, _oracle(H, V, SIMPLE, 9))


def test_protein_local_window_vs_oracle(ctx):
    H, V = sa.workload(sa.WL_PROTEIN, 0x5EED0003, 1000, 1500)
    got = ctx.align_score(H, V, sa.blosum62(-1, -11), sa.LOCAL)
    assert np.array_equal(got, _oracle(H, V, sa.blosum62(-1, -11), sa.LOCAL))


@pytest.mark.parametrize("gaps", [(-1, -4), (-1, -1), (-2, -11)])
def test_table_local_lookup_forms(ctx, gaps):
    """Local scores on the table kernels with every s - gapOpen >= 0 (gap
    open -4 is Blosum62's minimum: t = 0 entries) and with some below 0
    (linear gaps -1): the two lookup forms of the packed sweep.  Pairs of 1 to
    1,100 residues (several row blocks, skewed ends, padding rows and
    columns), and identical runs of W (11 per residue) whose scores reach the
    top of the int16 window."""
    e, o = gaps
    sc = sa.blosum62(e, o)
    rng = np.random.default_rng(11 - o)
    hs, vs = [], []
    for i in range(600):
        h = rng.integers(0, 24, int(rng.integers(1, 1100))).astype(np.uint8)
        if i % 3 == 0:
            v = rng.integers(0, 24, int(rng.integers(1, 1100))).astype(np.uint8)
        else:
            v = h.copy()
            mut = rng.random(len(v)) < 0.2
            v[mut] = rng.integers(0, 24, int(mut.sum()))
        hs.append(h)
        vs.append(v)
    w = sa.AMINO_ACIDS.index("W")
    for n in (2000, 2700):
        hs.append(np.full(n, w, np.uint8))
        vs.append(np.full(n + 7, w, np.uint8))
    H, V = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    _say("table local", gaps, len(hs))
    got = ctx.align_score(H, V, sc, sa.LOCAL)
    assert got[-1] == 2700 * 11 and ctx.stats()["pairs_int16"] == len(hs)
    assert np.array_equal(got, _oracle(H, V, sc, sa.LOCAL))


def test_device_resident_inputs(ctx):
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 0, 3000)
    dh = sa.DeviceBuffer.from_array(H.residues)
    dv = sa.DeviceBuffer.from_array(V.residues)
    out = sa.DeviceBuffer(4 * len(H))
    ctx.align_score(H, V, SIMPLE, 9, h_dev=dh.address, v_dev=dv.address, scores_dev=out.address)
    host = ctx.align_score(H, V, SIMPLE, 9)
    assert np.array_equal(out.to_array(np.int32, len(H)), host)
    assert np.array_equal(host, _oracle(H, V, SIMPLE, 9))
