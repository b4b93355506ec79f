"""Every compiled packed-kernel shape (lanes per pair group G x rows per lane
R) and the plain one-pair-per-launch form give the oracle's scores."""
import numpy as np
import pytest

import pyoracle as po
import seqan_amd as sa

pytestmark = pytest.mark.gpu

# (16, R): a group of 16 lanes per pair couple; (64, R): a wavefront per
# pair couple (kFwd16); (1, R): one pair couple per lane (kLane16).
SHAPES = [(16, 16), (16, 8), (64, 4), (64, 8), (64, 16), (1, 16), (1, 32)]


def _random_set(rng, count, lo, hi, alpha):
    hs, vs = [], []
    for _ in range(count):
        h = rng.integers(0, alpha, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        if rng.random() < 0.6:
            v = h.copy()
            mut = rng.random(len(v)) < 0.1
            v[mut] = rng.integers(0, alpha, int(mut.sum()))
            if len(v) > 10:
                a = int(rng.integers(0, len(v) - 5))
                v = np.concatenate([v[:a], v[a + int(rng.integers(1, 5)):]])
        else:
            v = rng.integers(0, alpha, int(rng.integers(lo, hi + 1))).astype(np.uint8)
        hs.append(h)
        vs.append(v)
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


def _oracle(H, V, s, mode):
    return po.score_batch(H.residues, H.offsets, V.residues, V.offsets,
                          po.Scoring(s.table, s.gap_open, s.gap_extend), mode)


@pytest.mark.parametrize("shape", SHAPES)
def test_forced_shape_matches_oracle(ctx, shape):
    G, R = shape
    rng = np.random.default_rng(G * 100 + R)
    dna = sa.simple_score(2, -3, -1, -5)
    H, V = _random_set(rng, 150, 1, 600, 4)
    for mode in (0, 9, 15, 4, 2):
        print("shape", shape, "dna mode", mode, flush=True)
        got = ctx.align_score(H, V, dna, mode, flags=sa.force_shape(G, R))
        assert np.array_equal(got, _oracle(H, V, dna, mode)), (shape, mode)
    P, Q = _random_set(rng, 80, 1, 500, 24)
    b62 = sa.blosum62(-1, -11)
    for mode in (16, 0, 9):
        print("shape", shape, "protein mode", mode, flush=True)
        got = ctx.align_score(P, Q, b62, mode, flags=sa.force_shape(G, R))
        assert np.array_equal(got, _oracle(P, Q, b62, mode)), (shape, mode)


def test_plain_per_pair_matches_batched(ctx):
    rng = np.random.default_rng(11)
    dna = sa.simple_score(2, -3, -1, -5)
    H, V = _random_set(rng, 12, 1, 400, 4)
    print("plain: 12 pairs", flush=True)
    got = ctx.align_score(H, V, dna, 9, flags=sa.PLAIN_PER_PAIR)
    assert np.array_equal(got, _oracle(H, V, dna, 9))
    assert np.array_equal(got, ctx.align_score(H, V, dna, 9))


@pytest.mark.parametrize("min_blocks", ["1", "2", "3", "0"])
def test_cooperative_items_match_oracle(ctx, monkeypatch, min_blocks):
    # Work items with at least SQMI_COOP_MIN_BLOCKS row blocks are swept by
    # the four waves of a workgroup together (0: never); every setting gives
    # the oracle's scores, including items with fewer blocks than waves.
    monkeypatch.setenv("SQMI_COOP_MIN_BLOCKS", min_blocks)
    rng = np.random.default_rng(int(min_blocks) + 40)
    dna = sa.simple_score(2, -3, -1, -5)
    H, V = _random_set(rng, 200, 1, 1500, 4)
    for mode in (0, 9, 15, 4, 2):
        print("coop", min_blocks, "dna mode", mode, flush=True)
        got = ctx.align_score(H, V, dna, mode)
        assert np.array_equal(got, _oracle(H, V, dna, mode)), mode
    P, Q = _random_set(rng, 60, 1, 1200, 24)
    b62 = sa.blosum62(-1, -11)
    for mode in (16, 0):
        print("coop", min_blocks, "protein mode", mode, flush=True)
        got = ctx.align_score(P, Q, b62, mode)
        assert np.array_equal(got, _oracle(P, Q, b62, mode)), mode


@pytest.mark.parametrize("R", [16, 32])
def test_lane_shape_full_items_and_tall_pairs(ctx, R):
    # One pair couple per lane with full items of 128 pairs, pairs of many
    # row blocks, and pairs taller than 255 blocks of R rows (which take
    # G = 16), under the AlignConfigs that track differently and local.
    rng = np.random.default_rng(40 + R)
    dna = sa.simple_score(2, -3, -1, -5)
    H, V = _random_set(rng, 700, 1, 300, 4)
    Ht, Vt = _random_set(rng, 40, 3000, 5200, 4)
    Hs = sa.SeqSet.from_list([H[i] for i in range(700)] + [Ht[i] for i in range(40)])
    Vs = sa.SeqSet.from_list([V[i] for i in range(700)] + [Vt[i] for i in range(40)])
    for mode in (0, 1, 9, 15, 6, sa.LOCAL):
        print("lane shape R", R, "mode", mode, flush=True)
        got = ctx.align_score(Hs, Vs, dna, mode, flags=sa.force_shape(1, R))
        assert ctx.stats()["pairs_lane"] > 0
        assert np.array_equal(got, _oracle(Hs, Vs, dna, mode)), (R, mode)
        assert np.array_equal(got, ctx.align_score(Hs, Vs, dna, mode)), (R, mode)


def test_lane_shape_device_residues_unaligned(ctx):
    # Residues resident on the device at odd byte offsets: the lane kernel
    # reads aligned dwords around each pair's residues.
    rng = np.random.default_rng(77)
    dna = sa.simple_score(2, -3, -1, -5)
    H, V = _random_set(rng, 300, 1, 700, 4)
    dh = sa.DeviceBuffer.from_array(np.concatenate([np.zeros(3, np.uint8), H.residues]))
    dv = sa.DeviceBuffer.from_array(np.concatenate([np.zeros(1, np.uint8), V.residues]))
    try:
        for mode in (0, 9, sa.LOCAL):
            print("lane shape, device residues, mode", mode, flush=True)
            got = ctx.align_score(H, V, dna, mode, flags=sa.force_shape(1, 16), h_dev=dh.address + 3,
                                  v_dev=dv.address + 1)
            assert np.array_equal(got, _oracle(H, V, dna, mode)), mode
    finally:
        dh.free()
        dv.free()


@pytest.mark.parametrize("min_blocks", ["2", "0"])
def test_wavefront_shape_cooperative_and_long_rows(ctx, monkeypatch, min_blocks):
    # A wavefront per pair couple over pairs of many row blocks (1,500-row
    # pairs are six 256-row blocks of G = 64, R = 4), swept by one wave or by
    # the four waves of a workgroup together.
    monkeypatch.setenv("SQMI_COOP_MIN_BLOCKS", min_blocks)
    rng = np.random.default_rng(int(min_blocks) + 60)
    dna = sa.simple_score(2, -3, -1, -5)
    H, V = _random_set(rng, 120, 1, 1500, 4)
    for shape in ((64, 4), (64, 8)):
        for mode in (0, 9, 15, sa.LOCAL):
            print("wavefront shape", shape, "coop", min_blocks, "mode", mode, flush=True)
            got = ctx.align_score(H, V, dna, mode, flags=sa.force_shape(*shape))
            assert np.array_equal(got, _oracle(H, V, dna, mode)), (shape, mode)


@pytest.mark.parametrize("shape", [(16, 16), (16, 8), (64, 4), (64, 8), (64, 16)])
def test_fine_row_units_match_oracle(ctx, monkeypatch, shape):
    # Every pair in quarter-block row units: block 0 of an item holds 1..4
    # quarters of the shape's rows (R / 4 .. R rows per lane), chosen by the
    # items' tallest pair; every quarter count occurs among these lengths.
    monkeypatch.setenv("SQMI_ROW_UNITS", "quarter")
    G, R = shape
    rng = np.random.default_rng(G * 7 + R)
    dna = sa.simple_score(2, -3, -1, -5)
    H, V = _random_set(rng, 300, 1, 1100, 4)
    for mode in (0, 9, 15, 4, 2, sa.LOCAL):
        print("fine units", shape, "dna mode", mode, flush=True)
        got = ctx.align_score(H, V, dna, mode, flags=sa.force_shape(G, R))
        assert np.array_equal(got, _oracle(H, V, dna, mode)), (shape, mode)
    P, Q = _random_set(rng, 80, 1, 500, 24)
    b62 = sa.blosum62(-1, -11)
    for mode in (16, 0, 9):
        print("fine units", shape, "protein mode", mode, flush=True)
        got = ctx.align_score(P, Q, b62, mode, flags=sa.force_shape(G, R))
        assert np.array_equal(got, _oracle(P, Q, b62, mode)), (shape, mode)
