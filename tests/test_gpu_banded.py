"""Banded global alignment scores (sqmi_align_score_banded): the reference's
banded globalAlignmentScore (global_alignment_banded.h:394-470, :555-583).

Checked against fixtures written by the compiled reference (its set path on
collections of one H length and one V length, the case its documentation
supports, and its per-pair function on varying lengths), against the oracle
on random collections of varying lengths, and against the unbanded kernels
for bands that hold every diagonal.  Bands of up to 64 diagonals run on
kBandedNarrow (several pairs per wave), wider ones on kBandedWide."""
import os

import numpy as np
import pytest

import pyoracle as po
import seqan_amd as sa

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BANDED = ["ref_dna_banded_150.npz", "ref_dna_banded_300x280.npz", "ref_dna_banded_linear.npz",
          "ref_protein_banded.npz", "ref_dna_banded_varying.npz"]


def _score_of(z, fixture):
    if "match" in z:
        return sa.simple_score(int(z["match"]), int(z["mismatch"]), int(z["gap_extend"]), int(z["gap_open"]))
    return sa.blosum62(int(z["gap_extend"]), int(z["gap_open"]))


@pytest.mark.parametrize("fixture", BANDED)
def test_banded_matches_reference_fixtures(ctx, fixture):
    z = np.load(os.path.join(GOLDEN, fixture), allow_pickle=False)
    score = _score_of(z, fixture)
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    for (mode, lo, up), want in zip(z["cases"], z["scores"]):
        print("banded fixture", fixture, "mode", mode, "band", lo, up, flush=True)
        got = ctx.align_score_banded(Hs, Vs, score, int(lo), int(up), mode=int(mode))
        assert np.array_equal(got, want), (mode, lo, up, np.nonzero(got != want)[0][:5])


@pytest.mark.parametrize("width", [3, 5, 17, 40, 64, 65, 130, 700])
def test_banded_matches_oracle(ctx, width):
    # Varying lengths (|H| - |V| up to 6 and inside the band with a margin),
    # bands of the given width around diagonal 0; several modes; two scorings.
    rng = np.random.default_rng(width)
    lim = min(6, (width - 3) // 2)
    hs, vs = [], []
    for _ in range(300):
        n = int(rng.integers(2, 400))
        h = rng.integers(0, 4, n).astype(np.uint8)
        v = h.copy()
        f = rng.random(n) < 0.12
        v[f] = rng.integers(0, 4, int(f.sum()))
        d = int(rng.integers(-lim, lim + 1))
        if d > 0:
            v = v[:max(1, n - d)]
        elif d < 0:
            v = np.concatenate([v, rng.integers(0, 4, -d).astype(np.uint8)])
        hs.append(h)
        vs.append(v.astype(np.uint8))
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    lo = -((width - 1) // 2)
    up = lo + width - 1
    assert all(lo < min(0, len(a) - len(b)) and max(0, len(a) - len(b)) < up for a, b in zip(hs, vs))
    for score in (sa.simple_score(2, -3, -1, -5), sa.simple_score(3, -2, -3, -2)):
        sc = po.Scoring(score.table, score.gap_open, score.gap_extend)
        for mode in (0, 9, 15, 6):
            print("banded vs oracle, width", width, "mode", mode, flush=True)
            got = ctx.align_score_banded(Hs, Vs, score, lo, up, mode=mode)
            want = np.array([po.score_banded(h, v, sc, mode, lo, up) for h, v in zip(hs, vs)], np.int32)
            assert np.array_equal(got, want), (mode, np.nonzero(got != want)[0][:5])


@pytest.mark.parametrize("band", [(-5, 4), (-13, 12), (-16, 16), (-40, 30), (-3, 120)])
@pytest.mark.parametrize("alphabet", ["dna", "protein"])
def test_banded_packed_route_matches_int32_and_oracle(ctx, band, alphabet):
    """Global bands of at least 9 diagonals run on the packed int16 kernel
    (band16.hip; 16 rows per block from 25 diagonals on, else 8): its scores
    equal the int32 kernels' (SQMI_FORCE_INT32) and the oracle's, on pairs of
    many |V| (several per wavefront and one alone, |V| below, at and above a
    block's rows), |H| spread over the band, and 300 pairs of one length."""
    lo, up = band
    rng = np.random.default_rng(1000 + up - lo)
    A = 4 if alphabet == "dna" else 24
    score = sa.simple_score(2, -3, -1, -5) if alphabet == "dna" else sa.blosum62(-1, -11)
    hs, vs = [], []
    for k in range(900):
        m = int(rng.integers(1, 300)) if k < 600 else 150
        if k == 599:
            m = 1
        d = int(rng.integers(lo + 1, up))          # |H| - |V| strictly inside the band
        n = max(2, m + d)
        if not (lo < min(0, n - m) and max(0, n - m) < up):
            n = m + (1 if up > 1 and lo < 0 else 0)
            if n < 2 or not (lo < min(0, n - m) and max(0, n - m) < up):
                continue
        v = rng.integers(0, A, m).astype(np.uint8)
        h = v[:n].copy() if n <= m else np.concatenate([v, rng.integers(0, A, n - m).astype(np.uint8)])
        f = rng.random(n) < 0.15
        h[f] = rng.integers(0, A, int(f.sum()))
        hs.append(h)
        vs.append(v)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    print("packed banded", band, alphabet, len(hs), flush=True)
    got = ctx.align_score_banded(Hs, Vs, score, lo, up)
    assert ctx.stats()["pairs_int16"] == len(hs)
    ref32 = ctx.align_score_banded(Hs, Vs, score, lo, up, flags=sa.FORCE_INT32)
    assert ctx.stats()["pairs_int16"] == 0
    assert np.array_equal(got, ref32), np.nonzero(got != ref32)[0][:5]
    sc = po.Scoring(score.table, score.gap_open, score.gap_extend)
    want = np.array([po.score_banded(h, v, sc, 0, lo, up) for h, v in zip(hs, vs)], np.int32)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:5]


def test_banded_full_band_equals_unbanded(ctx):
    H, V = sa.workload(sa.WL_DNA150, 0x5EED0401, 0, 2048)
    score = sa.simple_score(2, -3, -1, -5)
    for mode in (0, 9):
        got = ctx.align_score_banded(H, V, score, -10000, 10000, mode=mode)
        assert np.array_equal(got, ctx.align_score(H, V, score, mode))


def test_banded_empty_pairs_and_errors(ctx):
    Hs = sa.SeqSet.from_list([[0, 1, 2, 3], [], [1, 1]])
    Vs = sa.SeqSet.from_list([[0, 1, 2], [1], [1, 1]])
    score = sa.simple_score(2, -3, -1, -5)
    got = ctx.align_score_banded(Hs, Vs, score, -2, 2)
    assert got[1] == np.iinfo(np.int32).min
    sc = po.Scoring(score.table, -5, -1)
    assert got[0] == po.score_banded(np.array([0, 1, 2, 3], np.uint8), np.array([0, 1, 2], np.uint8), sc, 0, -2, 2)
    assert got[2] == 4
    # One diagonal; missing diagonal 0; an edge on diagonal |H| - |V| = 1; an empty band.
    for lo, up in ((0, 0), (2, 5), (-5, 1), (3, 1)):
        with pytest.raises(sa.SqmiError):
            ctx.align_score_banded(Hs, Vs, score, lo, up)
    with pytest.raises(sa.SqmiError):   # |H| = 1
        ctx.align_score_banded(sa.SeqSet.from_list([[1]]), sa.SeqSet.from_list([[1, 2]]), score, -3, 3)


@pytest.mark.parametrize("band", [(-4, 4), (-20, 7), (3, 50), (-90, -10), (-300, 300)])
def test_banded_local_matches_oracle(ctx, band):
    # Local banded scores (localAlignmentScore with lowerDiag, upperDiag):
    # bands need not hold diagonal 0; narrow and wide kernels.
    rng = np.random.default_rng(abs(band[0]) * 1000 + abs(band[1]))
    hs, vs = [], []
    for _ in range(200):
        n = int(rng.integers(100, 400))
        h = rng.integers(0, 4, n).astype(np.uint8)
        a = int(rng.integers(0, n // 2))
        seg = h[a:a + int(rng.integers(20, 200))].copy()
        f = rng.random(len(seg)) < 0.12
        seg[f] = rng.integers(0, 4, int(f.sum()))
        v = np.concatenate([rng.integers(0, 4, int(rng.integers(0, 60))).astype(np.uint8), seg,
                            rng.integers(0, 4, int(rng.integers(0, 60))).astype(np.uint8)])
        hs.append(h)
        vs.append(v.astype(np.uint8))
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    lo, up = band
    for score in (sa.simple_score(2, -3, -1, -5), sa.simple_score(1, -1, -1, -1)):
        sc = po.Scoring(score.table, score.gap_open, score.gap_extend)
        print("local banded vs oracle, band", band, flush=True)
        got = ctx.align_score_banded(Hs, Vs, score, lo, up, mode=sa.LOCAL)
        want = np.array([po.score_banded(h, v, sc, sa.LOCAL, lo, up) for h, v in zip(hs, vs)], np.int32)
        assert np.array_equal(got, want), np.nonzero(got != want)[0][:5]


def test_banded_local_full_band_equals_unbanded(ctx):
    H, V = sa.workload(sa.WL_PROTEIN, 0x5EED0402, 0, 2048)
    score = sa.blosum62(-1, -11)
    got = ctx.align_score_banded(H, V, score, -5000, 5000, mode=sa.LOCAL)
    assert np.array_equal(got, ctx.align_score(H, V, score, sa.LOCAL))


def test_banded_local_unsupported(ctx):
    score = sa.simple_score(2, -3, -1, -5)
    Hs = sa.SeqSet.from_list([[0, 1, 2, 3, 0, 1]])
    Vs = sa.SeqSet.from_list([[0, 1, 2, 3]])
    for lo, up in ((2, 2), (5, 9), (-9, -3)):   # one diagonal; touching a corner only
        with pytest.raises(sa.SqmiError):
            ctx.align_score_banded(Hs, Vs, score, lo, up, mode=sa.LOCAL)


def test_banded_device_resident(ctx):
    # Residues and scores in device memory (SQMI_RESIDUES_ON_DEVICE,
    # SQMI_SCORES_ON_DEVICE) give the host-buffer results.
    H, V = sa.workload(sa.WL_DNA150, 0x5EED0403, 0, 3000)
    score = sa.simple_score(2, -3, -1, -5)
    want = ctx.align_score_banded(H, V, score, -12, 12)
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    ds = sa.DeviceBuffer(4 * len(H))
    try:
        ctx.align_score_banded(H, V, score, -12, 12, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
        assert np.array_equal(ds.to_array(np.int32, len(H)), want)
    finally:
        for b in (dh, dv, ds):
            b.free()


def test_banded_rows_device_resident(ctx):
    # Residues and outputs in device memory (rows packed from ops) give the
    # host-buffer results, local windows included.
    H, V = sa.workload(sa.WL_DNA150, 0x5EED0404, 0, 2000)
    score = sa.simple_score(2, -3, -1, -5)
    n, S = len(H), score.alphabet_size ** 2 + 2
    for mode in (0, sa.LOCAL):
        want = ctx.align_rows_banded(H, V, score, -12, 12, mode=mode)
        bufs = [sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues),
                sa.DeviceBuffer(4 * n), sa.DeviceBuffer(4 * S * n),
                sa.DeviceBuffer(int((H.lengths() + V.lengths()).sum())), sa.DeviceBuffer(8 * n),
                sa.DeviceBuffer(16 * n)]
        try:
            out = tuple(b.address for b in bufs[2:(7 if mode else 6)])
            assert ctx.align_rows_banded(H, V, score, -12, 12, mode=mode, h_dev=bufs[0].address,
                                         v_dev=bufs[1].address, out_dev=out) is None
            lens = bufs[5].to_array(np.int64, n)
            assert np.array_equal(bufs[2].to_array(np.int32, n), want[0])
            assert np.array_equal(bufs[3].to_array(np.int32, n * S).reshape(n, S), want[1])
            assert np.array_equal(lens, np.diff(want[2]))
            assert np.array_equal(bufs[4].to_array(np.uint8, int(lens.sum())), want[3])
            if mode:
                assert np.array_equal(bufs[6].to_array(np.int64, 2 * n).reshape(n, 2), want[4])
        finally:
            for b in bufs:
                b.free()


def test_python_mirror_functions(ctx):
    # The SeqAn-named functions of the Python mirror reach the same entry points.
    hs = [[0, 1, 2, 3, 0, 1, 2, 3], [3, 3, 2, 1, 0, 0, 1]]
    vs = [[0, 1, 2, 0, 0, 1, 2, 3], [3, 2, 2, 1, 0, 1]]
    score = sa.simple_score(2, -3, -1, -5)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    assert np.array_equal(sa.global_alignment_score(hs, vs, score, ctx=ctx), ctx.align_score(Hs, Vs, score, 0))
    assert np.array_equal(sa.global_alignment_score(hs, vs, score, ctx=ctx, lower_diag=-3, upper_diag=3),
                          ctx.align_score_banded(Hs, Vs, score, -3, 3))
    assert np.array_equal(sa.local_alignment_score(hs, vs, score, alphabet=sa.DNA, ctx=ctx, lower_diag=-3,
                                                   upper_diag=3),
                          ctx.align_score_banded(Hs, Vs, score, -3, 3, mode=sa.LOCAL))
    g = sa.global_alignment(hs, vs, score, align_config=(True, False, False, True), ctx=ctx)
    r = ctx.align_rows(Hs, Vs, score, mode=sa.FREE_TOP | sa.FREE_BOTTOM)
    assert all(np.array_equal(x, y) for x, y in zip(g, r))
    l = sa.local_alignment(hs, vs, score, alphabet=sa.DNA, ctx=ctx)
    lr = ctx.align_local_rows(Hs, Vs, score)
    assert all(np.array_equal(x, y) for x, y in zip(l, lr))
    gb = sa.global_alignment(hs, vs, score, ctx=ctx, lower_diag=-3, upper_diag=3)
    rb = ctx.align_rows_banded(Hs, Vs, score, -3, 3)
    assert all(np.array_equal(x, y) for x, y in zip(gb, rb))
    lb = sa.local_alignment(hs, vs, score, alphabet=sa.DNA, ctx=ctx, lower_diag=-3, upper_diag=3)
    lrb = ctx.align_rows_banded(Hs, Vs, score, -3, 3, mode=sa.LOCAL)
    assert all(np.array_equal(x, y) for x, y in zip(lb, lrb))
    with pytest.raises(ValueError):
        sa.global_alignment(hs, vs, score, ctx=ctx, lower_diag=-3)


# ---------------------------------------------------------------------------
# Banded rows (sqmi_align_rows_banded): globalAlignment / localAlignment on
# Gaps with lowerDiag, upperDiag (global_alignment_banded.h:152-181,
# local_alignment_banded.h:242-259), on the int32 traceback kernel with the
# band masked in.
# ---------------------------------------------------------------------------
BANDED_ROWS = ["ref_dna_banded_rows.npz", "ref_dna_banded_ties_rows.npz", "ref_dna_banded_linear_rows.npz",
               "ref_protein_banded_rows.npz"]


@pytest.fixture(params=["lanes", "int32"])
def band_route(request, monkeypatch):
    # Bands of at most 64 diagonals run on kBandedTrace + kBandedWalk;
    # SQMI_BANDED_TRACE_INT32=1 keeps them on the int32 traceback kernel.
    if request.param == "int32":
        monkeypatch.setenv("SQMI_BANDED_TRACE_INT32", "1")
    else:
        monkeypatch.delenv("SQMI_BANDED_TRACE_INT32", raising=False)
    return request.param


@pytest.mark.parametrize("fixture", BANDED_ROWS)
def test_banded_rows_match_reference_fixtures(ctx, fixture, band_route):
    z = np.load(os.path.join(GOLDEN, fixture), allow_pickle=False)
    score = _score_of(z, fixture)
    Hs, Vs = sa.SeqSet(z["h_res"], z["h_off"]), sa.SeqSet(z["v_res"], z["v_off"])
    for k, (mode, lo, up) in enumerate(z["cases"]):
        mode, lo, up = int(mode), int(lo), int(up)
        print("banded rows fixture", fixture, "mode", mode, "band", lo, up, flush=True)
        out = ctx.align_rows_banded(Hs, Vs, score, lo, up, mode=mode)
        s, c, off, ops = out[:4]
        assert np.array_equal(s, z["scores"][k]), (mode, lo, up)
        assert np.array_equal(c, z["counts"][k]), (mode, lo, up)
        assert np.array_equal(off, z["op_off"][k]), (mode, lo, up)
        assert np.array_equal(ops, z["ops"][z["case_off"][k]:z["case_off"][k + 1]]), (mode, lo, up)
        if mode == sa.LOCAL:
            assert np.array_equal(out[4], z["begins"][k]), (lo, up)
        s2, c2 = ctx.align_rows_banded(Hs, Vs, score, lo, up, mode=mode, rows=False)
        assert np.array_equal(s2, s) and np.array_equal(c2, c)
        assert np.array_equal(ctx.align_score_banded(Hs, Vs, score, lo, up, mode=mode), s)


def test_banded_rows_end_cell_on_lower_diagonal(ctx, band_route):
    # End cells on the band's lower diagonal in the last row, where the
    # reference takes the cell's own direction (one pair per call: each has
    # its own scoring).
    z = np.load(os.path.join(GOLDEN, "ref_dna_banded_end_rows.npz"), allow_pickle=False)
    print("banded end-cell rows", len(z["cases"]), "pairs", flush=True)
    for k, (mode, lo, up, ma, mi, ge, go) in enumerate(z["cases"]):
        h = z["h_res"][z["h_off"][k]:z["h_off"][k + 1]]
        v = z["v_res"][z["v_off"][k]:z["v_off"][k + 1]]
        score = sa.simple_score(int(ma), int(mi), int(ge), int(go))
        s, c, off, ops = ctx.align_rows_banded(sa.SeqSet.from_list([h]), sa.SeqSet.from_list([v]), score,
                                               int(lo), int(up), mode=int(mode))
        assert s[0] == z["scores"][k] and np.array_equal(c[0], z["counts"][k]), k
        assert np.array_equal(ops, z["ops"][z["op_off"][k]:z["op_off"][k + 1]]), k


@pytest.mark.parametrize("mode", [0, 8, 9, 15, 4, sa.LOCAL])
def test_banded_rows_match_oracle(ctx, mode):
    # Varying lengths up to 600, bands of several widths, two scorings.
    rng = np.random.default_rng(100 + mode)
    for ma, mi, ge, go in ((2, -3, -1, -5), (1, -1, -1, -2)):
        score = sa.simple_score(ma, mi, ge, go)
        sc = po.Scoring(score.table, go, ge)
        hs, vs = [], []
        for _ in range(150):
            n = int(rng.integers(2, 600))
            h = rng.integers(0, 4, n).astype(np.uint8)
            v = np.delete(h, np.nonzero(rng.random(n) < 0.03)[0])
            if len(v) == 0:
                v = h.copy()
            f = rng.random(len(v)) < 0.1
            v[f] = rng.integers(0, 4, int(f.sum()))
            hs.append(h)
            vs.append(v.astype(np.uint8))
        Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
        lim = int(np.max(np.abs(Hs.lengths() - Vs.lengths())))
        for lo, up in ((-lim - 2, lim + 3), (-lim - 15, lim + 1), (-lim - 1, lim + 40)):
            print("banded rows vs oracle, mode", mode, "scoring", (ma, mi, ge, go), "band", lo, up, flush=True)
            out = ctx.align_rows_banded(Hs, Vs, score, lo, up, mode=mode)
            want = po.rows_batch(Hs.residues, Hs.offsets, Vs.residues, Vs.offsets, sc, mode,
                                 with_begins=True, band=(lo, up))
            for a, b in zip(out, want):
                assert np.array_equal(a, b), (mode, lo, up)


def test_banded_rows_full_band_is_unbanded(ctx):
    rng = np.random.default_rng(5)
    hs = [rng.integers(0, 4, int(rng.integers(2, 300))).astype(np.uint8) for _ in range(100)]
    vs = [rng.integers(0, 4, int(rng.integers(2, 300))).astype(np.uint8) for _ in range(100)]
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    score = sa.simple_score(2, -3, -1, -5)
    for mode in (0, 9):
        got = ctx.align_rows_banded(Hs, Vs, score, -400, 400, mode=mode)
        want = ctx.align_rows(Hs, Vs, score, mode=mode)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), mode
    got = ctx.align_rows_banded(Hs, Vs, score, -400, 400, mode=sa.LOCAL)
    want = ctx.align_local_rows(Hs, Vs, score)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_banded_rows_errors_and_empty(ctx):
    Hs = sa.SeqSet.from_list([[0, 1, 2, 3], [], [1, 1, 1]])
    Vs = sa.SeqSet.from_list([[0, 1, 2], [1], [1, 1, 1]])
    score = sa.simple_score(2, -3, -1, -5)
    s, c, off, ops = ctx.align_rows_banded(Hs, Vs, score, -2, 3)
    assert s[1] == np.iinfo(np.int32).min and (c[1] == 0).all() and off[2] - off[1] == 0
    assert s[2] == 6 and bytes(ops[off[2]:off[3]]) == b"MMM"
    with pytest.raises(sa.SqmiError):
        ctx.align_rows_banded(Hs, Vs, score, 0, 3)        # misses the (0, 0) corner diagonal margin
    with pytest.raises(sa.SqmiError):
        ctx.align_rows_banded(Hs, Vs, score, 3, -2)       # lower > upper


@pytest.mark.parametrize("band", [(-1, 2), (-2, 0), (0, 1), (-3, 4), (-4, 3)])
def test_banded_rows_narrow_local_bands(ctx, band):
    # Local bands of 2-8 diagonals: segments of 2, 4 and 8 lanes of the
    # lane-per-diagonal trace kernel, against the oracle's banded traceback
    # (rows, counts, window begins), with many ties under the second scoring.
    lo, up = band
    rng = np.random.default_rng(300 + 10 * lo + up)
    for ma, mi, ge, go in ((2, -3, -1, -5), (1, -1, -1, -1)):
        score = sa.simple_score(ma, mi, ge, go)
        sc = po.Scoring(score.table, go, ge)
        hs, vs = [], []
        for _ in range(200):
            n = int(rng.integers(max(2, lo + 2), 260))
            h = rng.integers(0, 4, n).astype(np.uint8)
            v = h[:max(2, n + int(rng.integers(-2, 3)))].copy()
            f = rng.random(len(v)) < 0.15
            v[f] = rng.integers(0, 4, int(f.sum()))
            hs.append(h)
            vs.append(v)
        Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
        print("narrow local band", band, "scoring", (ma, mi, ge, go), flush=True)
        out = ctx.align_rows_banded(Hs, Vs, score, lo, up, mode=sa.LOCAL)
        want = po.rows_batch(Hs.residues, Hs.offsets, Vs.residues, Vs.offsets, sc, sa.LOCAL,
                             with_begins=True, band=(lo, up))
        for a, b in zip(out, want):
            assert np.array_equal(a, b), (band, (ma, mi, ge, go))


@pytest.mark.parametrize("band", [(-1, 1), (-2, 2), (-3, 3)])
def test_banded_rows_narrow_global_bands(ctx, band):
    # Global bands of 3-7 diagonals on pairs of equal length (both corner
    # diagonals inside the band): segments of 4 and 8 lanes.
    lo, up = band
    rng = np.random.default_rng(400 + up)
    score = sa.simple_score(2, -3, -1, -5)
    sc = po.Scoring(score.table, -5, -1)
    hs, vs = [], []
    for _ in range(200):
        n = int(rng.integers(2, 260))
        h = rng.integers(0, 4, n).astype(np.uint8)
        v = h.copy()
        f = rng.random(n) < 0.15
        v[f] = rng.integers(0, 4, int(f.sum()))
        hs.append(h)
        vs.append(v)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    for mode in (0, 9, 15):
        print("narrow global band", band, "mode", mode, flush=True)
        out = ctx.align_rows_banded(Hs, Vs, score, lo, up, mode=mode)
        want = po.rows_batch(Hs.residues, Hs.offsets, Vs.residues, Vs.offsets, sc, mode, band=(lo, up))
        for a, b in zip(out, want):
            assert np.array_equal(a, b), (band, mode)


def test_banded_rows_chunked_trace_budget(ctx, monkeypatch):
    """Banded rows and counts with a 1 MB trace budget: the lane-per-diagonal
    trace and walk run in chunks whose trace bytes fit it (3,000 pairs of
    ~150 bp, band [-16, 16]: 64 trace bytes per row, ~29 chunks), and give
    what one chunk gives."""
    rng = np.random.default_rng(5)
    hs, vs = [], []
    for _ in range(3000):
        n = int(rng.integers(140, 160))
        h = rng.integers(0, 4, n).astype(np.uint8)
        v = h.copy()
        f = rng.random(n) < 0.1
        v[f] = rng.integers(0, 4, int(f.sum()))
        v = v[:max(2, n - int(rng.integers(0, 6)))]
        hs.append(h)
        vs.append(v)
    Hs, Vs = sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)
    score = sa.simple_score(2, -3, -1, -5)
    print("banded rows, one chunk", flush=True)
    want = ctx.align_rows_banded(Hs, Vs, score, -16, 16)
    monkeypatch.setenv("SQMI_TRACE_BUDGET_MB", "1")
    print("banded rows, 1 MB trace budget", flush=True)
    got = ctx.align_rows_banded(Hs, Vs, score, -16, 16)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
