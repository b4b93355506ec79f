"""Replay of identical sqmi_align_score calls: the first call captures its
launches in a HIP graph, the next identical calls relaunch it.  Every
replay is checked against the eager result (SQMI_GRAPH=0), with eager work
of another context and host synchronisation between the replays; the
residues may change between replays (only the lengths fix the plan), and
any other call on the context drops the graph."""
import numpy as np
import pytest

import pyoracle as po
import seqan_amd as sa

pytestmark = pytest.mark.gpu
SIMPLE = sa.simple_score(2, -3, -1, -5)


def _eager(ctx, monkeypatch, *args, **kw):
    monkeypatch.setenv("SQMI_GRAPH", "0")
    try:
        return ctx.align_score(*args, **kw)
    finally:
        monkeypatch.delenv("SQMI_GRAPH")


@pytest.mark.parametrize("mode", [sa.FREE_TOP | sa.FREE_BOTTOM, 0, sa.LOCAL])
def test_replays_match_eager(ctx, monkeypatch, mode):
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 7000, 4000)
    want = _eager(ctx, monkeypatch, H, V, SIMPLE, mode)
    other = sa.Context(0)
    try:
        P, Q = sa.workload(sa.WL_PROTEIN, 0x5EED0003, 0, 500)
        for k in range(4):
            print("replay", k, "mode", mode, flush=True)
            got = ctx.align_score(H, V, SIMPLE, mode)
            assert ctx.stats()["replayed"] == (1 if k > 0 else 0)
            assert np.array_equal(got, want), k
            other.align_score(P, Q, sa.blosum62(-1, -11), sa.LOCAL)   # eager work elsewhere
            sa.synchronize()
    finally:
        other.close()


def test_replay_with_new_residues_and_device_buffers(ctx, monkeypatch):
    # Same lengths, different residues: the plan is reused, the scores are
    # the new residues' scores.  Device-resident residues and scores too.
    rng = np.random.default_rng(5)
    H, V = sa.workload(sa.WL_DNA1K, 0x5EED0004, 0, 512)
    for k in range(3):
        h2 = sa.SeqSet(rng.integers(0, 4, H.residues.size).astype(np.uint8), H.offsets)
        v2 = sa.SeqSet(rng.integers(0, 4, V.residues.size).astype(np.uint8), V.offsets)
        print("new residues", k, flush=True)
        got = ctx.align_score(h2, v2, SIMPLE, 0)
        assert ctx.stats()["replayed"] == (1 if k > 0 else 0)
        want = po.score_batch(h2.residues, h2.offsets, v2.residues, v2.offsets,
                              po.Scoring(SIMPLE.table, -5, -1), 0)
        assert np.array_equal(got, want), k
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    ds = sa.DeviceBuffer(4 * 512)
    try:
        want = _eager(ctx, monkeypatch, H, V, SIMPLE, 9)
        for k in range(3):
            ctx.align_score(H, V, SIMPLE, 9, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
            assert ctx.stats()["replayed"] == (1 if k > 0 else 0)
            assert np.array_equal(ds.to_array(np.int32, 512), want), k
    finally:
        for b in (dh, dv, ds):
            b.free()


def test_replay_dropped_by_other_calls(ctx):
    H, V = sa.workload(sa.WL_DNA150, 0x5EED0001, 0, 300)
    first = ctx.align_score(H, V, SIMPLE, 0)
    ctx.align_counts(H, V, SIMPLE, 0)                     # rewrites the context's buffers
    assert np.array_equal(ctx.align_score(H, V, SIMPLE, 0), first)
    assert ctx.stats()["replayed"] == 0
    assert np.array_equal(ctx.align_score(H, V, SIMPLE, 0), first)
    assert ctx.stats()["replayed"] == 1
    ctx.align_score(H, V, SIMPLE, 9)                      # another mode: a new plan
    assert ctx.stats()["replayed"] == 0


def test_long_pairs_graph_replays(ctx, monkeypatch):
    # The long-pair route: one launch per tile anti-diagonal on three
    # streams, captured with its fork / join, replayed.
    H, V = sa.workload(sa.WL_DNA100K, 0x5EED0005, 0, 6, 16000)
    want = _eager(ctx, monkeypatch, H, V, SIMPLE, 0)
    assert ctx.stats()["pairs_long"] == 6
    for k in range(3):
        print("long replay", k, flush=True)
        got = ctx.align_score(H, V, SIMPLE, 0)
        assert ctx.stats()["replayed"] == (1 if k > 0 else 0)
        assert np.array_equal(got, want), k
