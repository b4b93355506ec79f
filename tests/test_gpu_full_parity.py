"""Bit-exactness at full size on every benchmark workload.

The fixtures (tests/golden/full_*.npz, written by tests/golden/make_full.py
in a container holding the reference tree) hold the outputs of EVERY pair of
each workload: the reference library's own scores (SeqAn 2
globalAlignmentScore / localAlignmentScore under
ExecutionPolicy<Parallel,Vectorial>, parallel_align_interface.h:332-365) and,
for the backward, the counts of the reference's globalAlignment rows
(computeAlignmentStats, evaluate_alignment.h:215-300), or of the oracle's
traceback for the 100 kbp pairs, which the reference's traceback cannot count
(dp_traceback_impl.h:559, 607 truncate trace positions past 2^32).  The
inputs are regenerated here by the library's seeded generator; a SHA-256 of
them, stored with the fixture, must match first."""
import hashlib
import os

import numpy as np
import pytest

import seqan_amd as sa

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIMPLE = sa.simple_score(2, -3, -1, -5)


def _digest(H, V):
    h = hashlib.sha256()
    for a in (H.residues, np.ascontiguousarray(H.offsets, np.int64), V.residues,
              np.ascontiguousarray(V.offsets, np.int64)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _load(name, config, seed, count):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not written (tests/golden/make_full.py)")
    z = np.load(path, allow_pickle=False)
    assert int(z["config"]) == config and int(z["seed"]) == seed and int(z["pairs"]) == count
    print(f"generating {count} pairs of workload {config}", flush=True)
    H, V = sa.workload(config, seed, 0, count)
    assert _digest(H, V) == str(z["sha256"]), "the generator's output changed since the fixture was written"
    return z, H, V


def _device_scores(ctx, H, V, score, mode):
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    n = len(H.offsets) - 1
    ds = sa.DeviceBuffer(4 * n)
    try:
        ctx.align_score(H, V, score, mode, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
        return ds.to_array(np.int32, n)
    finally:
        for b in (dh, dv, ds):
            b.free()


def test_all_uneven_dna_semiglobal_scores(ctx):
    z, H, V = _load("full_dna_lognormal_semiglobal.npz", sa.WL_DNA_LOGNORMAL, 0x5EED0002, 1 << 20)
    got = _device_scores(ctx, H, V, SIMPLE, int(z["mode"]))
    bad = np.nonzero(got != z["scores"])[0]
    assert len(bad) == 0, f"{len(bad)} of {len(got)} scores differ (first {bad[:5]})"
    # The same pairs on the one-pair-per-lane and group shapes.
    for G, R in ((1, 16), (16, 16)):
        print("shape", G, R, flush=True)
        got = ctx.align_score(H, V, SIMPLE, int(z["mode"]), flags=sa.force_shape(G, R))
        assert np.array_equal(got, z["scores"]), (G, R)


def test_all_protein_local_scores(ctx):
    z, H, V = _load("full_protein_local.npz", sa.WL_PROTEIN, 0x5EED0003, 1 << 18)
    got = _device_scores(ctx, H, V, sa.blosum62(-1, -11), sa.LOCAL)
    bad = np.nonzero(got != z["scores"])[0]
    assert len(bad) == 0, f"{len(bad)} of {len(got)} scores differ (first {bad[:5]})"


def test_all_dna1k_scores_and_counts(ctx):
    z, H, V = _load("full_dna1k_counts.npz", sa.WL_DNA1K, 0x5EED0004, 1 << 16)
    s, c = ctx.align_counts(H, V, SIMPLE, 0)
    assert np.array_equal(s, z["scores"])
    bad = np.nonzero((c != z["counts"]).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} of {len(c)} count rows differ (first {bad[:5]})"


def test_all_dna100k_scores_and_counts(ctx):
    z, H, V = _load("full_dna100k_counts.npz", sa.WL_DNA100K, 0x5EED0005, 64)
    s, c = ctx.align_counts(H, V, SIMPLE, 0)
    assert np.array_equal(s, z["scores"])
    bad = np.nonzero((c != z["counts"]).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} of {len(c)} count rows differ (first {bad[:5]})"
