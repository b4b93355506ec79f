"""sqmi_align_score_noalloc, the entry point that allocates no device memory
(no counterpart in the reference; for callers that cannot allocate): called
through the C ABI, its scores equal sqmi_align_score's on the same inputs,
on two calls in a row (the second smaller, so that it runs over what the
first left in the shared scratch array), for host and device-resident
buffers, every route it takes (packed group and lane shapes, int32, empty
pairs), and a request larger than the scratch array is refused before
anything runs."""
import numpy as np
import pytest

import seqan_amd as sa

pytestmark = pytest.mark.gpu
SIMPLE = sa.simple_score(2, -3, -1, -5)


def _mixed_set(rng, count, lo, hi):
    hs, vs = [], []
    for k in range(count):
        n = int(rng.integers(lo, hi + 1))
        h = rng.integers(0, 4, n).astype(np.uint8)
        v = h.copy()
        f = rng.random(n) < 0.1
        v[f] = rng.integers(0, 4, int(f.sum()))
        if k % 23 == 5:
            v = v[:0]                                  # an empty pair: INT32_MIN
        hs.append(h)
        vs.append(v)
    return sa.SeqSet.from_list(hs), sa.SeqSet.from_list(vs)


def test_noalloc_equals_align_score_twice_then_smaller(ctx):
    rng = np.random.default_rng(2024)
    first = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 0, 20000)
    second = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 50000, 3000)
    for k, (H, V) in enumerate((first, second)):
        for mode in (sa.FREE_TOP | sa.FREE_BOTTOM, 0, sa.LOCAL):
            print("noalloc call", k, "pairs", len(H.offsets) - 1, "mode", mode, flush=True)
            want = ctx.align_score(H, V, SIMPLE, mode)
            got = ctx.align_score_noalloc(H, V, SIMPLE, mode)
            assert np.array_equal(got, want), (k, mode)


def test_noalloc_routes_and_device_buffers(ctx):
    rng = np.random.default_rng(7)
    H, V = _mixed_set(rng, 900, 1, 900)
    b62 = sa.blosum62(-1, -11)
    P, Q = sa.workload(sa.WL_PROTEIN, 0x5EED0003, 0, 600)
    # The int32 kernel (one thread per pair): a scoring whose values leave
    # the int16 window on the identical 40-residue pair (1000 per match),
    # beside short pairs that stay packed.
    heavy = sa.simple_score(1000, -1000, -1, -5)
    short_h, short_v = _mixed_set(rng, 300, 1, 25)
    ident = rng.integers(0, 4, 40).astype(np.uint8)
    H3 = sa.SeqSet.from_list([short_h[i] for i in range(300)] + [ident])
    V3 = sa.SeqSet.from_list([short_v[i] for i in range(300)] + [ident])
    cases = [(H3, V3, heavy, 0, 0), (H, V, SIMPLE, 9, sa.force_shape(1, 16)),
             (H, V, SIMPLE, sa.LOCAL, sa.force_shape(16, 8)), (P, Q, b62, sa.LOCAL, 0), (P, Q, b62, 0, 0)]
    for k, (A, B, sc, mode, flags) in enumerate(cases):
        print("noalloc routes case", k, "mode", mode, "flags", flags, flush=True)
        want = ctx.align_score(A, B, sc, mode, flags=flags)
        got = ctx.align_score_noalloc(A, B, sc, mode, flags=flags)
        assert np.array_equal(got, want), k
        if k == 0:
            assert ctx.stats()["pairs_int32"] == 1 and got[-1] == 40000
    # Inputs and outputs resident on the device.
    dh, dv = sa.DeviceBuffer.from_array(H.residues), sa.DeviceBuffer.from_array(V.residues)
    ds = sa.DeviceBuffer(4 * (len(H.offsets) - 1))
    try:
        print("noalloc device buffers", flush=True)
        ctx.align_score_noalloc(H, V, SIMPLE, 9, h_dev=dh.address, v_dev=dv.address, scores_dev=ds.address)
        got = ds.to_array(np.int32, len(H.offsets) - 1)
        assert np.array_equal(got, ctx.align_score(H, V, SIMPLE, 9))
    finally:
        for b in (dh, dv, ds):
            b.free()


def test_noalloc_refuses_requests_beyond_the_scratch(ctx):
    # 80 MB of host residues do not fit the 64 MiB array: refused with
    # SQMI_ERR_OUT_OF_MEMORY before any copy or launch, after which the
    # entry point still works.
    n = 40000
    h = np.zeros(n * 1000, np.uint8)
    off = np.arange(n + 1, dtype=np.int64) * 1000
    H, V = sa.SeqSet(h, off), sa.SeqSet(h.copy(), off.copy())
    print("noalloc oversize request", flush=True)
    with pytest.raises(sa.SqmiError) as e:
        ctx.align_score_noalloc(H, V, SIMPLE, 0)
    assert e.value.status == 6
    small = sa.workload(sa.WL_DNA150, 0x5EED0001, 0, 256)
    assert np.array_equal(ctx.align_score_noalloc(*small, SIMPLE, 0), ctx.align_score(*small, SIMPLE, 0))
    # The long-pair route is refused (its checkpoints take gigabytes).
    L = sa.workload(sa.WL_DNA100K, 0x5EED0005, 0, 2)
    with pytest.raises(sa.SqmiError) as e:
        ctx.align_score_noalloc(*L, SIMPLE, 0)
    assert e.value.status == 3
