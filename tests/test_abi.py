"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol include/seqan_amd.h declares; validation, routing and
bucketing of the planner; the workload generator."""
import os
import re

import numpy as np
import pytest

import seqan_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "seqan_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sqmi_[a-z0-9_]+)\s*\(", txt)) - {"sqmi_COUNTS_STRIDE"})


def test_library_exports_every_declared_symbol():
    L = sa.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(sa.ABI_SYMBOLS)


def test_status_strings_and_version():
    L = sa.lib()
    assert L.sqmi_abi_version() == 7 == sa.ABI_VERSION
    assert L.sqmi_status_string(2) == b"sequence sets differ in length"


SIMPLE = sa.simple_score(2, -3, -1, -5)


def test_plan_validation_errors():
    H = sa.SeqSet.from_list(["ACGT", "AC"])
    V = sa.SeqSet.from_list(["ACGT"])
    with pytest.raises(sa.SqmiError) as e:
        sa.plan_stats(H, V, SIMPLE, 0)
    assert e.value.status == 2
    V2 = sa.SeqSet.from_list(["ACGT", "A"])
    with pytest.raises(sa.SqmiError) as e:
        sa.plan_stats(H, V2, SIMPLE, 64)          # unknown mode bit
    assert e.value.status == 1
    bad = sa.SeqSet(H.residues, np.array([0, 4, 2], np.int64))
    with pytest.raises(sa.SqmiError):
        sa.plan_stats(bad, V2, SIMPLE, 0)


def test_plan_routing_int16_int32_empty():
    rng = np.random.default_rng(1)
    long = rng.integers(0, 4, 20000).astype(np.uint8)
    H = sa.SeqSet.from_list(["", "ACGT", long, rng.integers(0, 4, 9000).astype(np.uint8)])
    V = sa.SeqSet.from_list(["AC", "ACG", long, rng.integers(0, 4, 9000).astype(np.uint8)])
    st = sa.plan_stats(H, V, SIMPLE, 0)
    assert st["pairs"] == 4
    assert st["pairs_int16"] == 2          # 4x3 and 9000x9000 stay packed
    assert st["pairs_int32"] == 1          # 20 kbp identical: score 40,000 > int16
    assert st["cells"] == 12 + 20000 * 20000 + 9000 * 9000
    forced = sa.plan_stats(H, V, SIMPLE, 0, flags=sa.FORCE_INT32)
    assert forced["pairs_int16"] == 0 and forced["pairs_int32"] == 3


def test_plan_buckets_cover_all_pairs_with_bounded_padding():
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 0, 50000)
    st = sa.plan_stats(H, V, SIMPLE, sa.FREE_TOP | sa.FREE_BOTTOM)
    assert st["pairs_int16"] == 50000
    assert st["cells"] == int((H.lengths() * V.lengths()).sum())
    assert st["work_items"] >= 50000 // 8
    assert st["cells"] <= st["padded_cells"] <= 1.5 * st["cells"]


def test_plan_one_pair_per_lane_shapes():
    # Forced G = 1 shapes put every packed pair on the one-pair-per-lane
    # kernel, in items of up to 128 pairs with rows padded to R only; pairs
    # that would need more than 255 blocks of R rows fall back to G = 16.
    H, V = sa.workload(sa.WL_DNA_LOGNORMAL, 0x5EED0002, 0, 20000)
    mode = sa.FREE_TOP | sa.FREE_BOTTOM
    for R in (16, 32):
        st = sa.plan_stats(H, V, SIMPLE, mode, flags=sa.force_shape(1, R))
        m = V.lengths()
        fits = int(((m + 1 + R - 1) // R <= 255).sum())
        assert st["pairs_int16"] == 20000 and st["pairs_lane"] == fits
        assert st["cells"] <= st["padded_cells"]
    auto = sa.plan_stats(H, V, SIMPLE, mode)
    assert auto["pairs_int16"] == 20000
    with pytest.raises(sa.SqmiError) as e:
        sa.plan_stats(H, V, SIMPLE, mode, flags=sa.force_shape(1, 8))   # not compiled
    assert e.value.status == 3


def test_shape_table_json_is_well_formed():
    # The committed table the build turns into the planner's C++ table.
    from seqan_amd import build as b
    src = b.shape_table_source()
    assert "kTabChoice[4][16][4]" in src and "kTabProvenance" in src


def test_workload_windows_are_reproducible():
    H, V = sa.workload(sa.WL_PROTEIN, 7, 0, 300)
    H2, V2 = sa.workload(sa.WL_PROTEIN, 7, 120, 10)
    for k in range(10):
        assert np.array_equal(H[120 + k], H2[k]) and np.array_equal(V[120 + k], V2[k])
    assert H.residues.max() < 27 and V.residues.max() < 27
    D, E = sa.workload(sa.WL_DNA1K, 1, 0, 100)
    assert set(np.unique(D.lengths())) == {1000}
    assert D.residues.max() < 4 and E.residues.max() < 4


def test_no_device_here_is_reported():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(sa.SqmiError) as e:
        sa.Context(0)
    assert e.value.status == 4
