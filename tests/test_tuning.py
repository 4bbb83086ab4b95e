"""The batched engine's tuning table and the autotuner's candidate grid hold
no shape that yk_batch_set_tuning refuses (CPU only)."""
import json

from yukawamd_amd import autotune
from yukawamd_amd.batch import TUNING_FILE


def test_committed_table_has_only_offered_shapes():
    table = json.loads(TUNING_FILE.read_text())["buckets"]
    assert table
    for b in table:
        assert not autotune.disabled_shape(b["list_items"], b["chunk_tiles"],
                                           b["waves_per_block"]), b


def test_candidates_exclude_refused_shapes():
    for hi in autotune.DEFAULT_BUCKETS:
        cands = autotune.candidates(hi)
        assert cands, hi
        for c in cands:
            assert not autotune.disabled_shape(c["list_items"], c["chunk_tiles"],
                                               c["waves_per_block"]), (hi, c)
            if hi > autotune.CLIST_MAX_N:
                assert c["list_items"] == 0


def test_disabled_shape_rule():
    assert autotune.disabled_shape(8, 8, 1)
    assert autotune.disabled_shape(2, 8, 4) and autotune.disabled_shape(4, 8, 4)
    for ct in (2, 16, 32, 64):
        assert autotune.disabled_shape(0, ct, 4)
    assert not autotune.disabled_shape(8, 16, 4)    # chunk_tiles unused by list items
    assert not autotune.disabled_shape(0, 128, 4)
    assert not autotune.disabled_shape(16, 8, 2)
