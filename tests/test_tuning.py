"""The batched engine's tuning table and the autotuner's candidate grid hold
only shapes that yk_batch_set_tuning accepts (CPU only)."""
import json

from yukawamd_amd import autotune
from yukawamd_amd.batch import TUNING_FILE


def accepted(b):
    """yk_batch_set_tuning's argument rules."""
    return (b["chunk_tiles"] >= 1 and 0 <= b["list_items"] <= 1024
            and b["waves_per_block"] in (1, 2, 4, 8))


def test_committed_table_has_only_accepted_shapes():
    table = json.loads(TUNING_FILE.read_text())["buckets"]
    assert table
    uppers = [b["n_upper"] for b in table]
    assert uppers == sorted(set(uppers))
    for b in table:
        assert accepted(b), b
        if b["n_upper"] > autotune.CLIST_MAX_N:
            assert b["list_items"] == 0, b


def test_candidates():
    for hi in autotune.DEFAULT_BUCKETS:
        cands = autotune.candidates(hi)
        assert cands, hi
        nt = (hi + 63) // 64
        for c in cands:
            assert accepted(c), (hi, c)
            if c["list_items"] == 0:
                assert -(-nt // c["chunk_tiles"]) <= autotune.MAX_ITEMS_PER_TILE, (hi, c)
            if hi > autotune.CLIST_MAX_N:
                assert c["list_items"] == 0
        keys = [tuple(sorted(c.items())) for c in cands]
        assert len(keys) == len(set(keys)), hi
