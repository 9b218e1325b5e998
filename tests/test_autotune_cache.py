"""Autotuner cache file -> library table (host logic only, no GPU)."""
import json

import gasal2_amd as g
from gasal2_amd import autotune


def test_bucket_of():
    assert autotune.bucket_of(150) == (0, 256)
    assert autotune.bucket_of(400) == (256, 512)
    assert autotune.bucket_of(10000) == (4096, 1 << 31)


def test_apply_and_clear(tmp_path):
    path = tmp_path / "tune.json"
    entries = [{"algo": int(g.LOCAL), "start": int(g.WITHOUT_START), "second": 0, "lq_lo": lo, "lq_hi": hi,
                "min_target_words": 32, "min_cells": 1 << 20, "max_blocks": 1024, "best_ms": 1.0}
               for lo, hi in autotune.BUCKETS[:3]]
    path.write_text(json.dumps({"device": "x", "n_cu": 256, "entries": entries}))
    L = g.lib()
    autotune.clear()
    assert L.gasal2_tune_count() == 0
    assert autotune.apply(str(path)) == 3
    assert L.gasal2_tune_count() == 3
    assert autotune.apply(str(path)) == 3      # same keys update in place
    assert L.gasal2_tune_count() == 3
    autotune.clear()
    assert L.gasal2_tune_count() == 0
    assert autotune.apply(str(tmp_path / "missing.json")) == 0
