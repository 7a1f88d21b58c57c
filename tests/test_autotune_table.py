"""Shape of the committed block-size table (CPU): every kind the library
tunes, ascending size bounds ending with an open-ended entry, block sizes
from the tuner's candidate list."""
import json
import os

from slate_amd import autotune


def test_committed_table_shape():
    assert os.path.exists(autotune.DEFAULT_PATH), "slate_amd/tuning.json is missing"
    with open(autotune.DEFAULT_PATH) as f:
        data = json.load(f)
    table = data["table"]
    assert set(table) == set(autotune.KINDS)
    for kind, rows in table.items():
        bounds = [r[0] for r in rows]
        assert bounds == sorted(bounds) and len(set(bounds)) == len(bounds), kind
        assert bounds[-1] >= 1 << 40, kind
        assert all(r[1] in autotune.CANDIDATES for r in rows), kind
    assert data["timings_ms"], "raw timings kept beside the table"


def test_committed_batched_entry_shape():
    with open(autotune.DEFAULT_PATH) as f:
        data = json.load(f)
    if "batched" not in data:
        return                                    # the batched entry is added by autotune --batched
    cfg = data["batched"]
    assert (cfg["nbf"], cfg["bt"]) in autotune.BATCHED_CONFIGS
    assert cfg["wgs"] in (1, 2)
    key = "%d/%d/%d" % (cfg["nbf"], cfg["bt"], cfg["wgs"])
    raw = data["batched_timings_ms"]
    assert key in raw
    assert sum(raw[key].values()) == min(sum(v.values()) for v in raw.values())
