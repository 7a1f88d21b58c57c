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
