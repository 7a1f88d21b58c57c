"""The committed block-size table (slate_amd/tuning.json, written by
slate_amd/autotune.py from timings on an MI355X) is installed into every new
context and is what the factorizations use when Options::block_size is 0."""
import json
import os

import pytest
import torch

import slate_amd as S
from slate_amd import autotune
from helpers import dev_rands, say

pytestmark = pytest.mark.gpu


def test_committed_table_is_installed_and_used(gpu):
    with open(autotune.DEFAULT_PATH) as f:
        table = json.load(f)["table"]
    S.context(0)                                  # loads the table
    for kind, rows in table.items():
        k = autotune.KINDS[kind]
        lo = 1
        for n_upper, nb in sorted(rows):
            for n in (lo, min(n_upper, 70000)):
                got = S.lib().slate_amd_tune_get_nb(S.context(0), k, n)
                say("%s n=%d: nb %d (table %d)" % (kind, n, got, nb))
                assert got == nb
            lo = n_upper + 1
    # a factorization with block_size 0 runs with the tuned nb: same pivots as
    # an explicit request for that nb
    n = 3000
    nb = S.lib().slate_amd_tune_get_nb(S.context(0), autotune.KINDS["lu_fp64"], n)
    A1, A2 = dev_rands(7, n, n), dev_rands(7, n, n)
    _, ip1 = S.getrf(A1)
    _, ip2 = S.getrf(A2, S.Options(block_size=nb))
    assert (ip1 == ip2).all() and torch.equal(A1, A2)


def test_committed_batched_configuration_is_installed(gpu):
    with open(autotune.DEFAULT_PATH) as f:
        data = json.load(f)
    got = S.tune_get_batched()
    say("batched configuration in the context: %s, committed %s" % (got, data.get("batched")))
    if "batched" in data:
        cfg = data["batched"]
        assert got[:3] == (cfg["nbf"], cfg["bt"], cfg["wgs"])
    else:
        assert got[:3] == (32, 256, 1)
    assert 1 <= got[2] <= got[3]
