"""The committed selector networks against the solver-labelled systems they
were trained on (CPU, host restatement of the forward pass): the same 20%
held-out split as training, misselection rate and summed system time
against the best choice per system and against every fixed choice.

Labels were written by the solver itself on an MI355X (slate_amd/batched.py
and slate_amd/selector.py 'label'); they are this library's data, not
reference fixtures."""
import os

import numpy as np
import pytest

import slate_amd as S
from slate_amd import batched, selector

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def test_batched_selector_holdout():
    params = S.batched_params()
    assert params is not None, "slate_amd/selector_batched_params.npy is missing"
    assert params.size == S.SELECTOR_FEATURES * 64 + 64 + 64 * 64 + 64 + 64 * len(S.CHOICES) + len(S.CHOICES)
    rows = batched.load_rows(os.path.join(DATA, "batched_labels_fp32_probe.npz"))
    _, ho = selector.split(rows, 0.2)
    rep = batched.evaluate(params, ho)
    print(rep)
    assert rep["systems"] >= 800
    assert rep["misselection_rate"] <= 0.2
    assert rep["mlp_over_best"] <= 1.10
    # no worse than the best single fixed choice (FP32 + IR, SLATE's gesv_mixed form)
    assert rep["mlp_over_best_fixed"] <= 1.0


def test_main_selector_holdout():
    """The committed precision/scheme network on the 809 systems held out of
    its training (n in [512, 8192] plus a few up to 65536, kappa up to 1e14)."""
    P = np.load(selector.DEFAULT_PARAMS, allow_pickle=False)
    assert P.size == S.SELECTOR_FEATURES * 64 + 64 + 64 * 64 + 64 + 64 * len(S.CHOICES) + len(S.CHOICES)
    rows = selector.load_rows(os.path.join(DATA, "selector_holdout.npz"))
    rep = selector.evaluate(P, rows)
    print(rep)
    assert rep["systems"] >= 800
    assert rep["misselection_rate"] <= 0.2
    assert rep["mlp_over_best"] <= 1.10
    assert rep["mlp_over_best_fixed"] < 1.0          # gains over every single fixed choice
    assert rep["mlp_over_default"] < 1.0             # and over SLATE's default (FP32 + classic IR)
