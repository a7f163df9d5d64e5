"""The config-sized oracle fixtures (tests/golden/config_*.npz) are this oracle's output (no GPU).

* every fixture's stored per-restart scores are the oracle's scoring restatement of the stored
  final poses, and its best score / restart their maximum / argmax;
* the cheap fixtures (C1, C3, the vol_color default call) are re-run end to end and must come out
  bit for bit: same final poses, scores and step counts;
* the C1 and vol_color inputs are regenerated from the seeded generators they were drawn from.
"""
import os
import sys

import numpy as np
import pytest
import torch

from fast_shepherd_score_amd import synth
from oracle import fss_oracle as orc
from tests._cases import load_config as load

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from make_config_vectors import c1, oracle_pair, vol_color_default  # noqa: E402

NAMES = ["c1", "c2", "c2_es", "c3", "vol_color_default"]


def _pair(d, i):
    return oracle_pair(d, i)


@pytest.mark.parametrize("name", NAMES)
def test_stored_scores_are_the_oracle_scores_of_the_stored_poses(name):
    d = load(name)
    n, B = d["q0"].shape[:2]
    assert d["scores"].shape == (n, B) and d["se3"].shape == (n, B, 7)
    for i in range(n):
        s = _pair(d, i).score_poses(torch.as_tensor(d["se3"][i], dtype=torch.float64)).numpy()
        np.testing.assert_allclose(s, d["scores"][i], rtol=0, atol=1e-12)
        assert d["best"][i] == d["scores"][i].max() and d["best_idx"][i] == int(np.argmax(d["scores"][i]))
        if not bool(d["early_stop"]):
            assert d["n_steps"][i] == int(d["steps"])


@pytest.mark.parametrize("name,pairs", [("c1", [0, 4]), ("c3", [0, 7]), ("vol_color_default", range(6))])
def test_cheap_fixtures_rerun_bit_for_bit(name, pairs):
    d = load(name)
    for i in pairs:
        o = orc.optimize(_pair(d, i), d["q0"][i], d["t0"][i], lr=float(d["lr"]), max_steps=int(d["steps"]),
                         early_stop=bool(d["early_stop"]))
        np.testing.assert_array_equal(o["se3"], d["se3"][i])
        np.testing.assert_array_equal(o["scores"], d["scores"][i])
        assert o["n_steps"] == d["n_steps"][i]


def test_inputs_come_from_the_seeded_generators():
    for name, gen in (("c1", c1), ("vol_color_default", vol_color_default)):
        d, g = load(name), gen()
        assert d["idx"].tolist() == list(g["idx"])
        for side in ("ref", "fit"):
            for a, b in zip(d[side], g[side]):
                np.testing.assert_array_equal(a, np.asarray(b, np.float32))
        np.testing.assert_array_equal(d["q0"], np.asarray(g["q0"], np.float32))
        np.testing.assert_array_equal(d["t0"], np.asarray(g["t0"], np.float32))
    # C2 / C3 sizes span the configured ranges
    c2, c3 = load("c2"), load("c3")
    assert c2["ref_n"].min() >= 200 and max(c2["ref_n"].max(), c2["fit_n"].max()) <= 400
    assert int(c2["restarts"]) == 50 and int(c2["steps"]) == 200
    sizes = np.concatenate([c3["ref_n"], c3["fit_n"]])
    assert sizes.min() <= 6 and sizes.max() >= 59
    assert synth is not None
