"""Host logic of the autotuner: the database format (shared with the CLI's
`libthom tune w4a16`), the search-space validity rules, the regret/gain
evaluation, and -- once a GPU sweep has produced them -- the committed
tuning database and trained model: regret against the exhaustive optimum on
held-out shapes and the gain over the untuned size-based choice.
"""
import math
import os

import numpy as np
import pytest

from libthom_amd import tuner


def test_db_roundtrip(tmp_path):
    path = str(tmp_path / "w4a16.csv")
    tuner.append_db(path, (1, 512, 1024, 128), (1, 64, 16, 4), 0.0123)
    tuner.append_db(path, (1, 512, 1024, 128), (1, 64, 16, 4), 0.0100)
    tuner.append_db(path, (1, 512, 1024, 128), (2, 128, 128, 1), 0.05)
    assert open(path).readline().strip() == ",".join(tuner.DB_COLUMNS)
    db = tuner.read_db(path)
    assert db == {(1, 512, 1024, 128): {(1, 64, 16, 4): 0.0100, (2, 128, 128, 1): 0.05}}


def test_space_and_validity():
    assert len(tuner.W4A16_SPACE) == 35
    assert tuner.valid((6, 128, 128, 1), 128, 4096, 11008, 128) and not tuner.valid((6, 128, 128, 1), 128, 4096, 520, 128)
    assert not tuner.valid((6, 128, 128, 1), 128, 4096, 11008, 32) and not tuner.valid((6, 128, 128, 1), 128, 4096)
    assert tuner.valid((4, 64, 0, 1), 4, 4096, 11008, 32) and not tuner.valid((4, 64, 0, 1), 5, 4096, 11008, 128)
    assert tuner.valid((4, 32, 0, 1), 2, 4096, 4096, 64) and not tuner.valid((4, 32, 0, 1), 2, 4096, 4096, 32)
    assert not tuner.valid((4, 64, 0, 1), 4, 16384, 4096, 128)
    assert tuner.valid((5, 64, 0, 1), 64, 4096, 520, 32) and not tuner.valid((5, 64, 0, 1), 65, 4096, 520, 32)
    assert not tuner.valid((5, 64, 0, 1), 16, 4096)
    assert tuner.valid((5, 32, 0, 1), 64, 4096, 520, 32) and not tuner.valid((5, 32, 0, 1), 65, 4096, 520, 32)
    assert tuner.valid((3, 128, 64, 1), 200, 4096, 11008, 128) and not tuner.valid((3, 128, 64, 1), 200, 4096, 520, 128)
    assert tuner.valid((1, 64, 16, 8), 128, 4096) and not tuner.valid((1, 64, 16, 8), 129, 4096)
    assert tuner.valid((2, 256, 256, 1), 1, 512)
    # kind 3: N % 32 == 0 and group % 64 == 0, known only with N and group
    assert tuner.valid((3, 256, 256, 1), 2048, 4096, 11008, 128)
    assert not tuner.valid((3, 256, 256, 1), 2048, 4096, 11008, 32)
    assert not tuner.valid((3, 256, 256, 1), 2048, 4096, 520, 128)
    assert not tuner.valid((3, 256, 256, 1), 2048, 4096)
    shapes = tuner.random_shapes(500)
    for M, N, K, G in shapes:
        assert 1 <= M <= 4096 and N % 64 == 0 and 512 <= N <= 16384 and K % 128 == 0 and K % G == 0


def test_evaluate_regret_and_gain():
    db = {(1, 64, 128, 128): {(1, 64, 16, 1): 2.0, (1, 64, 16, 2): 1.0},
          (2, 64, 128, 128): {(1, 64, 16, 1): 1.0, (1, 64, 16, 2): 1.5}}
    shapes = list(db)
    r = tuner.evaluate(db, shapes, [(1, 64, 16, 2), (1, 64, 16, 2)], {shapes[0]: 3.0, shapes[1]: 3.0})
    assert r["mean_regret"] == pytest.approx(0.25) and r["within_5pct"] == 0.5
    assert r["mean_speedup_vs_default"] == pytest.approx((3.0 + 2.0) / 2)


DB = os.path.join(tuner.TUNING_DIR, "w4a16_sweep.csv")
MODEL = os.path.join(tuner.TUNING_DIR, "w4a16_mlp.json")
DEFAULT = os.path.join(tuner.TUNING_DIR, "w4a16_default.csv")


@pytest.mark.skipif(not (os.path.exists(DB) and os.path.exists(MODEL)), reason="no committed sweep/model yet")
def test_committed_model_regret_and_gain():
    db = tuner.read_db(DB)
    shapes = sorted(db)
    assert len(shapes) >= 1000
    _, test = tuner.split_shapes(shapes)
    model = tuner.CostModel.load(MODEL, device="cpu")
    chosen = model.choose(test, model.predict_logtime_host(test))
    res = tuner.evaluate(db, test, chosen, tuner.read_default(DEFAULT))
    # the learned choice lands within 5% of the exhaustive optimum on average,
    # and beats the untuned size-based choice
    assert res["mean_regret"] <= 0.05, res
    assert res["geomean_speedup_vs_default"] > 1.0, res
