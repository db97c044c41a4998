"""Autotuner table logic (no device): classes, nearest-class lookup, storage."""
import json

import numpy as np

from katric_amd import tune as T


def test_graph_and_batch_keys():
    off = np.array([0, 2, 4, 6, 6], np.uint64)  # degrees 2, 2, 2, 0
    assert T.graph_key(4, off) == "plain/n2/d1/D1"
    assert T.batch_key(3, off).startswith("batch/g2/")


def test_candidates_respect_bin_order():
    for stage in T.STAGES:
        for p in T.candidates(stage, T.DEFAULTS):
            assert p["group_max_deg"] >= p["wave_max_deg"]
            assert set(p) == set(T.DEFAULTS)
    assert len(T.candidates(T.STAGES[0], T.DEFAULTS)) == 24  # 6 list walks x 2 orders x 2 walk sides


def test_save_lookup_nearest(tmp_path):
    path = str(tmp_path / "t.json")
    p1 = dict(T.DEFAULTS, strategy=1, order=1)
    p2 = dict(T.DEFAULTS, wave_max_deg=64)
    T.save_result("batch/g13/d5/D12", {"params": p1, "ms": 1.0}, path)
    T.save_result("plain/n22/d5/D18", {"params": p2, "ms": 2.0}, path)
    table = T.load_table(path)
    assert T.lookup("batch/g13/d5/D12", table) == p1
    assert T.lookup("batch/g6/d4/D11", table) == p1  # nearest batch class
    assert T.lookup("plain/n16/d5/D6", table) == p2
    assert T.lookup("plain/n16/d5/D6", {}) == T.DEFAULTS
    with open(path) as f:
        assert json.load(f)["classes"]["plain/n22/d5/D18"]["ms"] == 2.0


def test_package_table_is_valid():
    for key, entry in T.load_table(T.PACKAGE_TABLE).items():
        assert key.split("/")[0] in ("plain", "batch")
        assert set(entry["params"]) == set(T.DEFAULTS)
