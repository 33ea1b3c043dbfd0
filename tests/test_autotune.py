"""Launch-schedule table of the contour pipeline (caustics_amd/autotune.py): key
bucketing, lookup fallbacks and the committed table's schema.  Timing itself needs the
GPU; tests/test_gpu_extended.py checks that schedules change no result."""
import json
import os

import pytest

from caustics_amd import autotune as T


def test_buckets_and_keys():
    assert [T.bucket(n) for n in (4, 256, 257, 1000, 4096, 5000)] == [256, 256, 512, 1024,
                                                                        4096, 4096]
    assert T.key(3, True, True, 300) == "nl3_ld1_vjp1_n512"


def test_lookup_exact_nearest_and_default(monkeypatch):
    tab = {"entries": {"nl2_ld1_vjp1_n256": {"schedule": {"streams": 1}},
                       "nl2_ld1_vjp1_n4096": {"schedule": {"chunk_bytes": 8 << 30,
                                                           "flat_per_cu": 16}}}}
    monkeypatch.setattr(T, "_table", tab)
    assert T.lookup(2, True, True, 200) == {**T.DEFAULT, "streams": 1}
    assert T.lookup(2, True, True, 2048)["chunk_bytes"] == 8 << 30   # nearest bucket
    assert T.lookup(2, True, True, 512)["streams"] == 1              # nearest: 256
    assert T.lookup(3, True, True, 256) == T.DEFAULT                 # nothing tuned


def test_committed_table_schema():
    if not os.path.exists(T.TABLE):
        pytest.skip("no tuned table committed")
    with open(T.TABLE) as f:
        tab = json.load(f)
    assert tab["entries"]
    for k, e in tab["entries"].items():
        nl, ld, vjp, n = k.split("_")
        assert nl[:2] == "nl" and int(n[1:]) in T.BUCKETS
        for knob, v in e["schedule"].items():
            assert v in T.CANDIDATES[knob], (k, knob, v)
        assert e["ms"] <= e["default_ms"] * 1.0001
