"""The autotuner's cache file (violet_amd.autotune.TuningCache), host side:
round trip, validation of stored shapes against the tuned dimensions,
tolerance of a damaged file, and a workload key that changes with every
part of the workload it covers."""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import pytest

from violet_amd import autotune


@dataclass
class _View:
    width: int
    height: int
    config: dict = field(default_factory=lambda: {"storage": 1})


BEST = dict(trace_tile_width=16, atrous_block=8, trace_groups=32, pixel_groups=0, atrous_groups=4)


def test_round_trip_and_table(tmp_path):
    c = autotune.TuningCache(tmp_path / "sub" / "tune.json")
    assert c.get("k") is None
    table = [(dict(BEST), 12.5, 0.25), (dict(BEST, atrous_block=16), 13.0, 0.5)]
    c.put("k", BEST, table)
    assert c.get("k") == BEST
    d = json.loads((tmp_path / "sub" / "tune.json").read_text())
    assert d["k"]["table"][1] == [dict(BEST, atrous_block=16), 13.0, 0.5]
    c.put("k2", dict(BEST, trace_tile_width=8))
    assert c.get("k") == BEST and c.get("k2")["trace_tile_width"] == 8
    assert not list((tmp_path / "sub").glob("*.tmp"))


def test_invalid_entries_and_damaged_file(tmp_path):
    p = tmp_path / "tune.json"
    c = autotune.TuningCache(p)
    p.write_text(json.dumps({"bad": {"shapes": dict(BEST, trace_tile_width=12)}, "nos": {"x": 1}, "odd": 5}))
    assert c.get("bad") is None and c.get("nos") is None and c.get("odd") is None
    p.write_text("{not json")
    assert c.get("k") is None
    c.put("k", BEST)  # a damaged file is replaced
    assert c.get("k") == BEST


def test_workload_key_covers_the_workload():
    v = [_View(640, 360), _View(1920, 1080)]
    k = autotune.workload_key(v, 1000, "gpu")
    assert k == autotune.workload_key([_View(640, 360), _View(1920, 1080)], 1000, "gpu")
    assert k != autotune.workload_key(v[::-1], 1000, "gpu")
    assert k != autotune.workload_key([_View(640, 360), _View(1920, 1088)], 1000, "gpu")
    assert k != autotune.workload_key(v, 1001, "gpu")
    assert k != autotune.workload_key(v, 1000, "other gpu")
    assert k != autotune.workload_key([_View(640, 360, {"storage": 0}), _View(1920, 1080)], 1000, "gpu")
