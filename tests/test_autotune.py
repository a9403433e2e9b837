"""Autotuner search and cache on the CPU with synthetic cost functions."""
import json

from isce3_amd import autotune


def cost(opts):
    # separable bowl with its minimum at tile_rows 2, order 1, chunk 64, bpc 1, staged 0
    if opts["lane_rows"] == 2 and opts["staged"]:
        raise ValueError("not offered")
    return (abs(opts["tile_rows"] - 2) + 3 * (opts["order"] != 1) + abs(opts["chunk"] - 64) / 16
            + 2 * abs(opts["blocks_per_cu"] - 1) + 5 * opts["staged"] - (opts["lane_rows"] == 2)
            + 10.0)


def test_coordinate_descent_finds_the_minimum(tmp_path):
    t = autotune.Autotuner(str(tmp_path / "c.json"), repeats=1)
    calls = []
    best = t.tune("k", "split", lambda o: calls.append(o) or cost(o))
    assert best == {"tile_rows": 2, "order": 1, "chunk": 64, "blocks_per_cu": 1, "staged": 0,
                    "lane_rows": 2}
    n_space = 4 * 2 * 3 * 3 * 2 * 2
    assert len(calls) < n_space  # fewer runs than the exhaustive search
    assert len({tuple(sorted(c.items())) for c in calls}) == len(calls)  # nothing timed twice


def test_cache_roundtrip_and_force(tmp_path):
    path = str(tmp_path / "sub" / "c.json")
    t = autotune.Autotuner(path, repeats=2)
    best = t.tune("k", "plain", lambda o: abs(o["tile_rows"] - 4))
    assert best == {"tile_rows": 4}
    data = json.load(open(path))
    assert data["k"]["options"] == {"tile_rows": 4}
    t2 = autotune.Autotuner(path)
    assert t2.lookup("k") == {"tile_rows": 4}
    called = []
    assert t2.tune("k", "plain", lambda o: called.append(o) or 0.0) == {"tile_rows": 4}
    assert not called  # cached
    again = t2.tune("k", "plain", lambda o: abs(o["tile_rows"] - 8), force=True)
    assert again == {"tile_rows": 8}


def test_defaults_and_keys(tmp_path):
    t = autotune.Autotuner(str(tmp_path / "none.json"))
    assert t.options_for("missing", "plain") == {"tile_rows": autotune.DEFAULTS["tile_rows"]}
    assert t.options_for("missing", "split") == autotune.DEFAULTS
    k1 = autotune.workload_key("gfx950", "split", 8, 1 << 20, 300.0)
    k2 = autotune.workload_key("gfx950", "split", 8, (1 << 20) + 5, 400.0)
    k3 = autotune.workload_key("gfx950", "split", 8, 1 << 22, 300.0)
    assert k1 == k2 and k1 != k3


def test_corrupt_cache_is_ignored(tmp_path):
    p = tmp_path / "c.json"
    p.write_text("{not json")
    t = autotune.Autotuner(str(p))
    assert t.lookup("k") is None


def test_nearest_bucket(tmp_path):
    t = autotune.Autotuner(str(tmp_path / "c.json"), repeats=1)
    k_small = autotune.workload_key("gfx950", "split", 8, 1 << 12, 64.0)
    k_big = autotune.workload_key("gfx950", "split", 8, 1 << 24, 512.0)
    t.tune(k_small, "split", lambda o: abs(o["tile_rows"] - 8), space={"tile_rows": (1, 8)})
    t.tune(k_big, "split", lambda o: abs(o["tile_rows"] - 2), space={"tile_rows": (1, 2)})
    assert t.options_for(autotune.workload_key("gfx950", "split", 8, 1 << 13, 64.0),
                         "split") == {"tile_rows": 8}
    assert t.options_for(autotune.workload_key("gfx950", "split", 8, 1 << 22, 300.0),
                         "split") == {"tile_rows": 2}
    assert t.nearest(autotune.workload_key("gfx950", "plain", 8, 1 << 22, 300.0)) is None
