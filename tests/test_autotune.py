"""Host logic of the autotuner: coordinate descent, signature rejection and
the tuned-config table."""
import os

from back40computing_amd import autotune


def test_descent_finds_separable_minimum():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30]}
    target = {"a": 3, "b": 20}

    def measure(cfg):
        ms = 1.0 + abs(cfg["a"] - target["a"]) + 0.1 * abs(cfg["b"] - target["b"])
        return ms, 0

    best, ms, trials = autotune.coordinate_descent(measure, {"a": 1, "b": 10}, space, passes=3)
    assert best == target
    assert ms == 1.0
    assert all(ok for _, _, ok in trials)


def test_descent_rejects_changed_output():
    space = {"a": [1, 2, 3]}

    def measure(cfg):
        # a == 3 is fastest but produces a different output: never chosen
        return (0.5, 99) if cfg["a"] == 3 else (float(cfg["a"]), 7)

    best, ms, trials = autotune.coordinate_descent(measure, {"a": 2}, space, passes=2)
    assert best == {"a": 1}
    assert ms == 1.0
    rejected = [c for c, _, ok in trials if not ok]
    assert rejected and all(c["a"] == 3 for c in rejected)


def test_descent_keeps_start_within_noise():
    space = {"a": [1, 2]}

    def measure(cfg):
        return (1.0 if cfg["a"] == 1 else 0.995), 0

    best, _, _ = autotune.coordinate_descent(measure, {"a": 1}, space, passes=1)
    assert best == {"a": 1}  # a 0.5 % gain is inside the 1 % noise margin


def test_table_round_trip(tmp_path):
    path = os.path.join(tmp_path, "tuned.json")
    assert autotune.tuned_config("batch", path=path, device="dev0") == {}
    autotune.save_entry("batch", {"blocks_per_cu": 4}, 12.5, "w", path=path, device="dev0")
    autotune.save_entry("single", {"tile_grab_max": 32}, 1.0, "w", path=path, device="dev0")
    assert autotune.tuned_config("batch", path=path, device="dev0") == {"blocks_per_cu": 4}
    # another device name falls back to any entry of the same kind
    assert autotune.tuned_config("single", path=path, device="other") == {"tile_grab_max": 32}


def test_search_space_names_are_config_fields():
    from back40computing_amd import _native
    fields = {f for f, _ in _native.BfsConfig._fields_}
    for space in (autotune.SPACE_SINGLE, autotune.SPACE_LATTICE, autotune.SPACE_BATCH):
        assert set(space) <= fields
