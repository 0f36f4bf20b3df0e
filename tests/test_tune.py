"""The autotuner's search and cache logic (CPU; the knobs are host-side
settings of libfqsl.so, so setting them needs no device)."""
import pytest

from fastqsl_amd import _lib, tune


def test_coordinate_descent_finds_the_separable_optimum():
    target = {"order": 2, "order_buckets": 16, "scott_waves": 0, "refill_min": 16, "queues": 8,
              "grid_pct": 75}
    cost = lambda c: 10 + sum(abs(hash((k, c[k])) % 7) * (c[k] != target[k]) + (c[k] != target[k])
                              for k in c)
    best, ms, seen = tune.coordinate_descent(cost)
    assert best == target and ms == 10
    assert len(seen) <= sum(len(v) for v in tune.SPACE.values()) * 2


def test_coordinate_descent_keeps_defaults_without_a_clear_gain():
    best, ms, _ = tune.coordinate_descent(lambda c: 5.0 - 0.001 * (c["refill_min"] == 4))
    assert best == tune.DEFAULTS


def test_cache_roundtrip_and_key(tmp_path, monkeypatch):
    monkeypatch.setenv("FQSL_TUNE_CACHE", str(tmp_path / "t.json"))
    k = tune.cache_key("MI355X", (256, 256, 256), "rkf45", "fp32c", 1e-6)
    assert k == "MI355X|256x256x256|rkf45|fp32c|tol1e-6"
    assert tune.cache_key("d", (1, 2, 3), "rk4", "fp64", 0.0).endswith("tolnone")
    assert tune.load_cache() == {}
    tune.save_cache({k: {"cfg": tune.DEFAULTS}})
    assert tune.load_cache()[k]["cfg"] == tune.DEFAULTS


def test_apply_sets_the_library_knobs():
    before = tune.current()
    try:
        tune.apply({"refill_min": 16, "queues": 8})
        assert _lib.tune_get("refill_min") == 16 and _lib.tune_get("queues") == 8
        with pytest.raises(_lib.FqslError):
            tune.apply({"queues": 3})
    finally:
        tune.apply(before)
    assert tune.current() == before


def test_knob_ranges_are_checked():
    """The C ABI rejects out-of-range knob values with a message and keeps the
    old value (host code only: no GPU needed)."""
    from fastqsl_amd import _lib
    before = _lib.tune_get("plain_refill")
    for knob, bad in (("plain_refill", 0), ("plain_refill", 65), ("refill_min", 0), ("queues", 3),
                      ("scott_waves", 1), ("order", 3), ("order_buckets", 3), ("grid_pct", 0)):
        with pytest.raises(_lib.FqslError):
            _lib.tune_set(knob, bad)
    assert _lib.tune_get("plain_refill") == before
    _lib.tune_set("plain_refill", 16)
    try:
        assert _lib.tune_get("plain_refill") == 16
    finally:
        _lib.tune_set("plain_refill", before)
