"""The autotuner's host logic: size buckets, the JSON cache, and cache hits that apply
the stored choice without timing anything (no GPU needed)."""
import json

import pytest

from gravhopper_amd import _native, autotune


class FakeContext:
    def __init__(self):
        self.calls = []

    def set_tuning(self, key, value):
        self.calls.append((key, value))


@pytest.fixture
def cache(tmp_path, monkeypatch):
    p = tmp_path / "tune.json"
    monkeypatch.setenv("GRAVHOP_TUNE_CACHE", str(p))
    return p


def test_buckets():
    assert [autotune.bucket(n) for n in (0, 1, 2, 3, 1024, 1025)] == [1, 1, 2, 4, 1024, 2048]


def test_cache_round_trip(cache):
    assert autotune._load() == {}
    autotune._store({"a": {"choice": 4}})
    assert json.loads(cache.read_text()) == {"a": {"choice": 4}}
    assert autotune._load() == {"a": {"choice": 4}}
    cache.write_text("{not json")
    assert autotune._load() == {}


def test_cache_hits_apply_the_stored_choice(cache):
    autotune._store({
        "gfx950:direct_self:32768": {"choice": 2},
        "gfx950:walk_group:131072:0.7": {"choice": 32},
        "gfx950:ensemble:3:%d" % autotune.bucket(3 * 1000 ** 2): {"choice": [4, True]},
    })
    ctx = FakeContext()
    assert autotune.tune_direct_self(ctx, 20000) == 2
    assert autotune.tune_walk_group(ctx, [[0.0, 0.0, 0.0]] * 100000, [1.0] * 100000,
                                    1e-3) == 32
    assert ctx.calls == [(_native.TUNE_DIRECT_SELF, 2), (_native.TUNE_WALK_GROUP, 32)]
    sims = [{"mass": [1.0] * 1000}] * 3
    assert autotune.tune_ensemble(ctx, sims) == (4, True)


def test_tune_for_run_dispatch(cache):
    autotune._store({"gfx950:direct_self:4096": {"choice": -1},
                     "gfx950:walk_group:4096:0.5": {"choice": 8}})
    ctx = FakeContext()
    pos, mass = [[0.0, 0.0, 0.0]] * 3000, [1.0] * 3000
    assert autotune.tune_for_run(ctx, pos, mass, 1e-3, "direct", "fp64") is None
    assert autotune.tune_for_run(ctx, pos, mass, 1e-3, "direct", "fp32") == (
        _native.TUNE_DIRECT_SELF, -1)
    assert autotune.tune_for_run(ctx, pos, mass, 1e-3, "tree", "fp32", theta=0.5) == (
        _native.TUNE_WALK_GROUP, 8)
