"""The autotuner on the GPU: it times every candidate shape, applies the fastest, and
Simulation(autotune=True) uses its choice only around its own runs."""
import numpy as np
import pytest

from gravhopper_amd import Simulation, _native, autotune, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def cache(tmp_path, monkeypatch):
    monkeypatch.setenv("GRAVHOP_TUNE_CACHE", str(tmp_path / "tune.json"))


@pytest.fixture
def reset(ctx):
    yield
    for k in (_native.TUNE_DIRECT_SELF, _native.TUNE_WALK_GROUP):
        ctx.set_tuning(k, 0)


def test_tune_direct_self(ctx, reset):
    best = autotune.tune_direct_self(ctx, 20000)
    assert best in autotune.SELF_CANDIDATES
    assert ctx.get_tuning(_native.TUNE_DIRECT_SELF) == best
    entry = autotune._load()["gfx950:direct_self:32768"]
    assert entry["choice"] == best and len(entry["seconds"]) == len(autotune.SELF_CANDIDATES)
    assert min(entry["seconds"].values()) == entry["seconds"][str(best)]


def test_tune_walk_group(ctx, reset):
    ic = workloads.plummer(100000, seed=71)
    best = autotune.tune_walk_group(ctx, ic["pos"], ic["mass"], 1e-4)
    assert best in autotune.GROUP_CANDIDATES
    assert ctx.get_tuning(_native.TUNE_WALK_GROUP) == best


def test_tune_ensemble(ctx):
    rng = np.random.default_rng(72)
    sims = []
    for k in range(3):
        ic = workloads.plummer(int(rng.integers(1000, 4000)), seed=720 + k)
        ic["eps"], ic["dt"] = 2e-5, 1e-3
        sims.append(ic)
    assert autotune.tune_ensemble(ctx, sims) in autotune.ENSEMBLE_CANDIDATES


@pytest.mark.parametrize("algorithm,n", [("direct", 20000), ("tree", 70000)])
def test_simulation_autotune(ctx, reset, algorithm, n):
    ic = workloads.plummer(n, seed=73)
    res = {}
    for tuned in (False, True):
        sim = Simulation(dt=0.001, eps=1e-5, algorithm=algorithm, precision='fp32',
                         autotune=tuned)
        sim.add_IC(ic)
        sim.run(2)
        res[tuned] = sim.positions[-1]
        # the choice is applied around the runs only
        assert ctx.get_tuning(_native.TUNE_DIRECT_SELF) == 0
        assert ctx.get_tuning(_native.TUNE_WALK_GROUP) == 0
    key = sim._tuned[2]
    assert key == (_native.TUNE_DIRECT_SELF if algorithm == 'direct'
                   else _native.TUNE_WALK_GROUP)
    disp = np.linalg.norm(res[False] - ic["pos"], axis=1)
    diff = np.linalg.norm(res[True] - res[False], axis=1)
    # the same forces to FP32 rounding (direct) or to the walk's lists (tree)
    assert np.median(diff / disp) < (1e-5 if algorithm == 'direct' else 5e-3)
