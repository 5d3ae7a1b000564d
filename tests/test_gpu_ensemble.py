"""Batched ensembles (gh_ensemble): every simulation of the batch must advance exactly
as it would alone through the same pair-once kernel, with or without the persistent
work queue and with or without HIP-graph replay (replayed several times with eager
work and synchronisation in between)."""
import numpy as np
import pytest

from conftest import assert_bitwise, rel_err
from gravhopper_amd import _native, workloads

pytestmark = pytest.mark.gpu


def small_ensemble(seed=5, n_sims=6):
    rng = np.random.default_rng(seed)
    sims = []
    for k in range(n_sims):
        n = int(rng.integers(300, 5000))
        b = rng.uniform(0.5e-3, 2e-3)
        m = 10 ** rng.uniform(5, 7)
        ic = workloads.plummer(n, b=b, mass=m, seed=seed * 1000 + k)
        ic["eps"] = 0.02 * b
        ic["dt"] = 0.005 * workloads.plummer_crossing_time(b, m)
        sims.append(ic)
    return sims


@pytest.mark.parametrize("ipt", [1, 2, 4])
def test_ensemble_equals_single_runs(ctx, ipt):
    sims = small_ensemble()
    ens = _native.Ensemble(ctx, sims, ipt=ipt)
    ens.run(4, graph=False)
    pos, vel = ens.state()
    ens.close()
    ctx.set_tuning(_native.TUNE_DIRECT_SELF, ipt)
    try:
        for k, s in enumerate(sims):
            x, v = ctx.run(s["pos"], s["vel"], s["mass"], s["dt"], s["eps"], 4,
                           precision=_native.GH_FP32)
            assert_bitwise(pos[k], x, "sim %d pos" % k)
            assert_bitwise(vel[k], v, "sim %d vel" % k)
    finally:
        ctx.set_tuning(_native.TUNE_DIRECT_SELF, 0)


def test_ensemble_graph_and_persistent_replays(ctx):
    sims = small_ensemble(seed=6, n_sims=5)
    eager = _native.Ensemble(ctx, sims, ipt=2)
    graph = _native.Ensemble(ctx, sims, ipt=2, persistent=True)
    for rep in range(4):
        eager.run(2, graph=False)
        graph.run(2, graph=True)
        pe, ve = eager.state()      # eager host work and a synchronisation between replays
        pg, vg = graph.state()
        for k in range(len(sims)):
            assert_bitwise(pg[k], pe[k], "replay %d sim %d pos" % (rep, k))
            assert_bitwise(vg[k], ve[k], "replay %d sim %d vel" % (rep, k))
    eager.close()
    graph.close()


def test_ensemble_accuracy_vs_fp64(ctx, oracle):
    sims = small_ensemble(seed=7, n_sims=3)
    ens = _native.Ensemble(ctx, sims, ipt=1)
    ens.run(3)
    pos, _ = ens.state()
    ens.close()
    for k, s in enumerate(sims):
        xo, _ = oracle.dkd_direct(s["pos"], s["vel"], s["mass"], s["dt"], s["eps"], 3)
        disp = np.linalg.norm(xo - s["pos"], axis=1)
        err = np.linalg.norm(pos[k] - xo, axis=1)
        assert np.median(err / disp) < 1e-5, k


def test_ensemble_info_and_errors(ctx):
    sims = small_ensemble(seed=8, n_sims=2)
    ens = _native.Ensemble(ctx, sims, ipt=1)
    info = ens.info()
    n = [len(s["mass"]) for s in sims]
    assert info["pairs_per_step"] == sum(k * (k - 1) for k in n)
    ens.close()
    with pytest.raises(_native.GravhopError):
        _native.Ensemble(ctx, sims, ipt=3)
