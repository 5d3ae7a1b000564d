"""Allocation-failure paths, driven by an injected cap on the context's device arenas.

The reference exits the process when its tree allocation fails (_jbgrav.c:364,779); the
library instead returns GH_ENOMEM with a message and keeps the context usable.  The cap
(GH_TUNE_ALLOC_LIMIT_MB) makes that path reachable without exhausting the GPU.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from gravhopper_amd import _native, workloads

pytestmark = pytest.mark.gpu

FP64, FP32 = _native.GH_FP64, _native.GH_FP32


@pytest.fixture
def capped():
    c = _native.Context(0)
    yield c
    c.close()


def test_arena_cap_fails_cleanly_and_context_recovers(ctx, capped):
    g = workloads.galaxy(n_halo=50000, n_disk=50000, seed=83)
    pos, vel, mass = g["pos"], g["vel"], g["mass"]
    capped.set_tuning(_native.TUNE_ALLOC_LIMIT_MB, 1)
    calls = [
        ("direct", lambda c: c.direct(pos, mass, 0.05, FP32)),
        ("tree", lambda c: c.tree(pos, mass, 0.05, 0.7, FP32)),
        ("tree_fp64", lambda c: c.tree(pos, mass, 0.05, 0.7, FP64)),
        ("run_tree", lambda c: c.run(pos, vel, mass, 0.01, 0.05, 3,
                                     algorithm=_native.GH_TREE, precision=FP32)[:2]),
    ]
    for name, f in calls:
        print("capped call", name, flush=True)
        with pytest.raises(_native.GravhopError, match="arena limit") as e:
            f(capped)
        assert e.value.code == _native.GH_ENOMEM, name
    # lifting the cap: the same context computes what a fresh one does
    capped.set_tuning(_native.TUNE_ALLOC_LIMIT_MB, 0)
    for name, f in calls:
        print("uncapped call", name, flush=True)
        got, want = f(capped), f(ctx)
        if isinstance(got, tuple):
            for a, b in zip(got, want):
                assert_bitwise(a, b, name)
        else:
            assert_bitwise(got, want, name)


def test_arena_cap_above_need_is_transparent(ctx, capped):
    # a cap the call fits under changes nothing; a later, larger call then hits it
    small = workloads.plummer(4096, seed=84)
    capped.set_tuning(_native.TUNE_ALLOC_LIMIT_MB, 64)
    got = capped.tree(small["pos"], small["mass"], 1e-3, 0.7, FP64)
    assert_bitwise(got, ctx.tree(small["pos"], small["mass"], 1e-3, 0.7, FP64))
    rng = np.random.default_rng(85)
    big = rng.normal(size=(2_000_000, 3))
    with pytest.raises(_native.GravhopError, match="arena limit"):
        capped.tree(big, np.ones(len(big)), 1e-3, 0.7, FP32)
    got = capped.tree(small["pos"], small["mass"], 1e-3, 0.7, FP64)
    assert_bitwise(got, ctx.tree(small["pos"], small["mass"], 1e-3, 0.7, FP64))


def test_ensemble_create_destroy_under_cap(capped):
    # every ensemble's buffers go back to the arena count when it is destroyed, so a
    # cap that fits one ensemble fits any number of them made one after another
    sims = []
    for k in range(4):
        ic = workloads.plummer(20000, seed=860 + k)
        ic["eps"], ic["dt"] = 2e-5, 1e-3
        sims.append(ic)
    capped.set_tuning(_native.TUNE_ALLOC_LIMIT_MB, 256)
    first = None
    for rep in range(12):
        ens = _native.Ensemble(capped, sims, ipt=2)
        ens.run(1)
        pos, _ = ens.state()
        ens.close()
        if first is None:
            first = pos
        for a, b in zip(pos, first):
            assert_bitwise(a, b, "repeat %d" % rep)
