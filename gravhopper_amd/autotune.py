"""Kernel-shape autotuner.

Times the candidate shapes of a kernel on the problem at hand (device-resident,
HIP-event timing inside libgravhop.so), keeps the fastest, and caches the choice per
(GPU architecture, kernel, size bucket) in a JSON file ($GRAVHOP_TUNE_CACHE, default
~/.cache/gravhopper_amd/tune.json).  Shapes only change FP32 rounding, never the
algorithm.

    from gravhopper_amd import autotune
    autotune.tune_direct_self(ctx, n)              # sets GH_TUNE_DIRECT_SELF on ctx
    autotune.tune_walk_group(ctx, pos, mass, eps)  # sets GH_TUNE_WALK_GROUP on ctx
    ipt, persistent = autotune.tune_ensemble(ctx, sims)
    key, value = autotune.tune_for_run(ctx, pos, mass, eps, 'direct', 'fp32')

Simulation(..., autotune=True) calls tune_for_run before its FP32 runs and applies
the choice only around its own runs.
"""
import json
import os
import time

import numpy as np

from . import _native, workloads

ARCH = "gfx950"
SELF_CANDIDATES = (-1, 1, 2, 4)   # one-sided kernel, pair-once with 1/2/4 per lane
GROUP_CANDIDATES = (8, 16, 32)    # targets per FP32 walk group
ENSEMBLE_CANDIDATES = tuple((ipt, p) for ipt in (1, 2, 4) for p in (False, True))


def _cache_path():
    return os.environ.get("GRAVHOP_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "gravhopper_amd",
                                       "tune.json"))


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def bucket(n):
    """Power-of-two size bucket."""
    return 1 << max(0, int(np.ceil(np.log2(max(int(n), 1)))))


def _time_direct(ctx, torch, x, m, n, eps, reps):
    acc = torch.empty((n, 3), dtype=torch.float64, device=x.device)
    ctx.direct_dev(x.data_ptr(), m.data_ptr(), n, eps, _native.GH_FP32, acc.data_ptr())
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.direct_dev(x.data_ptr(), m.data_ptr(), n, eps, _native.GH_FP32, acc.data_ptr())
    ctx.sync()
    return (time.perf_counter() - t0) / reps


def tune_direct_self(ctx, n, eps=1e-5, reps=2, use_cache=True):
    """Chooses the FP32 self-field kernel for N particles; returns the
    GH_TUNE_DIRECT_SELF value and sets it on ctx."""
    key = "%s:direct_self:%d" % (ARCH, bucket(n))
    cache = _load() if use_cache else {}
    if key in cache:
        ctx.set_tuning(_native.TUNE_DIRECT_SELF, cache[key]["choice"])
        return cache[key]["choice"]
    import torch
    ic = workloads.plummer(n, seed=bucket(n))
    x = torch.tensor(ic["pos"], dtype=torch.float64, device="cuda")
    m = torch.tensor(ic["mass"], dtype=torch.float64, device="cuda")
    times = {}
    for cand in SELF_CANDIDATES:
        ctx.set_tuning(_native.TUNE_DIRECT_SELF, cand)
        times[cand] = _time_direct(ctx, torch, x, m, n, eps, reps)
    best = min(times, key=times.get)
    ctx.set_tuning(_native.TUNE_DIRECT_SELF, best)
    cache[key] = {"choice": best, "seconds": {str(k): v for k, v in times.items()}}
    if use_cache:
        _store(cache)
    return best


def tune_ensemble(ctx, sims, steps=2, use_cache=True):
    """Chooses (particles per lane, persistent) for an ensemble; returns the pair."""
    sig = bucket(sum(len(s["mass"]) ** 2 for s in sims))
    key = "%s:ensemble:%d:%d" % (ARCH, len(sims), sig)
    cache = _load() if use_cache else {}
    if key in cache:
        return tuple(cache[key]["choice"])
    times = {}
    for ipt, persistent in ENSEMBLE_CANDIDATES:
        ens = _native.Ensemble(ctx, sims, ipt=ipt, persistent=persistent)
        ens.run(1)
        ctx.sync()
        t0 = time.perf_counter()
        ens.run(steps)
        ctx.sync()
        times[(ipt, persistent)] = (time.perf_counter() - t0) / steps
        ens.close()
    best = min(times, key=times.get)
    cache[key] = {"choice": list(best),
                  "seconds": {"%d,%d" % k: v for k, v in times.items()}}
    if use_cache:
        _store(cache)
    return best


def _time_tree(ctx, torch, x, m, n, eps, theta, reps):
    acc = torch.empty((n, 3), dtype=torch.float64, device=x.device)
    ctx.tree_dev(x.data_ptr(), m.data_ptr(), n, eps, theta, _native.GH_FP32, acc.data_ptr())
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.tree_dev(x.data_ptr(), m.data_ptr(), n, eps, theta, _native.GH_FP32,
                     acc.data_ptr())
    ctx.sync()
    return (time.perf_counter() - t0) / reps


def tune_walk_group(ctx, pos, mass, eps, theta=0.7, reps=2, use_cache=True):
    """Chooses the FP32 walk's group size for these particles (the whole tree step:
    build and walk); returns the GH_TUNE_WALK_GROUP value and sets it on ctx.  Cached
    per size bucket and theta."""
    n = len(mass)
    key = "%s:walk_group:%d:%g" % (ARCH, bucket(n), float(theta))
    cache = _load() if use_cache else {}
    if key in cache:
        ctx.set_tuning(_native.TUNE_WALK_GROUP, cache[key]["choice"])
        return cache[key]["choice"]
    import torch
    x = torch.tensor(np.ascontiguousarray(pos, dtype=np.float64), device="cuda")
    m = torch.tensor(np.ascontiguousarray(mass, dtype=np.float64), device="cuda")
    times = {}
    for cand in GROUP_CANDIDATES:
        ctx.set_tuning(_native.TUNE_WALK_GROUP, cand)
        times[cand] = _time_tree(ctx, torch, x, m, n, eps, theta, reps)
    best = min(times, key=times.get)
    ctx.set_tuning(_native.TUNE_WALK_GROUP, best)
    cache[key] = {"choice": best, "seconds": {str(k): v for k, v in times.items()}}
    if use_cache:
        _store(cache)
    return best


def tune_for_run(ctx, pos, mass, eps, algorithm, precision, theta=0.7, use_cache=True):
    """Tunes the kernel shape a run of these particles uses: the FP32 self-field kernel
    for 'direct', the FP32 walk's group size for 'tree'.  Returns (tuning key, value),
    or None for FP64 runs (their kernels are chosen by size and give the same bits
    whatever the shape)."""
    if precision != "fp32":
        return None
    if algorithm == "direct":
        return _native.TUNE_DIRECT_SELF, tune_direct_self(ctx, len(mass), eps=eps,
                                                          use_cache=use_cache)
    return _native.TUNE_WALK_GROUP, tune_walk_group(ctx, pos, mass, eps, theta=theta,
                                                    use_cache=use_cache)
