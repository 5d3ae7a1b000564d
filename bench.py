"""Benchmark of the gravity engine on one MI355X.

Headline workload (the metric's direct-sum case): Plummer sphere, N = 262144 equal
masses, b = 1 pc, M = 1e6 Msun, eps = 0.01 pc, dt = 0.001 Myr; drift-kick-drift steps
with FP32 pair forces, compensated accumulation and FP64 state, run device-resident
(inputs in HBM before the timed region).  One step = one direct-sum force evaluation
(N(N-1) pair interactions) + one fused leapfrog update.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  `value` is pair interactions per second over the K timed steps.
Extra fields report the FP64 comparison kernel, force errors against the FP64 direct
sum, energy drift, the tree step at N = 1M (when built) and the reference C engine
timed on one host core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0         # MI355X published HBM3E bandwidth
FLOP_PER_PAIR = 20            # 3 sub, 3 mul + 3 add (r^2 + eps^2), rsqrt = 1, 2 mul (r^-3),
                              # 1 mul (m r^-3), 3 FMA = 6
PEAK_VALU_LANE_OPS = PEAK_FP32_TFLOPS * 1e12 / 2  # lane instructions/s (an FMA is 2 FLOP)
G_KMS2_KPC_MSUN = 4.398600412921223e-09 / 1.022712165045695e-03  # G in (km/s)^2 kpc / Msun
N_DIRECT = 262144
EPS_DIRECT = 1e-5             # kpc (0.01 pc)
DT_DIRECT = 0.001             # Myr
SEED_DIRECT = 20260002


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy-kernel calibration), committed under profiles/."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def counters_from_profiles(name):
    """Per-launch counters of the headline kernel measured with rocprofv3 --pmc,
    committed under profiles/ (None when absent)."""
    p = os.path.join(ROOT, "profiles", name)
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def rel_err_stats(got, want):
    e = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    return float(np.median(e)), float(np.percentile(e, 99))


def total_energy(ctx, pos, vel, mass, eps):
    ke, pe, _ = ctx.energy(pos, vel, mass, eps)
    return ke + G_KMS2_KPC_MSUN * pe


def host_description():
    """CPU model and the cores this process may use (the baselines run on one of them)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "usable_cores": usable, "baseline_cores": 1}


def cpu_baseline(budget_s=30.0):
    """The reference's own C direct sum (gravhopper/_jbgrav.c, built from its source
    into oracle/_ref) on one host core; falls back to the oracle restatement."""
    from gravhopper_amd import workloads
    from oracle import oracle as O
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    ref = O.ref_module()
    n = 16384 if ref is not None else 32768
    ic = workloads.plummer(n, b=1e-3, mass=1e6, seed=SEED_DIRECT)
    times = []
    spent = 0.0
    while len(times) < 3 and spent < budget_s:
        t0 = time.perf_counter()
        if ref is not None:
            ref.direct_summation(ic["pos"], ic["mass"], EPS_DIRECT)
        else:
            O.direct(ic["pos"], ic["mass"], EPS_DIRECT, nthreads=1)
        times.append(time.perf_counter() - t0)
        spent += times[-1]
    if ref is not None:
        kind = "reference"
        what = "gravhopper/_jbgrav.c direct_summation (built from the reference source)"
    else:
        kind = "port"
        what = "oracle/jbgrav_oracle.c orc_direct (restatement, same FP64 op order)"
    dt = float(np.median(times))
    return {"value": n * (n - 1) / dt, "unit": "pair interactions/s", "cores": 1,
            "kind": kind,
            "sample": "%s, Plummer N=%d, median of %d FP64 force evaluations (%.1f s each, "
                      "%.1f s total); the reference cannot run N=262144 (N^2 scratch of "
                      "2.2 TB, 32-bit index overflow above N=26754)" % (
                          what, n, len(times), dt, spent)}


def cpu_tree_baseline(n_targets=131072):
    """The reference's own C tree (gravhopper/_jbgrav.c tree_force_position, built from
    its source into oracle/_ref) on one host core at N = 1048576: one build with a single
    target, then one build with a bounded sample of targets; the walk time per target is
    extrapolated to all N targets."""
    from gravhopper_amd import workloads
    from oracle import oracle as O
    ref = O.ref_module()
    if ref is None:
        return None
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    g = workloads.galaxy(seed=3)
    n = len(g["mass"])
    sub = np.random.default_rng(1).choice(n, n_targets, replace=False)
    t0 = time.perf_counter()
    ref.tree_force_position(g["pos"], g["mass"], g["pos"][sub[:1]], 0.05, 0.7)
    t1 = time.perf_counter()
    ref.tree_force_position(g["pos"], g["mass"], g["pos"][sub], 0.05, 0.7)
    t2 = time.perf_counter()
    build = t1 - t0
    walk = max((t2 - t1) - build, 0.0) / n_targets * n
    return {"value": 1.0 / (build + walk), "unit": "tree steps/s", "cores": 1,
            "kind": "reference",
            "build_s": build, "walk_s_extrapolated": walk,
            "sample": "gravhopper/_jbgrav.c tree_force_position (built from the reference "
                      "source), halo+disk N=%d, theta=0.7: one build, then one build plus "
                      "a walk over %d random targets; walk time scaled to all %d targets "
                      "(%.1f s measured)" % (n, n_targets, n, t2 - t0)}


def bench_plummer2000_run(ctx):
    """The reference's CPU-runnable case: Plummer N=2000 (b=1 pc, 1e6 Msun), eps=0.05 pc,
    dt=0.005 Myr, 400 FP64 DKD steps with the direct sum, end to end through gh_run with
    the full history, as Simulation.run does it."""
    from gravhopper_amd import _native, workloads
    ic = workloads.plummer(2000, b=1e-3, mass=1e6, seed=20260001)
    ctx.run(ic["pos"], ic["vel"], ic["mass"], 0.005, 5e-5, 4, history=True)  # warm
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.run(ic["pos"], ic["vel"], ic["mass"], 0.005, 5e-5, 400, precision=_native.GH_FP64,
                history=True)
        runs.append(time.perf_counter() - t0)
    el = sorted(runs)[1]
    return {"workload": "Plummer N=2000, eps=0.05 pc, dt=0.005 Myr, 400 FP64 DKD steps, "
                        "direct sum, history of 400 snapshots (median of 3 runs)",
            "seconds": el, "ms_per_step": el / 400 * 1e3, "runs_s": runs}


def cpu_plummer2000_baseline(n_forces=40):
    """The reference's C direct sum (built from its source) for the N=2000 case, on one
    core: n_forces force evaluations, scaled to the run's 400 (the leapfrog itself is a
    few numpy passes per step, negligible at this N)."""
    from gravhopper_amd import workloads
    from oracle import oracle as O
    ref = O.ref_module()
    if ref is None:
        return None
    ic = workloads.plummer(2000, b=1e-3, mass=1e6, seed=20260001)
    t0 = time.perf_counter()
    for _ in range(n_forces):
        ref.direct_summation(ic["pos"], ic["mass"], 5e-5)
    el = time.perf_counter() - t0
    return {"value": el / n_forces * 400, "unit": "s per 400-step run (extrapolated)",
            "cores": 1, "kind": "reference",
            "sample": "gravhopper/_jbgrav.c direct_summation, Plummer N=2000, %d force "
                      "evaluations (%.1f s), scaled to 400" % (n_forces, el)}


def bench_tree(ctx, torch, steps, warmup, lite=False):
    """Barnes-Hut step at N = 1048576: Hernquist halo + exponential disk, theta = 0.7,
    eps = 0.05 kpc, dt = 0.1 Myr, FP32 monopoles, FP64 state."""
    from gravhopper_amd import _native, workloads
    g = workloads.galaxy(seed=3)
    n = len(g["mass"])
    eps, dt, theta = 0.05, 0.1, 0.7
    dev = torch.device("cuda", ctx.device)
    x = torch.tensor(g["pos"], dtype=torch.float64, device=dev).contiguous()
    v = torch.tensor(g["vel"], dtype=torch.float64, device=dev).contiguous()
    m = torch.tensor(g["mass"], dtype=torch.float64, device=dev).contiguous()
    torch.cuda.synchronize()  # library kernels run on the context's own stream
    FP32, FP64, TREE = _native.GH_FP32, _native.GH_FP64, _native.GH_TREE
    if not lite:
        ke0, pe0, _ = ctx.energy(g["pos"], g["vel"], g["mass"], eps)
    ctx.run_dev(x.data_ptr(), v.data_ptr(), m.data_ptr(), n, dt, eps, warmup,
                algorithm=TREE, theta=theta, precision=FP32)
    ctx.sync()
    torch.cuda.synchronize()
    # the step time by the wall clock without per-kernel events, then the same number of
    # steps with HIP events per kernel class for the build / walk / update split
    t0 = time.perf_counter()
    ctx.run_dev(x.data_ptr(), v.data_ptr(), m.data_ptr(), n, dt, eps, steps,
                algorithm=TREE, theta=theta, precision=FP32)
    ctx.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ctx.timing_enable(True)
    ctx.run_dev(x.data_ptr(), v.data_ptr(), m.data_ptr(), n, dt, eps, steps,
                algorithm=TREE, theta=theta, precision=FP32)
    ctx.sync()
    torch.cuda.synchronize()
    b_ms, b_n = ctx.timing_read(_native.T_TREE_BUILD)
    w_ms, w_n = ctx.timing_read(_native.T_TREE_WALK)
    u_ms, u_n = ctx.timing_read(_native.T_UPDATE)
    ctx.timing_enable(False)
    out = {"workload": "Hernquist halo 524288 (a=10 kpc, 1e12 Msun) + exp disk 524288 "
                       "(Rd=3 kpc, z0=0.3 kpc, 5e10 Msun), theta=0.7, eps=0.05 kpc, "
                       "dt=0.1 Myr, FP32 monopoles, FP64 state",
           "n_particles": n, "steps": steps, "steps_per_s": steps / el,
           "ms_per_step": el / steps * 1e3,
           "build_ms": b_ms / max(b_n, 1), "walk_ms": w_ms / max(w_n, 1),
           "update_ms": u_ms / max(u_n, 1)}
    if lite:
        return out
    # force accuracy on a 16384-particle subsample against the FP64 direct sum
    pos = x.cpu().numpy()
    mass = g["mass"]
    # total-energy drift over the warmup and timed steps (FP64 pair sums)
    ke1, pe1, _ = ctx.energy(pos, v.cpu().numpy(), mass, eps)
    G = G_KMS2_KPC_MSUN
    e0, e1 = ke0 + G * pe0, ke1 + G * pe1
    out["energy_rel_drift"] = abs(e1 - e0) / abs(e0)
    out["energy_drift_steps"] = warmup + 2 * steps
    rng = np.random.default_rng(16384)
    sub = rng.choice(n, 16384, replace=False)
    acc_t = torch.empty((n, 3), dtype=torch.float64, device=dev)
    ctx.tree_dev(x.data_ptr(), m.data_ptr(), n, eps, theta, FP32, acc_t.data_ptr())
    tsub = torch.tensor(pos[sub], dtype=torch.float64, device=dev).contiguous()
    acc_d = torch.empty((len(sub), 3), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.direct_dev(x.data_ptr(), m.data_ptr(), n, eps, FP64, acc_d.data_ptr(),
                   d_force_pos=tsub.data_ptr(), nf=len(sub))
    ctx.sync()
    at = acc_t.cpu().numpy()[sub]
    ad = acc_d.cpu().numpy()
    e = np.linalg.norm(at - ad, axis=1) / np.linalg.norm(ad, axis=1)
    out["force_rel_err_median"] = float(np.median(e))
    out["force_rel_err_p99"] = float(np.percentile(e, 99))
    # interaction count of the reference walk (FP64 walk, same accepted nodes)
    _, cnt = ctx.tree_interactions(pos, mass, eps, theta)
    out["interactions_per_target"] = float(cnt.mean())
    out["walk_achieved_tflops"] = 20.0 * cnt.sum() / (out["walk_ms"] * 1e-3) / 1e12
    out["walk_roofline_frac"] = out["walk_achieved_tflops"] / PEAK_FP32_TFLOPS
    # opening angle trade-off: step time and force error at smaller theta
    sweep = []
    for th in (0.7, 0.5, 0.4, 0.3):
        xs = torch.tensor(pos, dtype=torch.float64, device=dev)
        vs = torch.tensor(g["vel"], dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.run_dev(xs.data_ptr(), vs.data_ptr(), m.data_ptr(), n, dt, eps, 1,
                    algorithm=TREE, theta=th, precision=FP32)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.run_dev(xs.data_ptr(), vs.data_ptr(), m.data_ptr(), n, dt, eps, 3,
                    algorithm=TREE, theta=th, precision=FP32)
        ctx.sync()
        st = (time.perf_counter() - t0) / 3
        ctx.tree_dev(x.data_ptr(), m.data_ptr(), n, eps, th, FP32, acc_t.data_ptr())
        ctx.sync()
        e = np.linalg.norm(acc_t.cpu().numpy()[sub] - ad, axis=1) / np.linalg.norm(ad, axis=1)
        sweep.append({"theta": th, "ms_per_step": st * 1e3,
                      "force_rel_err_median": float(np.median(e)),
                      "force_rel_err_p99": float(np.percentile(e, 99))})
    out["theta_sweep"] = sweep
    # total-energy drift over the configuration's 50 steps from the initial state
    xs = torch.tensor(g["pos"], dtype=torch.float64, device=dev)
    vs = torch.tensor(g["vel"], dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.run_dev(xs.data_ptr(), vs.data_ptr(), m.data_ptr(), n, dt, eps, 50,
                algorithm=TREE, theta=theta, precision=FP32)
    ctx.sync()
    e50 = total_energy(ctx, xs.cpu().numpy(), vs.cpu().numpy(), mass, eps)
    out["energy_rel_drift_50_steps"] = abs(e50 - e0) / abs(e0)
    # the FP64 walk (the default precision of Simulation): the reference's accepted
    # nodes, FP64 monopoles
    xs = torch.tensor(pos, dtype=torch.float64, device=dev)
    vs = torch.tensor(g["vel"], dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.run_dev(xs.data_ptr(), vs.data_ptr(), m.data_ptr(), n, dt, eps, 1,
                algorithm=TREE, theta=theta, precision=FP64)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run_dev(xs.data_ptr(), vs.data_ptr(), m.data_ptr(), n, dt, eps, 3,
                algorithm=TREE, theta=theta, precision=FP64)
    ctx.sync()
    out["fp64_ms_per_step"] = (time.perf_counter() - t0) / 3 * 1e3
    # the FP32 direct sum at the same N, one evaluation
    ctx.timing_enable(True)
    ctx.direct_dev(x.data_ptr(), m.data_ptr(), n, eps, FP32, acc_t.data_ptr())
    ctx.sync()
    d_ms, _ = ctx.timing_read(_native.T_DIRECT_F32)
    ctx.timing_enable(False)
    out["direct_fp32_ms_same_n"] = d_ms
    out["tree_speedup_over_direct"] = d_ms / out["ms_per_step"]
    return out


def bench_ensemble(ctx, torch, steps):
    """128 independent Plummer spheres, N log-uniform in [1024, 65536], b in
    [0.5, 2] pc, M in [1e5, 1e7] Msun, eps = 0.02 b, dt = 0.005 t_cr; FP32 pairs in one
    launch per step (graph-replayed), against each simulation stepped on its own."""
    from gravhopper_amd import _native, autotune, workloads
    sims = workloads.ensemble(seed=4)
    ipt, persistent = autotune.tune_ensemble(ctx, sims, use_cache=False)
    ens = _native.Ensemble(ctx, sims, ipt=ipt, persistent=persistent)
    info = ens.info()
    ens.run(1)
    ctx.sync()
    t0 = time.perf_counter()
    ens.run(steps, graph=True)
    ctx.sync()
    el = time.perf_counter() - t0
    pos_end, vel_end = ens.state()
    ens.close()
    pairs = info["pairs_per_step"]
    out = {"workload": "128 Plummer spheres, N log-uniform [1024, 65536], eps=0.02 b, "
                       "dt=0.005 t_cr, FP32 pairs, FP64 state",
           "n_sims": len(sims), "n_particles": int(sum(len(s["mass"]) for s in sims)),
           "pairs_per_step": pairs, "ipt": ipt, "persistent": bool(persistent),
           "steps": steps, "ms_per_step": el / steps * 1e3,
           "interactions_per_s": pairs * steps / el}
    # each simulation alone: device-resident, one step at a time
    per = []
    dev = torch.device("cuda", ctx.device)
    for s in sims:
        n = len(s["mass"])
        x = torch.tensor(s["pos"], dtype=torch.float64, device=dev)
        v = torch.tensor(s["vel"], dtype=torch.float64, device=dev)
        m = torch.tensor(s["mass"], dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.run_dev(x.data_ptr(), v.data_ptr(), m.data_ptr(), n, s["dt"], s["eps"], 1,
                    precision=_native.GH_FP32)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.run_dev(x.data_ptr(), v.data_ptr(), m.data_ptr(), n, s["dt"], s["eps"], 2,
                    precision=_native.GH_FP32)
        ctx.sync()
        per.append((time.perf_counter() - t0) / 2)
    per = np.array(per)
    out["sequential_ms_per_step_total"] = float(per.sum() * 1e3)
    out["sequential_interactions_per_s"] = pairs / float(per.sum())
    out["sequential_sim_step_ms_p50"] = float(np.percentile(per, 50) * 1e3)
    out["sequential_sim_step_ms_p99"] = float(np.percentile(per, 99) * 1e3)
    out["speedup_over_sequential"] = float(per.sum()) / (el / steps)
    # accuracy: every sphere advances bitwise as the single-sphere FP32 path does (GPU
    # test), so its force error is that path's against the FP64 kernel, on 8 spheres
    # across the size range; energy drift of every sphere over the 1 + steps steps
    order = np.argsort([len(s["mass"]) for s in sims])
    errs = []
    for k in order[np.linspace(0, len(sims) - 1, 8).round().astype(int)]:
        s = sims[k]
        a32 = ctx.direct(s["pos"], s["mass"], s["eps"], _native.GH_FP32)
        a64 = ctx.direct(s["pos"], s["mass"], s["eps"], _native.GH_FP64)
        errs.append(np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1))
    e = np.concatenate(errs)
    out["force_rel_err_median"] = float(np.median(e))
    out["force_rel_err_p99"] = float(np.percentile(e, 99))
    drift = []
    for k, s in enumerate(sims):
        e0 = total_energy(ctx, s["pos"], s["vel"], s["mass"], s["eps"])
        e1 = total_energy(ctx, pos_end[k], vel_end[k], s["mass"], s["eps"])
        drift.append(abs(e1 - e0) / abs(e0))
    out["energy_rel_drift_median"] = float(np.median(drift))
    out["energy_rel_drift_max"] = float(np.max(drift))
    out["energy_drift_steps"] = 1 + steps
    return out


def bench_adjoint(ctx, torch, steps=200, every=20):
    """Plummer N=16384, b=1 pc, 1e6 Msun, eps=0.05 pc, dt=0.005 Myr, FP32 forces, FP64
    state; loss = mean |x_final - x_target|^2 with the target run from masses scaled
    by 1 + 0.05 N(0,1); gradients w.r.t. x0, v0, m through the adjoint, checkpoint
    every 20 steps."""
    from gravhopper_amd import _native, workloads
    from gravhopper_amd.autograd import leapfrog
    n, dt, eps = 16384, 0.005, 5e-5
    ic = workloads.plummer(n, b=1e-3, mass=1e6, seed=5)
    rng = np.random.default_rng(6)
    m_t = ic["mass"] * (1 + 0.05 * rng.normal(size=n))
    dev = torch.device("cuda", ctx.device)
    xt = torch.tensor(ic["pos"], device=dev)
    vt = torch.tensor(ic["vel"], device=dev)
    mt = torch.tensor(m_t, device=dev)
    torch.cuda.synchronize()
    ctx.run_dev(xt.data_ptr(), vt.data_ptr(), mt.data_ptr(), n, dt, eps, steps,
                precision=_native.GH_FP32)
    ctx.sync()
    x0 = torch.tensor(ic["pos"], device=dev, requires_grad=True)
    v0 = torch.tensor(ic["vel"], device=dev, requires_grad=True)
    m = torch.tensor(ic["mass"], device=dev, requires_grad=True)
    # warm the kernels and workspaces
    x, _ = leapfrog(x0, v0, m, dt, eps, 2, checkpoint_every=every)
    ((x - xt) ** 2).sum(-1).mean().backward()
    torch.cuda.synchronize()
    x0.grad = v0.grad = m.grad = None
    ctx.timing_enable(True)
    t0 = time.perf_counter()
    x, v_end = leapfrog(x0, v0, m, dt, eps, steps, checkpoint_every=every)
    loss = ((x - xt) ** 2).sum(-1).mean()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    vjp_ms, vjp_n = ctx.timing_read(_native.T_VJP)
    f_ms, f_n = ctx.timing_read(_native.T_DIRECT_F32)
    ctx.timing_enable(False)
    fwd, bwd = (t1 - t0) / steps, (t2 - t1) / steps
    # accuracy: FP32 force against the FP64 kernel at the initial state, energy drift of
    # the forward run, and the gradients against the same adjoint in FP64
    f_med, f_p99 = rel_err_stats(ctx.direct(ic["pos"], ic["mass"], eps, _native.GH_FP32),
                                 ctx.direct(ic["pos"], ic["mass"], eps, _native.GH_FP64))
    e0 = total_energy(ctx, ic["pos"], ic["vel"], ic["mass"], eps)
    e1 = total_energy(ctx, x.detach().cpu().numpy(), v_end.detach().cpu().numpy(), ic["mass"],
                      eps)
    g32 = [t.grad.detach().cpu().numpy() for t in (x0, v0, m)]

    def grads(nsteps, precision):
        x0.grad = v0.grad = m.grad = None
        xe, ve = leapfrog(x0, v0, m, dt, eps, nsteps, checkpoint_every=every,
                          precision=precision)
        ((xe - xt) ** 2).sum(-1).mean().backward()
        torch.cuda.synchronize()
        return xe.detach().cpu().numpy(), ve.detach().cpu().numpy(), [
            t.grad.detach().cpu().numpy() for t in (x0, v0, m)]

    def grad_err(ga, gb):
        return {name: float(np.linalg.norm(a - b) / np.linalg.norm(b))
                for name, a, b in zip(("x0", "v0", "m"), ga, gb)}

    # the same run in FP64: at dt = 0.005 Myr (a third of a crossing time of the 1 pc
    # sphere) the 200-step run is chaotic, so FP32 and FP64 trajectories, and with them
    # the gradients, part ways; over one checkpoint segment (20 steps) they agree
    x64, v64, g64 = grads(steps, "fp64")
    e64 = total_energy(ctx, x64, v64, ic["mass"], eps)
    x32 = x.detach().cpu().numpy()
    disp = np.linalg.norm(x64 - ic["pos"], axis=1)
    _, _, g32s = grads(every, "fp32")
    _, _, g64s = grads(every, "fp64")
    return {"workload": "Plummer N=16384, eps=0.05 pc, dt=0.005 Myr, 200 steps, checkpoint "
                        "every 20, FP32 forces and VJP pairs, FP64 state",
            "force_rel_err_median": f_med, "force_rel_err_p99": f_p99,
            "energy_rel_drift": abs(e1 - e0) / abs(e0), "energy_drift_steps": steps,
            "energy_rel_drift_fp64_forward": abs(e64 - e0) / abs(e0),
            "final_pos_diff_vs_fp64_rel_median": float(np.median(
                np.linalg.norm(x32 - x64, axis=1) / disp)),
            "grad_rel_err_vs_fp64_adjoint_%d_steps" % steps: grad_err(g32, g64),
            "grad_rel_err_vs_fp64_adjoint_%d_steps" % every: grad_err(g32s, g64s),
            "loss": float(loss.detach()), "forward_ms_per_step": fwd * 1e3,
            "backward_ms_per_step": bwd * 1e3,
            "adjoint_over_forward": bwd / fwd,
            "vjp_kernel_ms": vjp_ms / max(vjp_n, 1),
            "force_kernel_ms": f_ms / max(f_n, 1),
            "vjp_over_force_kernel": (vjp_ms / max(vjp_n, 1)) / (f_ms / max(f_n, 1)),
            "grad_norms": [float(np.linalg.norm(g)) for g in g32]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_DIRECT)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-tree", action="store_true")
    ap.add_argument("--tree-steps", type=int, default=5)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=5)
    ap.add_argument("--only-tree", action="store_true",
                    help="run only the N=1M tree steps (for profiling)")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="run only the N=16384 adjoint benchmark (for profiling)")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="run only the 128-sphere ensemble benchmark (for profiling)")
    ap.add_argument("--only-plummer2000", action="store_true",
                    help="run only the N=2000 FP64 run with history (for profiling)")
    ap.add_argument("--tree-lite", action="store_true",
                    help="tree steps without the accuracy and comparison runs")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a known-size kernel "
                         "for calibrating HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs one process on one MI355X "
              "and has no multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from gravhopper_amd import _native, workloads

    ctx = _native.default_context()
    if args.calibrate_copy_mib > 0:
        nbytes = args.calibrate_copy_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("bench: calibration copy of %d MiB" % args.calibrate_copy_mib)
        del src, dst
    if args.only_tree:
        print(json.dumps({"tree": bench_tree(ctx, torch, args.tree_steps, 1,
                                             lite=args.tree_lite)}), flush=True)
        return
    if args.only_adjoint:
        print(json.dumps({"adjoint": bench_adjoint(ctx, torch)}), flush=True)
        return
    if args.only_ensemble:
        print(json.dumps({"ensemble": bench_ensemble(ctx, torch, args.ensemble_steps)}),
              flush=True)
        return
    if args.only_plummer2000:
        print(json.dumps({"plummer2000_fp64_run": bench_plummer2000_run(ctx)}), flush=True)
        return
    n = args.n
    ic = workloads.plummer(n, b=1e-3, mass=1e6, seed=SEED_DIRECT)
    dev = torch.device("cuda", ctx.device)
    x = torch.tensor(ic["pos"], dtype=torch.float64, device=dev).contiguous()
    v = torch.tensor(ic["vel"], dtype=torch.float64, device=dev).contiguous()
    m = torch.tensor(ic["mass"], dtype=torch.float64, device=dev).contiguous()
    torch.cuda.synchronize()  # library kernels run on the context's own stream
    FP32, FP64 = _native.GH_FP32, _native.GH_FP64

    log("bench: N=%d warmup=%d steps=%d" % (n, args.warmup, args.steps))
    ke0, pe0, _ = ctx.energy(ic["pos"], ic["vel"], ic["mass"], EPS_DIRECT)
    ctx.run_dev(x.data_ptr(), v.data_ptr(), m.data_ptr(), n, DT_DIRECT, EPS_DIRECT,
                args.warmup, precision=FP32)
    ctx.sync()
    torch.cuda.synchronize()
    ctx.timing_enable(True)
    t0 = time.perf_counter()
    ctx.run_dev(x.data_ptr(), v.data_ptr(), m.data_ptr(), n, DT_DIRECT, EPS_DIRECT,
                args.steps, precision=FP32)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    k_ms, k_n = ctx.timing_read(_native.T_DIRECT_F32)
    u_ms, u_n = ctx.timing_read(_native.T_UPDATE)
    ctx.timing_enable(False)
    elapsed = t1 - t0
    pairs = float(n) * (n - 1)
    value = pairs * args.steps / elapsed
    log("bench: %.3f s for %d steps, kernel avg %.3f ms" % (elapsed, args.steps, k_ms / max(k_n, 1)))

    k_avg_s = (k_ms / k_n) / 1e3 if k_n else float("nan")
    achieved_tflops = FLOP_PER_PAIR * pairs / k_avg_s / 1e12
    result = {
        "metric": "pair interactions/s (direct, N=262k FP32) + tree steps/s at N=1M; adjoint/fwd",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pairs, fp64 state",
        "data": "synthetic (seeded Plummer sphere)",
        "config": {"workload": "Plummer N=%d b=1pc M=1e6Msun eps=0.01pc dt=0.001Myr, "
                               "DKD leapfrog, FP32-compensated direct sum" % n,
                   "n_particles": n, "pairs_per_step": pairs, "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profiles("direct_f32_n262144_traffic.json"),
                     "kernel": "k_direct_sym_f32<4,false>", "kernel_avg_ms": k_avg_s * 1e3,
                     "flop_per_pair": FLOP_PER_PAIR},
    }
    cnt = counters_from_profiles("direct_f32_n262144_traffic.json")
    if cnt and cnt.get("valu_wave_instructions_per_launch") and n == cnt.get("n", n):
        # VALU issue from the counters: SQ_INSTS_VALU wave instructions per launch x 64
        # lanes over this run's kernel time, against the lane-instruction rate of the
        # FP32 peak (one FMA per lane = 2 FLOP)
        lane_ops = cnt["valu_wave_instructions_per_launch"] * 64 / k_avg_s
        result["roofline"].update({
            "valu_counter": "SQ_INSTS_VALU",
            "valu_wave_instructions_per_launch": cnt["valu_wave_instructions_per_launch"],
            "valu_lane_ops_per_s": lane_ops,
            "valu_issue_frac": lane_ops / PEAK_VALU_LANE_OPS,
            "valu_busy_frac_counter": cnt.get("valu_busy_frac")})
    extras = {"update_kernel_avg_ms": u_ms / max(u_n, 1)}

    if not args.no_extras:
        # precision against the FP64 direct sum (the reference's arithmetic) at the
        # final state, and the FP64 kernel's speed
        xs = x.cpu().numpy()
        vs = v.cpu().numpy()
        acc64 = torch.empty((n, 3), dtype=torch.float64, device=dev)
        acc32 = torch.empty((n, 3), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.timing_enable(True)
        ctx.direct_dev(x.data_ptr(), m.data_ptr(), n, EPS_DIRECT, FP64, acc64.data_ptr())
        ctx.direct_dev(x.data_ptr(), m.data_ptr(), n, EPS_DIRECT, FP32, acc32.data_ptr())
        ctx.sync()
        f64_ms, _ = ctx.timing_read(_native.T_DIRECT_F64)
        f32_ms, _ = ctx.timing_read(_native.T_DIRECT_F32)
        ctx.timing_enable(False)
        a64 = acc64.cpu().numpy()
        a32 = acc32.cpu().numpy()
        err = np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1)
        ke1, pe1, _ = ctx.energy(xs, vs, ic["mass"], EPS_DIRECT)
        G = G_KMS2_KPC_MSUN
        e0, e1 = ke0 + G * pe0, ke1 + G * pe1
        extras.update({
            "force_rel_err_median": float(np.median(err)),
            "force_rel_err_p99": float(np.percentile(err, 99)),
            "fp64_direct_kernel_ms": f64_ms,
            "fp32_direct_kernel_ms": f32_ms,
            "fp32_speedup_over_fp64": f64_ms / f32_ms if f32_ms > 0 else None,
            "energy_rel_drift": abs(e1 - e0) / abs(e0),
            "energy_drift_steps": args.warmup + args.steps,
        })
        # the same steps from the same initial state with the plain FP64 kernel
        x64 = torch.tensor(ic["pos"], dtype=torch.float64, device=dev).contiguous()
        v64 = torch.tensor(ic["vel"], dtype=torch.float64, device=dev).contiguous()
        torch.cuda.synchronize()
        ctx.run_dev(x64.data_ptr(), v64.data_ptr(), m.data_ptr(), n, DT_DIRECT, EPS_DIRECT,
                    args.warmup + args.steps, precision=FP64)
        ctx.sync()
        ke2, pe2, _ = ctx.energy(x64.cpu().numpy(), v64.cpu().numpy(), ic["mass"], EPS_DIRECT)
        e2 = ke2 + G * pe2
        dx = np.linalg.norm(x64.cpu().numpy() - xs, axis=1)
        extras.update({
            "energy_rel_drift_fp64_kernel": abs(e2 - e0) / abs(e0),
            "final_pos_diff_vs_fp64_kernel_kpc_median": float(np.median(dx)),
            "final_pos_diff_vs_fp64_kernel_kpc_max": float(dx.max()),
        })
        del x64, v64
    if not args.no_extras:
        log("bench: N=2000 FP64 run of 400 steps")
        extras["plummer2000_fp64_run"] = bench_plummer2000_run(ctx)
    if not args.no_extras and not args.no_tree:
        log("bench: tree step at N=1M")
        extras["tree"] = bench_tree(ctx, torch, args.tree_steps, 1)
    if not args.no_extras and not args.no_adjoint:
        log("bench: adjoint at N=16384")
        extras["adjoint"] = bench_adjoint(ctx, torch)
    if not args.no_extras and not args.no_ensemble:
        log("bench: ensemble of 128 spheres")
        extras["ensemble"] = bench_ensemble(ctx, torch, args.ensemble_steps)
    result["extras"] = extras
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        host = host_description()  # before the baselines pin this process to one core
        result["cpu_baseline"] = cpu_baseline()
        result["cpu_baseline"]["host"] = host
        if not args.no_extras:
            log("bench: cpu baseline, reference direct sum at N=2000")
            c1 = cpu_plummer2000_baseline()
            if c1 is not None:
                result["cpu_baseline"]["plummer2000_fp64_run"] = c1
                if "plummer2000_fp64_run" in extras:
                    extras["plummer2000_fp64_run"]["speedup_over_cpu_reference"] = (
                        c1["value"] / extras["plummer2000_fp64_run"]["seconds"])
        if not args.no_extras and not args.no_tree:
            log("bench: cpu baseline, reference tree at N=1M")
            tb = cpu_tree_baseline()
            if tb is not None:
                result["cpu_baseline"]["tree"] = tb
                if "tree" in extras:
                    extras["tree"]["speedup_over_cpu_reference"] = (
                        extras["tree"]["steps_per_s"] / tb["value"])
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
