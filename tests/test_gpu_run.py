"""Device-resident leapfrog (gh_run) and the Simulation drop-in on the GPU.

The FP64 direct run is bit-identical to the oracle's DKD restatement of
gravhopper.py:405-416 (so 400 steps of the README Plummer run agree to the last bit),
host-driven and device-resident Simulation runs agree bit for bit, and continued runs
equal one long run.
"""
import numpy as np
import pytest

from conftest import assert_bitwise, rel_err
from gravhopper_amd import Simulation, _native, workloads

pytestmark = pytest.mark.gpu


def test_run_fp64_bitwise_vs_oracle_short(ctx, oracle):
    ic = workloads.plummer(500, seed=31)
    x, v, hp, hv = ctx.run(ic["pos"], ic["vel"], ic["mass"], 0.005, 5e-5, 10, history=True)
    xo, vo, hpo, hvo = oracle.dkd_direct(ic["pos"], ic["vel"], ic["mass"], 0.005, 5e-5, 10,
                                         history=True)
    assert_bitwise(x, xo, "pos")
    assert_bitwise(v, vo, "vel")
    assert_bitwise(hp, hpo, "hist pos")
    assert_bitwise(hv, hvo, "hist vel")


def test_run_plummer_n2000_400_steps_bitwise(ctx, oracle):
    # README example: Plummer N=2000, b=1 pc, 1e6 Msun, eps=0.05 pc, dt=0.005 Myr, 400 steps
    ic = workloads.plummer(2000, b=1e-3, mass=1e6, seed=20260001)
    x, v = ctx.run(ic["pos"], ic["vel"], ic["mass"], 0.005, 5e-5, 400)
    xo, vo = oracle.dkd_direct(ic["pos"], ic["vel"], ic["mass"], 0.005, 5e-5, 400)
    assert_bitwise(x, xo, "final pos")
    assert_bitwise(v, vo, "final vel")


def test_simulation_device_vs_host_loop_bitwise(ctx):
    ic = workloads.plummer(300, seed=32)
    a = Simulation(dt=0.005, eps=5e-5, algorithm='direct')
    a.add_IC(ic)
    a.run(6)
    b = Simulation(dt=0.005, eps=5e-5, algorithm='direct', device_loop=False)
    b.add_IC(ic)
    # a zero external force forces the host-driven loop (gravhopper.py:317-320)
    b.add_external_force(lambda pos, args: np.zeros_like(pos))
    b.run(6)
    assert_bitwise(a.positions, b.positions, "positions")
    assert_bitwise(a.velocities, b.velocities, "velocities")
    np.testing.assert_array_equal(a.times, b.times)


def _add_external_forces(sim, velocity=True):
    # position-, time- and velocity-dependent accelerations in km/s/Myr
    sim.add_external_force(lambda pos, k: -k * pos, 3.0)
    sim.add_external_timedependent_force(
        lambda pos, t, a: a * np.sin(0.37 * t) * np.ones_like(pos) + 0.0 * pos, 0.05)
    if velocity:
        sim.add_external_velocitydependent_force(lambda pos, vel, g: -g * vel, 0.02)


@pytest.mark.parametrize("algorithm", ["direct", "tree"])
@pytest.mark.parametrize("velocity", [False, True])
def test_external_forces_device_loop_equals_host_loop(ctx, algorithm, velocity):
    """gh_run_ext keeps the state on the GPU and calls the callbacks per step; the
    trajectory must equal the host-driven loop of perform_timestep bit for bit, also
    across a continued run (kick times from the running time sum)."""
    ic = workloads.plummer(400, seed=36)
    sims = []
    for device_loop in (True, False):
        s = Simulation(dt=0.004, eps=5e-5, algorithm=algorithm, device_loop=device_loop)
        s.add_IC(ic)
        _add_external_forces(s, velocity)
        s.run(4)
        s.run(3)
        sims.append(s)
    a, b = sims
    assert_bitwise(a.positions, b.positions, "positions")
    assert_bitwise(a.velocities, b.velocities, "velocities")
    np.testing.assert_array_equal(a.times, b.times)
    # the external forces did act
    c = Simulation(dt=0.004, eps=5e-5, algorithm=algorithm)
    c.add_IC(ic)
    c.run(7)
    assert not np.array_equal(a.velocities[-1], c.velocities[-1])


def test_external_callback_error_propagates(ctx):
    ic = workloads.plummer(200, seed=37)
    calls = []

    def bad(pos, args):
        calls.append(1)
        if len(calls) == 3:
            raise ValueError("external force failed")
        return 0.0 * pos

    sim = Simulation(dt=0.004, eps=5e-5, algorithm='direct')
    sim.add_IC(ic)
    sim.add_external_force(bad)
    with pytest.raises(ValueError, match="external force failed"):
        sim.run(5)
    assert len(calls) == 3
    # the context stays usable
    x, v = ctx.run(ic["pos"], ic["vel"], ic["mass"], 0.004, 5e-5, 2)
    assert np.all(np.isfinite(x))


def test_simulation_continue_equals_long_run(ctx):
    ic = workloads.plummer(256, seed=33)
    a = Simulation(dt=0.002, eps=1e-4, algorithm='direct')
    a.add_IC(ic)
    a.run(5)
    a.run(7)
    b = Simulation(dt=0.002, eps=1e-4, algorithm='direct')
    b.add_IC(ic)
    b.run(12)
    assert a.positions.shape == (13, 256, 3)
    assert_bitwise(a.positions, b.positions)
    assert_bitwise(a.velocities, b.velocities)


def test_simulation_external_forces(ctx):
    # harmonic external potential on a single particle: DKD closed form
    # cos(Omega dt) = 1 - (omega dt)^2 / 2 (no N-body force for Np = 1)
    omega = 0.05  # 1/Myr
    sim = Simulation(dt=1.0, eps=0.1, algorithm='direct')
    sim.add_IC({'pos': np.array([[1.0, 0.0, 0.0]]), 'vel': np.zeros((1, 3)), 'mass': np.ones(1)})
    # acceleration in km/s/Myr for x in kpc: -omega^2 x converted from kpc/Myr^2
    sim.add_external_force(lambda pos, args: -omega ** 2 * pos / 1.022712165045695e-03)
    sim.run(200)
    x = sim.positions[:, 0, 0]
    big_omega = np.arccos(1 - (omega * 1.0) ** 2 / 2)
    t = np.arange(201)
    np.testing.assert_allclose(x, np.cos(big_omega * t) / np.cos(big_omega / 2) *
                               np.cos(big_omega / 2), atol=2e-3)


def test_dkd_map_powers_harmonic(ctx):
    # One particle (no N-body force) in a harmonic external field a = -K x, started
    # off the axis with a velocity.  With u = v D (kpc/Myr) and w^2 = K D, one DKD step
    # (gravhopper.py:405-416) is the linear map
    #   M = [[1 - (wh)^2/2, h (1 - (wh)^2/4)], [-w^2 h, 1 - (wh)^2/2]]
    # per coordinate, det M = 1, so M^n = (sin(n p) M - sin((n-1) p) I) / sin p with
    # cos p = 1 - (wh)^2/2.  Every snapshot must equal (x0, u0) M^n to 1e-12 of the
    # amplitude: a kick and drift in the other order, a dropped half-drift or the
    # force taken at x instead of the half-step position all fail by far more.
    D = 1.022712165045695e-03
    w, h, n = 0.05, 1.0, 300
    K = w * w / D
    x0 = np.array([[1.0, -0.5, 0.3]])
    v0 = np.array([[30.0, 10.0, -20.0]])
    sim = Simulation(dt=h, eps=0.1, algorithm='direct')
    sim.add_IC({'pos': x0, 'vel': v0, 'mass': np.ones(1)})
    sim.add_external_force(lambda pos, args: -K * pos)
    sim.run(n)
    x, v = sim.positions[:, 0, :], sim.velocities[:, 0, :]
    M = np.array([[1 - (w * h) ** 2 / 2, h * (1 - (w * h) ** 2 / 4)],
                  [-w * w * h, 1 - (w * h) ** 2 / 2]])
    p = np.arccos(M[0, 0])
    steps = np.arange(n + 1)
    s1 = np.sin(steps * p)[:, None, None]
    s0 = np.sin((steps - 1) * p)[:, None, None]
    Mn = (s1 * M - s0 * np.eye(2)) / np.sin(p)                   # (n+1, 2, 2)
    z0 = np.stack([x0[0], v0[0] * D])                            # (2, 3): x and u rows
    z = np.einsum("nij,jk->nik", Mn, z0)                         # (n+1, 2, 3)
    amp = np.sqrt(z0[0] ** 2 + (z0[1] / w) ** 2)                 # per coordinate
    assert np.all(np.abs(x - z[:, 0]) <= 1e-12 * amp), np.abs(x - z[:, 0]).max()
    assert np.all(np.abs(v * D - z[:, 1]) <= 1e-12 * amp * w), np.abs(v * D - z[:, 1]).max()


def test_simulation_fp32_close_to_fp64(ctx):
    ic = workloads.plummer(4096, seed=34)
    res = {}
    for prec in ('fp64', 'fp32'):
        s = Simulation(dt=0.001, eps=1e-5, algorithm='direct', precision=prec)
        s.add_IC(ic)
        s.run(5)
        res[prec] = s.positions[-1]
    disp = np.linalg.norm(res['fp64'] - ic['pos'], axis=1)
    diff = np.linalg.norm(res['fp32'] - res['fp64'], axis=1)
    assert np.median(diff / disp) < 1e-5


@pytest.mark.parametrize("nsteps", [7, 8])
@pytest.mark.parametrize("prec", [_native.GH_FP64, _native.GH_FP32])
def test_run_graph_replay_equals_eager(ctx, prec, nsteps):
    """Direct-sum runs replay two captured steps as a HIP graph (8 steps: an eager first
    step, three replays and an eager last step); every replayed step's snapshot must
    equal the eagerly launched run, over several calls with host work and
    synchronisation in between."""
    ic = workloads.plummer(20000 if prec == _native.GH_FP32 else 1500, seed=35)
    x, v = ic["pos"], ic["vel"]
    xe, ve = x, v
    for call in range(3):
        ctx.set_tuning(_native.TUNE_RUN_GRAPH, 0)
        x, v, hp, hv = ctx.run(x, v, ic["mass"], 0.002, 1e-5, nsteps, precision=prec,
                               history=True)
        ctx.set_tuning(_native.TUNE_RUN_GRAPH, -1)
        xe, ve, hpe, hve = ctx.run(xe, ve, ic["mass"], 0.002, 1e-5, nsteps, precision=prec,
                                   history=True)
        ctx.set_tuning(_native.TUNE_RUN_GRAPH, 0)
        assert_bitwise(hp, hpe, "call %d history pos" % call)
        assert_bitwise(hv, hve, "call %d history vel" % call)
        assert_bitwise(x, xe)
        assert_bitwise(v, ve)


def test_run_history_slices_equal_eager(ctx):
    """A replayed run copies its history out in slices while later steps run; 101 steps
    give 13 slices and an odd last step.  Every snapshot must equal the eager run's."""
    ic = workloads.plummer(3000, seed=36)
    try:
        ctx.set_tuning(_native.TUNE_RUN_GRAPH, -1)
        xe, ve, hpe, hve = ctx.run(ic["pos"], ic["vel"], ic["mass"], 0.003, 2e-5, 101,
                                   history=True)
    finally:
        ctx.set_tuning(_native.TUNE_RUN_GRAPH, 0)
    x, v, hp, hv = ctx.run(ic["pos"], ic["vel"], ic["mass"], 0.003, 2e-5, 101, history=True)
    assert_bitwise(hp, hpe, "history pos")
    assert_bitwise(hv, hve, "history vel")
    assert_bitwise(x, xe, "pos")
    assert_bitwise(v, ve, "vel")
    assert_bitwise(hp[-1], x, "last snapshot")


def test_tree_run_history_slices_equal_chained_steps(ctx):
    """Eager (tree) runs copy their history in slices of about 16 MiB, each once the
    next step is queued; at N=200000 a slice is one step.  The snapshots must equal
    those of single-step runs chained from each other."""
    g = workloads.galaxy(n_halo=100000, n_disk=100000, seed=37)
    args = (g["mass"], 0.1, 0.05)
    x, v, hp, hv = ctx.run(g["pos"], g["vel"], *args, 5, algorithm=_native.GH_TREE,
                           precision=_native.GH_FP32, history=True)
    xc, vc = g["pos"], g["vel"]
    for k in range(5):
        xc, vc = ctx.run(xc, vc, *args, 1, algorithm=_native.GH_TREE,
                         precision=_native.GH_FP32)
        assert_bitwise(hp[k], xc, "step %d pos" % k)
        assert_bitwise(hv[k], vc, "step %d vel" % k)
    assert_bitwise(x, xc, "final pos")
    assert_bitwise(v, vc, "final vel")


def test_external_run_history_slices_equal_chained_steps(ctx):
    """Runs with external forces copy their history in slices of about 16 MiB while the
    next step's force runs; at N=200000 a slice is one step.  The snapshots must equal
    those of single-step runs chained from each other."""
    g = workloads.galaxy(n_halo=100000, n_disk=100000, seed=38)
    dt = 0.1

    def external(xh, v, t):
        # a harmonic pull towards a centre that moves with time
        return -1e-3 * (xh - np.array([0.01 * t, 0.0, 0.0]))

    kw = dict(algorithm=_native.GH_TREE, precision=_native.GH_FP32)
    x, v, hp, hv = ctx.run_ext(g["pos"], g["vel"], g["mass"], dt, 0.05, 4, external, t0=2.0,
                               history=True, **kw)
    xc, vc, t = g["pos"], g["vel"], 2.0
    for k in range(4):
        xc, vc = ctx.run_ext(xc, vc, g["mass"], dt, 0.05, 1, external, t0=t, **kw)
        t = t + dt  # the time advances as in the run: t + dt per step
        assert_bitwise(hp[k], xc, "step %d pos" % k)
        assert_bitwise(hv[k], vc, "step %d vel" % k)
    assert_bitwise(x, xc, "final pos")
    assert_bitwise(v, vc, "final vel")
