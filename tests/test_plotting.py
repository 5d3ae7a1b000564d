"""Simulation.plot_particles and movie_particles (gravhopper.py:958-1162) on the CPU.

The history arrays are filled on the host (a uniformly moving cloud), so no GPU run is
needed: the tests check what is drawn (offsets, labels, title, limits) and the frames
of the animation."""
import numpy as np
import pytest

matplotlib = pytest.importorskip("matplotlib")
matplotlib.use("Agg")
import matplotlib.pyplot as plt  # noqa: E402

from gravhopper_amd import Simulation  # noqa: E402

NSNAP, NP, DT = 11, 50, 0.5


@pytest.fixture
def sim():
    rng = np.random.default_rng(0)
    pos0 = rng.normal(size=(NP, 3))
    vel = rng.normal(scale=100.0, size=(NP, 3))
    s = Simulation(dt=DT, eps=0.01, algorithm='direct')
    s.add_IC({'pos': pos0, 'vel': vel, 'mass': np.ones(NP)})
    s.init_run(NSNAP - 1)
    t = np.arange(NSNAP) * DT
    s._pos[:] = pos0[None] + vel[None] * t[:, None, None] * 1.022712165045695e-03
    s._vel[:] = vel[None]
    s._times[:] = t
    s.timestep = NSNAP - 1
    yield s
    plt.close('all')


def test_plot_final_positions(sim):
    fig, ax = plt.subplots()
    sc = sim.plot_particles(ax=ax, xlim=(-3, 3), ylim=(-2, 4), color='k')
    np.testing.assert_array_equal(sc.get_offsets(), sim.positions[-1][:, :2])
    assert ax.get_xlabel() == '$x$ (kpc)' and ax.get_ylabel() == '$y$ (kpc)'
    assert ax.get_title() == '5.0'
    assert ax.get_xlim() == (-3, 3) and ax.get_ylim() == (-2, 4)
    assert sc.get_sizes()[0] == 0.2


def test_plot_options(sim):
    fig, ax = plt.subplots()
    sc = sim.plot_particles(parm='vel', coords='zx', snap='IC', ax=ax, particle_range=(5, 20),
                            timeformat='t={0:.2f}', s=3.0)
    np.testing.assert_array_equal(sc.get_offsets(), sim.velocities[0][5:20][:, [2, 0]])
    assert ax.get_xlabel() == '$v_z$ (km/s)' and ax.get_ylabel() == '$v_x$ (km/s)'
    assert ax.get_title() == 't=0.00'
    assert sc.get_sizes()[0] == 3.0
    fig, ax = plt.subplots()
    sc = sim.plot_particles(coords='yz', snap=4, ax=ax, nolabels=True, timeformat=False)
    np.testing.assert_array_equal(sc.get_offsets(), sim.positions[4][:, 1:])
    assert ax.get_xlabel() == '' and ax.get_title() == ''


def test_plot_default_axis(sim):
    plt.figure()
    sc = sim.plot_particles()
    assert sc.axes.get_aspect() == 1.0


def test_plot_errors(sim):
    fig, ax = plt.subplots()
    with pytest.raises(ValueError, match="parm"):
        sim.plot_particles(parm='acc', ax=ax)
    with pytest.raises(ValueError, match="length 2"):
        sim.plot_particles(coords='xyz', ax=ax)
    with pytest.raises(ValueError, match="x, y, or z"):
        sim.plot_particles(coords='xw', ax=ax)


def test_movie_frames_and_limits(sim, tmp_path):
    from PIL import Image
    fig, ax = plt.subplots()
    out = tmp_path / "movie.gif"
    sim.movie_particles(str(out), fps=10, ax=ax, skip=2, coords='xz')
    xz = sim.positions[:, :, [0, 2]]
    np.testing.assert_allclose(ax.get_xlim(), (xz[..., 0].min(), xz[..., 0].max()))
    np.testing.assert_allclose(ax.get_ylim(), (xz[..., 1].min(), xz[..., 1].max()))
    with Image.open(out) as im:
        assert im.n_frames == NSNAP // 2
    # the animation leaves the plot at its last frame: snapshot (NSNAP // 2 - 1) * 2
    last = (NSNAP // 2 - 1) * 2
    sc = ax.collections[0]
    np.testing.assert_array_equal(sc.get_offsets(), xz[last])
    assert ax.get_title() == '{0:.1f}'.format(sim.times[last])
    assert plt.fignum_exists(fig.number)


def test_movie_closes_its_own_figure(sim, tmp_path):
    before = set(plt.get_fignums())
    sim.movie_particles(str(tmp_path / "m.gif"), xlim=(-5, 5), ylim=(-5, 5))
    assert set(plt.get_fignums()) == before
