"""Drop-in for `gravhopper.Simulation`: the N-body driver of the hot path.

Same public surface as gravhopper/gravhopper.py:91-473, :853-904 and :958-1162
(constructor, set_/get_ dt, eps and algorithm, add_IC, run, init_run, reset, snap,
current_snap, prev_snap, perform_timestep, calculate_acceleration,
calculate_extra_acceleration, the add_external_*force registrations, plot_particles and
movie_particles), with the same drift-kick-drift leapfrog.

What changes is where the steps run:
  - without external forces, run(N) hands the whole loop to the GPU (gh_run): the
    state stays in HBM, every step is one force evaluation plus one fused update
    kernel, and the (N+1, Np, 3) history is filled from device snapshots;
  - with external forces (host Python callbacks), the state still stays on the GPU
    (gh_run_ext): each step sends the half-step positions (and, for velocity-dependent
    forces, the pre-kick velocities) to the host, calls the callbacks there at the
    reference's kick time while the GPU computes the N-body force, and sends the
    summed external acceleration back for the kick (gravhopper.py:405-473);
  - with device_loop=False, external-force runs use a host loop instead: each step
    calls the GPU force and does the leapfrog in numpy, as the reference does.
All paths give bit-identical trajectories for the same force.

Units: internally kpc, km/s, Msun, Myr (gravhopper.py:150-154).  With astropy
installed, Quantities are accepted and the history arrays are Quantities; without it
plain numbers are taken in those units.  The galpy/gala/agama potential wrappers of
the reference are not provided (those packages are not part of this engine); plain
callables are.
"""
import numpy as np

from . import _native
from . import jbgrav
from .units import (ACCEL_SCALE, DRIFT_SCALE, HAVE_ASTROPY, check_dimension, is_quantity,
                    u, value_in)


class GravHopperException(Exception):
    """Parent class for all error exceptions (gravhopper.py:68-70)."""
    pass


class UninitializedSimulationException(GravHopperException):
    """Running a simulation without any initial conditions (gravhopper.py:72-74)."""
    pass


class ICException(GravHopperException):
    """Initial conditions that don't make sense (gravhopper.py:76-79)."""

    def __init__(self, msg):
        super().__init__(msg)
        print(msg)


class UnknownAlgorithmException(GravHopperException):
    """Unknown N-body algorithm name (gravhopper.py:81-83)."""
    pass


class ExternalPackageException(GravHopperException):
    """A pynbody/galpy/gala/agama feature was requested (gravhopper.py:85-88)."""

    def __init__(self, msg):
        super().__init__(msg)
        print(msg)


def _default_dt():
    return 1.0 * u.Myr if HAVE_ASTROPY else 1.0


def _default_eps():
    return 100.0 * u.pc if HAVE_ASTROPY else 0.1


class Simulation(object):
    """Main class for N-body simulation (gravhopper.py:91-114).

    Attributes
    ----------
    Np, Nsnap : int
    positions, velocities : (Nsnap, Np, 3) arrays (kpc, km/s)
    masses : (Np,) array (Msun)
    times : (Nsnap,) array (Myr)

    Extra keyword arguments (defaults reproduce the reference):
    precision : 'fp64' (reference arithmetic) or 'fp32' (FP32 pair forces,
        compensated sums, FP64 state)
    theta : opening angle of the tree (the reference always uses 0.7,
        gravhopper.py:446 with jbgrav.py:52)
    autotune : before the first FP32 run, time the candidate kernel shapes on this
        simulation's particles (gravhopper_amd.autotune.tune_for_run: the self-field
        kernel for 'direct', the walk's group size for 'tree') and use the fastest
        for this simulation's runs (FP32 rounding may differ between shapes; FP64
        runs are unaffected)
    device_loop : with external forces, keep the state on the GPU and call the
        external-force callbacks once per step from the library's loop (True), or
        step in a host loop as the reference does (False); same trajectories
    """

    def __init__(self, dt=None, eps=None, algorithm='tree', precision='fp64', theta=0.7,
                 device_loop=True, autotune=False):
        self.device_loop = bool(device_loop)
        self.autotune = bool(autotune)
        self._tuned = None  # (algorithm, Np, tuning key, value) once tuned
        self.ICarrays = False
        self.Np = 0
        self.Nsnap = 0
        self.timestep = 0
        self.running = False
        self.positions = None
        self.velocities = None
        self.masses = None
        self.times = None
        self.extra_force_functions = []
        self.extra_timedependent_force_functions = []
        self.extra_velocitydependent_force_functions = []
        if HAVE_ASTROPY:  # pragma: no cover
            self.lenunit = u.kpc
            self.velunit = u.km / u.s
            self.massunit = u.Msun
            self.timeunit = u.Myr
            self.accelunit = self.velunit / self.timeunit
        self.params = {}
        self.set_dt(_default_dt() if dt is None else dt)
        self.set_eps(_default_eps() if eps is None else eps)
        self.set_algorithm(algorithm)
        self.set_precision(precision)
        self.theta = float(theta)
        self._plot_parms = None

    # -- parameters (gravhopper.py:166-290) ------------------------------------
    def set_dt(self, dt):
        """Sets the time step; ValueError unless it has dimensions of time."""
        check_dimension(dt, 'Myr', "dt must have dimensions of time.")
        self.params['dt'] = dt

    def get_dt(self):
        return self.params['dt']

    def set_eps(self, eps):
        """Sets the softening length; ValueError unless it has dimensions of length."""
        check_dimension(eps, 'kpc', "eps must have dimensions of length.")
        self.params['eps'] = eps

    def get_eps(self):
        return self.params['eps']

    def set_algorithm(self, algorithm):
        """'tree' or 'direct'; ValueError otherwise."""
        if algorithm in ('tree', 'direct'):
            self.params['algorithm'] = algorithm
        else:
            raise ValueError("algorithm must be 'tree' or 'direct'.")

    def get_algorithm(self):
        return self.params['algorithm']

    def set_precision(self, precision):
        """'fp64' or 'fp32'; ValueError otherwise."""
        if precision not in _native.PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'fp32'.")
        self.params['precision'] = precision

    def get_precision(self):
        return self.params['precision']

    # -- internal-unit views -----------------------------------------------------
    def _dt(self):
        return float(value_in(self.params['dt'], 'Myr'))

    def _eps(self):
        return float(value_in(self.params['eps'], 'kpc'))

    def _quantity(self, arr, unit):
        if HAVE_ASTROPY:  # pragma: no cover
            return arr << unit
        return arr

    def _raw(self, x, unit_name):
        return np.asarray(value_in(x, unit_name), dtype=np.float64)

    # -- running (gravhopper.py:293-356) -----------------------------------------
    def run(self, N=1):
        """Runs N time steps, initialising the run first or continuing from the last
        snapshot."""
        N = int(N)
        if self.running is False:
            self.init_run(N)
        else:
            self.Nsnap += N
            self._pos = np.concatenate((self._pos, np.zeros((N, self.Np, 3))), axis=0)
            self._vel = np.concatenate((self._vel, np.zeros((N, self.Np, 3))), axis=0)
            self._times = np.concatenate((self._times, np.zeros(N)), axis=0)
            self._expose()

        if N <= 0:
            return
        ctx = _native.default_context()
        tuned = self._tuning_for_run(ctx)
        if tuned is not None:
            saved = ctx.get_tuning(tuned[0])
            ctx.set_tuning(*tuned)
        try:
            if not self._has_external_forces():
                self._run_on_device(N)
            elif self.device_loop:
                self._run_on_device_external(N)
            else:
                for _ in range(self.timestep, self.timestep + N):
                    self.timestep += 1
                    self.perform_timestep()
                    self._times[self.timestep] = self._times[self.timestep - 1] + self._dt()
        finally:
            if tuned is not None:
                ctx.set_tuning(tuned[0], saved)

    def _tuning_for_run(self, ctx):
        """(tuning key, value) this simulation's runs use, from the autotuner, or None."""
        if not self.autotune or self.params['precision'] != 'fp32':
            return None
        algo = self.params['algorithm']
        if self._tuned is None or self._tuned[:2] != (algo, self.Np):
            from . import autotune
            t = self.timestep
            saved = {k: ctx.get_tuning(k) for k in (_native.TUNE_DIRECT_SELF,
                                                    _native.TUNE_WALK_GROUP)}
            try:
                kv = autotune.tune_for_run(ctx, self._pos[t], self._mass, self._eps(), algo,
                                           'fp32', theta=self.theta)
            finally:
                for k, v in saved.items():
                    ctx.set_tuning(k, v)
            self._tuned = (algo, self.Np) + tuple(kv)
        return self._tuned[2:]

    def _has_external_forces(self):
        return bool(self.extra_force_functions or self.extra_timedependent_force_functions
                    or self.extra_velocitydependent_force_functions)

    def _run_on_device(self, N):
        t0 = self.timestep
        algo = _native.GH_DIRECT if self.params['algorithm'] == 'direct' else _native.GH_TREE
        prec = _native.PRECISIONS[self.params['precision']]
        # the snapshots go straight into the history arrays (rows t0+1 .. t0+N)
        _native.default_context().run(
            self._pos[t0], self._vel[t0], self._mass, self._dt(), self._eps(), N,
            algorithm=algo, theta=self.theta, precision=prec,
            hist_out=(self._pos[t0 + 1:t0 + N + 1], self._vel[t0 + 1:t0 + N + 1]))
        dt = self._dt()
        for k in range(t0 + 1, t0 + N + 1):
            self._times[k] = self._times[k - 1] + dt
        self.timestep = t0 + N

    def _run_on_device_external(self, N):
        """N steps of gh_run_ext; the callbacks get what perform_timestep gives them:
        half-step positions, kick time t + dt/2, pre-kick velocities."""
        t0 = self.timestep
        algo = _native.GH_DIRECT if self.params['algorithm'] == 'direct' else _native.GH_TREE
        prec = _native.PRECISIONS[self.params['precision']]
        pass_v = bool(self.extra_velocitydependent_force_functions)

        def external(xh, v, t):
            return self._extra_raw(xh, xh, t, v)

        _native.default_context().run_ext(
            self._pos[t0], self._vel[t0], self._mass, self._dt(), self._eps(), N, external,
            t0=self._times[t0], pass_velocity=pass_v, algorithm=algo, theta=self.theta,
            precision=prec,
            hist_out=(self._pos[t0 + 1:t0 + N + 1], self._vel[t0 + 1:t0 + N + 1]))
        dt = self._dt()
        for k in range(t0 + 1, t0 + N + 1):
            self._times[k] = self._times[k - 1] + dt
        self.timestep = t0 + N

    def init_run(self, Nsnap=None):
        """Allocates the history arrays and copies the ICs into snapshot 0."""
        if self.ICarrays is False:
            raise UninitializedSimulationException
        self.Nsnap = Nsnap + 1
        self.Np = len(self.ICarrays['pos'])
        self._pos = np.zeros((self.Nsnap, self.Np, 3))
        self._vel = np.zeros((self.Nsnap, self.Np, 3))
        self._mass = np.zeros(self.Np)
        self._times = np.zeros(self.Nsnap)
        self._pos[0, :, :] = self._raw(self.ICarrays['pos'], 'kpc').reshape(self.Np, 3)
        self._vel[0, :, :] = self._raw(self.ICarrays['vel'], 'km/s').reshape(self.Np, 3)
        self._mass[:] = self._raw(self.ICarrays['mass'], 'Msun').reshape(self.Np)
        self._expose()
        self.running = True

    def _expose(self):
        if HAVE_ASTROPY:  # pragma: no cover
            self.positions = self._pos << u.kpc
            self.velocities = self._vel << (u.km / u.s)
            self.masses = self._mass << u.Msun
            self.times = self._times << u.Myr
        else:
            self.positions = self._pos
            self.velocities = self._vel
            self.masses = self._mass
            self.times = self._times

    def reset(self):
        """Back to the state before init_run(); ICs and external forces are kept."""
        self.running = False
        self.positions = None
        self.velocities = None
        self.masses = None
        self.times = None
        self.Nsnap = 0
        self.timestep = 0

    # -- snapshots (gravhopper.py:361-402) ---------------------------------------
    def snap(self, step):
        return {'pos': self.positions[step, :, :], 'vel': self.velocities[step, :, :],
                'mass': self.masses[:]}

    def current_snap(self):
        return self.snap(self.timestep)

    def prev_snap(self):
        return self.snap(self.timestep - 1)

    # -- one host-driven step (gravhopper.py:405-473) -----------------------------
    def perform_timestep(self):
        """One drift-kick-drift step with the force on the GPU and any external forces
        on the host."""
        dt = self._dt()
        t = self.timestep
        self._pos[t] = self._pos[t - 1] + 0.5 * self._vel[t - 1] * dt * DRIFT_SCALE
        kick_time = self._times[t - 1] + 0.5 * dt
        accelerations = self._calculate_acceleration_raw(kick_time)
        self._vel[t] = self._vel[t - 1] + accelerations * dt
        self._pos[t] += 0.5 * self._vel[t] * dt * DRIFT_SCALE

    def _nbody_raw(self):
        """N-body acceleration at the current snapshot, km/s/Myr, plain array."""
        t = self.timestep
        if self.Np > 1:
            snap = {'pos': self._pos[t], 'mass': self._mass}
            prec = self.params['precision']
            if self.params['algorithm'] == 'direct':
                g = jbgrav.direct_summation(snap, self._eps(), precision=prec)
            elif self.params['algorithm'] == 'tree':
                g = jbgrav.tree_force(snap, self._eps(), theta=self.theta, precision=prec)
            else:
                raise UnknownAlgorithmException()
            return np.asarray(value_in(g, 'km/s/Myr'))
        return np.zeros((self.Np, 3))

    def _calculate_acceleration_raw(self, time):
        nbody = self._nbody_raw()
        t = self.timestep
        extra = self._extra_raw(self._pos[t], nbody, time, self._vel[t - 1])
        return nbody + extra

    def _extra_raw(self, pos, template, time, vel):
        ext = np.zeros_like(template)
        if HAVE_ASTROPY:  # pragma: no cover
            posq, timeq = pos << u.kpc, time * u.Myr
            velq = None if vel is None else vel << (u.km / u.s)
        else:
            posq, velq, timeq = pos, vel, time
        for fn, args in self.extra_force_functions:
            ext += np.asarray(value_in(fn(posq, args), 'km/s/Myr'))
        for fn, args in self.extra_timedependent_force_functions:
            ext += np.asarray(value_in(fn(posq, timeq, args), 'km/s/Myr'))
        for fn, args in self.extra_velocitydependent_force_functions:
            ext += np.asarray(value_in(fn(posq, velq, args), 'km/s/Myr'))
        return ext

    def calculate_acceleration(self, time=None):
        """N-body plus external acceleration at current_snap() (gravhopper.py:419-459);
        velocity-dependent forces see the previous snapshot's velocities."""
        t_raw = None if time is None else float(value_in(time, 'Myr'))
        acc = self._calculate_acceleration_raw(t_raw)
        return self._quantity(acc, u.km / u.s / u.Myr) if HAVE_ASTROPY else acc

    def calculate_extra_acceleration(self, pos, template_array, time=None, vel=None):
        """Acceleration from the registered external forces only (gravhopper.py:462-473)."""
        return self._extra_raw(self._raw(pos, 'kpc'), np.asarray(value_in(template_array,
                               'km/s/Myr')), None if time is None else
                               float(value_in(time, 'Myr')),
                               None if vel is None else self._raw(vel, 'km/s'))

    # -- external forces (gravhopper.py:603-845, plain callables) -----------------
    def add_external_force(self, fn, args=None):
        """fn(pos, args) -> acceleration; summed with the N-body force."""
        if isinstance(fn, list):
            for item in fn:
                self.add_external_force(item, args)
            return
        self._check_callable(fn)
        self.extra_force_functions.append((fn, args))

    def add_external_timedependent_force(self, fn, args=None):
        """fn(pos, time, args) -> acceleration, evaluated at t + dt/2."""
        if isinstance(fn, list):
            for item in fn:
                self.add_external_timedependent_force(item, args)
            return
        self._check_callable(fn)
        self.extra_timedependent_force_functions.append((fn, args))

    def add_external_velocitydependent_force(self, fn, args=None):
        """fn(pos, vel, args) -> acceleration, with the pre-kick velocities."""
        if isinstance(fn, list):
            for item in fn:
                self.add_external_velocitydependent_force(item, args)
            return
        self._check_callable(fn)
        self.extra_velocitydependent_force_functions.append((fn, args))

    @staticmethod
    def _check_callable(fn):
        if not callable(fn):
            raise ExternalPackageException(
                "External forces must be plain callables; galpy, gala and agama "
                "potential objects are not supported by this engine.")

    # -- plots and movies (gravhopper.py:958-1062, 1065-1088, 1091-1162) -----------
    def _plot_series(self, parm):
        if parm == 'pos':
            return self.positions, 'kpc'
        if parm == 'vel':
            return self.velocities, 'km/s'
        raise ValueError("parm must be 'pos' or 'vel'")

    @staticmethod
    def _in_unit(data, unit, default):
        """Plain values of data in `unit`.  With astropy, any unit data converts to;
        without it, data are plain numbers in the internal unit, so only that unit (or
        None) is accepted."""
        if is_quantity(data):
            return data.to(unit).value
        if unit is not None and str(unit) != default:
            raise ValueError("without astropy, data can only be plotted in %s" % default)
        return np.asarray(data)

    def _plot_xy(self, snapnum, sl=None):
        """(x, y) of the particles in the stored plot settings at snapshot(s) snapnum."""
        pp = self._plot_parms
        data, default = self._plot_series(pp['data_parm'])
        lo, hi = pp['particle_range']
        rows = data[snapnum, lo:hi] if sl is None else data[:, lo:hi]
        x = self._in_unit(rows[..., pp['xindex']], pp['unit'], default)
        y = self._in_unit(rows[..., pp['yindex']], pp['unit'], default)
        return x, y

    def plot_particles(self, parm='pos', coords='xy', snap='final', xlim=None, ylim=None,
                       s=0.2, unit=None, ax=None, timeformat='{0:.1f}', nolabels=False,
                       particle_range=None, **kwargs):
        """Scatter plot of one snapshot's positions ('pos') or velocities ('vel') in the
        projection coords ('xy', 'xz', 'yz', ...).  snap: a snapshot number, 'final'
        or 'IC'.  unit: the unit to plot in (default: that of the data, kpc or km/s).
        particle_range: (first, end) indices to plot.  timeformat: format of the title
        (the snapshot's time), False for none.  Extra keywords go to ax.scatter.
        Returns the PathCollection (gravhopper.py:958-1062)."""
        import matplotlib.pyplot as plt

        data, default = self._plot_series(parm)
        if len(coords) != 2:
            raise ValueError("coords must have length 2")
        if coords[0] not in 'xyz' or coords[1] not in 'xyz':
            raise ValueError("coords characters must be x, y, or z")
        if ax is None:
            ax = plt.subplot(111, aspect=1.0)
        snapnum = self.timestep if snap == 'final' else 0 if snap == 'IC' else snap
        if unit is None:
            unit = data.unit if is_quantity(data) else default
        if particle_range is None:
            particle_range = (0, self.Np)
        self._plot_parms = {'data_parm': parm, 'particle_range': particle_range,
                            'xindex': 'xyz'.index(coords[0]),
                            'yindex': 'xyz'.index(coords[1]), 'unit': unit}
        x, y = self._plot_xy(snapnum)
        output = ax.scatter(x, y, s=s, **kwargs)
        if xlim is not None:
            ax.set_xlim(xlim)
        if ylim is not None:
            ax.set_ylim(ylim)
        if not nolabels:
            pre = '' if parm == 'pos' else 'v_'
            ax.set_xlabel('${0}{1}$ ({2})'.format(pre, coords[0], unit))
            ax.set_ylabel('${0}{1}$ ({2})'.format(pre, coords[1], unit))
        self._plot_particles_settitle(ax, snapnum, timeformat)
        return output

    def _plot_particles_settitle(self, ax, snapnum, timeformat=False):
        """Titles ax with the time of snapshot snapnum in timeformat (False: no title)."""
        if timeformat is not False:
            ax.set_title(timeformat.format(self.times[snapnum]))

    def _plot_particles_setoffsets(self, scatterplot, snapnum):
        """Moves the points of a plot_particles() scatter plot to snapshot snapnum."""
        x, y = self._plot_xy(snapnum)
        scatterplot.set_offsets(np.column_stack([x, y]))

    def movie_particles(self, fname, fps=25, ax=None, skip=None, timeformat='{0:.1f}',
                        *args, **kwargs):
        """Writes an animation of plot_particles() over the snapshots to fname: every
        skip-th snapshot from the ICs, fps frames per second.  Unless xlim / ylim are
        given, the axes span every value the plotted particles take over the run.  With
        ax None a new figure is made and closed afterwards.  Other arguments go to
        plot_particles (gravhopper.py:1091-1162)."""
        import matplotlib.pyplot as plt
        from matplotlib.animation import FuncAnimation

        if ax is None:
            fig = plt.figure()
            ax = fig.add_subplot(111, aspect=1.0)
            close_plot = True
        else:
            fig = ax.get_figure()
            close_plot = False
        skip = 1 if skip is None else skip
        particles = self.plot_particles(*args, ax=ax, snap='IC', timeformat=timeformat,
                                        **kwargs)
        xall, yall = self._plot_xy(None, sl=True)
        if 'xlim' not in kwargs:
            ax.set_xlim([np.min(xall), np.max(xall)])
        if 'ylim' not in kwargs:
            ax.set_ylim([np.min(yall), np.max(yall)])

        def animate(frame):
            snapnum = frame * skip
            self._plot_particles_setoffsets(particles, snapnum)
            self._plot_particles_settitle(ax, snapnum, timeformat)
            return particles

        anim = FuncAnimation(fig, animate, frames=(self.timestep + 1) // skip,
                             interval=1000 / fps)
        anim.save(fname)
        if close_plot:
            plt.close(fig)

    # -- initial conditions (gravhopper.py:848-904) --------------------------------
    def nrows(self, array):
        """Number of rows in an array or scalar."""
        return array.shape[0] if np.ndim(array) > 1 else 1

    def add_IC(self, newIC):
        """Adds particles: a dict with 'pos', 'vel' and 'mass'."""
        if 'pos' not in newIC:
            raise ICException("Missing 'pos' key in initial conditions function.")
        if 'vel' not in newIC:
            raise ICException("Missing 'vel' key in initial conditions function.")
        if 'mass' not in newIC:
            raise ICException("Missing 'mass' key in initial conditions function.")
        npos = self.nrows(newIC['pos'])
        nvel = self.nrows(newIC['vel'])
        nmass = len(newIC['mass'])
        if (npos != nvel) | (npos != nmass):
            raise ICException('Inconsistent number of particles in initial conditions function.')
        if self.ICarrays is False:
            self.ICarrays = {}
            self.ICarrays['pos'] = newIC['pos']
            self.ICarrays['vel'] = newIC['vel']
            self.ICarrays['mass'] = newIC['mass']
        else:
            self.ICarrays['pos'] = np.vstack((self.ICarrays['pos'], newIC['pos']))
            self.ICarrays['vel'] = np.vstack((self.ICarrays['vel'], newIC['vel']))
            self.ICarrays['mass'] = np.hstack((self.ICarrays['mass'], newIC['mass']))
