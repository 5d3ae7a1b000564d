"""ctypes binding of libgravhop.so (include/gravhop.h).

The library is the only compute path: if it is missing, or no HIP device is visible,
every call raises RuntimeError.  There is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GRAVHOP_LIB", os.path.join(HERE, "libgravhop.so"))

GH_FP64 = 0
GH_FP32 = 1
GH_DIRECT = 0
GH_TREE = 1

GH_EARG = -1
GH_EHIP = -2
GH_ENOMEM = -3
GH_ECOINCIDENT = -4
GH_EDEPTH = -5
GH_ECALLBACK = -6

PRECISIONS = {"fp64": GH_FP64, "fp32": GH_FP32}

# names and argument types of every exported entry point of include/gravhop.h
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_I = ctypes.c_int
# gh_external_fn of gh_run_ext
EXTERNAL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.c_double, ctypes.POINTER(ctypes.c_double))
SIGNATURES = {
    "gh_ctx_create": ([_I, ctypes.POINTER(_P)], _I),
    "gh_ctx_destroy": ([_P], None),
    "gh_last_error": ([_P], ctypes.c_char_p),
    "gh_device_count": ([ctypes.POINTER(_I)], _I),
    "gh_sync": ([_P], _I),
    "gh_direct_summation": ([_P, _P, _P, _I64, _D, _I, _P], _I),
    "gh_direct_summation_position": ([_P, _P, _P, _I64, _P, _I64, _D, _I, _P], _I),
    "gh_direct_summation_dev": ([_P, _P, _P, _I64, _P, _I64, _D, _I, _P, _P], _I),
    "gh_tree_force": ([_P, _P, _P, _I64, _D, _D, _I, _P], _I),
    "gh_tree_force_position": ([_P, _P, _P, _I64, _P, _I64, _D, _D, _I, _P], _I),
    "gh_tree_force_dev": ([_P, _P, _P, _I64, _P, _I64, _D, _D, _I, _P, _P], _I),
    "gh_tree_dump": ([_P, _P, _P, _I64, _D, _P, _I64, ctypes.POINTER(_I64)], _I),
    "gh_tree_interactions": ([_P, _P, _P, _I64, _P, _I64, _D, _D, _P, _P], _I),
    "gh_tree_interactions_fp32": ([_P, _P, _P, _I64, _P, _I64, _D, _D, _P, _P], _I),
    "gh_run": ([_P, _P, _P, _P, _I64, _D, _D, _I, _D, _I, _I64, _P, _P], _I),
    "gh_run_dev": ([_P, _P, _P, _P, _I64, _D, _D, _I, _D, _I, _I64, _P, _P, _P], _I),
    "gh_run_ext": ([_P, _P, _P, _P, _I64, _D, _D, _I, _D, _I, _I64, _D, EXTERNAL_FN, _P, _I,
                    _P, _P], _I),
    "gh_timing_enable": ([_P, _I], _I),
    "gh_set_tuning": ([_P, _I, _I], _I),
    "gh_ensemble_create": ([_P, _I, _P, _P, _P, _P, _P, _P, _I, _I, ctypes.POINTER(_P)], _I),
    "gh_ensemble_run": ([_P, _I64, _I], _I),
    "gh_ensemble_get": ([_P, _P, _P], _I),
    "gh_ensemble_info": ([_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                          ctypes.POINTER(_I64)], _I),
    "gh_ensemble_destroy": ([_P], None),
    "gh_adjoint_forward": ([_P, _P, _P, _P, _I64, _D, _D, _I, _I64, _I64, _P, _P], _I),
    "gh_adjoint_backward": ([_P, _P, _P, _I64, _D, _D, _I, _I64, _I64, _P, _P, _P, _P], _I),
    "gh_timing_read": ([_P, _I, ctypes.POINTER(_D), ctypes.POINTER(_I64)], _I),
    "gh_energy": ([_P, _P, _P, _P, _I64, _D, _P, _P, _P], _I),
    "gh_energy_dev_noalloc": ([_P, _P, _P, _I64, _D, _P, _P], _I),
}

# gh_set_tuning keys
TUNE_DIRECT_SELF, TUNE_DIRECT_IPT, TUNE_WALK_GROUP, TUNE_RUN_GRAPH = 0, 1, 2, 3
TUNE_WALK_F64EVAL, TUNE_ALLOC_LIMIT_MB, TUNE_DIRECT_F64_TPW = 4, 5, 6
TUNE_SYM_BAND, TUNE_SORT_DIRECT, TUNE_TREE_SORT, TUNE_WALK_PREFIX = 7, 8, 9, 10

# kernel classes of gh_timing_read
T_DIRECT_F32, T_DIRECT_F64, T_TREE_BUILD, T_TREE_WALK, T_UPDATE, T_VJP = range(6)

NODE_DTYPE = np.dtype([
    ("level", np.int32), ("branch", np.int32), ("leaf", np.int64),
    ("center", np.float64, (3,)), ("size", np.float64), ("mass", np.float64),
    ("com", np.float64, (3,)),
])


class GravhopError(RuntimeError):
    """A failure reported by libgravhop.so; .code is the GH_E* value."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None
_lock = threading.RLock()


def load_library(path=None):
    """Loads and prototypes libgravhop.so (no GPU access)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(
                "libgravhop.so is not built (%s); build it with "
                "`python -m gravhopper_amd._build`" % p)
        L = ctypes.CDLL(p)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        if path is None:
            _lib = L
        return L


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _history_arrays(nsteps, n, history, hist_out):
    """(hist_pos, hist_vel) for a run: new arrays, the caller's (checked), or None."""
    if hist_out is None:
        if not history:
            return None, None
        return np.empty((nsteps, n, 3)), np.empty((nsteps, n, 3))
    hp, hv = hist_out
    for a in (hp, hv):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64
                and a.shape == (nsteps, n, 3) and a.flags.c_contiguous and a.flags.writeable):
            raise ValueError("hist_out arrays must be writable C-contiguous float64 arrays "
                             "of shape (%d, %d, 3)" % (nsteps, n))
    return hp, hv


class Context:
    """One device context: a HIP stream plus the library's device arenas."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.gh_ctx_create(int(device), ctypes.byref(h))
        if rc != 0 or not h.value:
            n = ctypes.c_int(0)
            self.lib.gh_device_count(ctypes.byref(n))
            raise RuntimeError("cannot create a gravhop context on HIP device %d "
                               "(%d device(s) visible, code %d)" % (device, n.value, rc))
        self.handle = h
        self.device = device
        self._tuning = {}  # values set through set_tuning (0 = automatic otherwise)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.gh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            msg = self.lib.gh_last_error(self.handle)
            raise GravhopError(rc, (msg or b"").decode() or "gravhop error %d" % rc)

    def sync(self):
        self.check(self.lib.gh_sync(self.handle))

    # -- forces (host arrays, code units) --------------------------------------
    def direct(self, pos, mass, eps, precision=GH_FP64, force_pos=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        np_ = len(mass)
        if force_pos is None:
            out = np.empty((np_, 3))
            self.check(self.lib.gh_direct_summation(self.handle, _ptr(pos), _ptr(mass), np_,
                                                    float(eps), int(precision), _ptr(out)))
        else:
            fpos = np.ascontiguousarray(force_pos, dtype=np.float64)
            out = np.empty((len(fpos), 3))
            self.check(self.lib.gh_direct_summation_position(
                self.handle, _ptr(pos), _ptr(mass), np_, _ptr(fpos), len(fpos), float(eps),
                int(precision), _ptr(out)))
        return out

    def tree(self, pos, mass, eps, theta=0.7, precision=GH_FP64, force_pos=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        np_ = len(mass)
        if force_pos is None:
            out = np.empty((np_, 3))
            self.check(self.lib.gh_tree_force(self.handle, _ptr(pos), _ptr(mass), np_,
                                              float(eps), float(theta), int(precision),
                                              _ptr(out)))
        else:
            fpos = np.ascontiguousarray(force_pos, dtype=np.float64)
            out = np.empty((len(fpos), 3))
            self.check(self.lib.gh_tree_force_position(
                self.handle, _ptr(pos), _ptr(mass), np_, _ptr(fpos), len(fpos), float(eps),
                float(theta), int(precision), _ptr(out)))
        return out

    def tree_dump(self, pos, mass, eps):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        n = ctypes.c_int64(0)
        self.check(self.lib.gh_tree_dump(self.handle, _ptr(pos), _ptr(mass), len(mass),
                                         float(eps), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=NODE_DTYPE)
        self.check(self.lib.gh_tree_dump(self.handle, _ptr(pos), _ptr(mass), len(mass),
                                         float(eps), _ptr(out), n.value, ctypes.byref(n)))
        return out

    def tree_interactions(self, pos, mass, eps, theta=0.7, force_pos=None, precision=GH_FP64):
        """Tree force and interaction count per target: FP64, the reference's accepted
        nodes; FP32, the group walk's list lengths."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        fpos = None if force_pos is None else np.ascontiguousarray(force_pos, dtype=np.float64)
        nt = len(mass) if fpos is None else len(fpos)
        out = np.empty((nt, 3))
        cnt = np.empty(nt, dtype=np.int64)
        fn = (self.lib.gh_tree_interactions if precision == GH_FP64
              else self.lib.gh_tree_interactions_fp32)
        self.check(fn(
            self.handle, _ptr(pos), _ptr(mass), len(mass), _ptr(fpos), nt if fpos is not None
            else 0, float(eps), float(theta), _ptr(out), _ptr(cnt)))
        return out, cnt

    # -- device-resident leapfrog ----------------------------------------------
    def run(self, pos, vel, mass, dt, eps, nsteps, algorithm=GH_DIRECT, theta=0.7,
            precision=GH_FP64, history=False, hist_out=None):
        """Advances copies of pos/vel by nsteps DKD steps; returns (pos, vel) or
        (pos, vel, hist_pos, hist_vel) with (nsteps, n, 3) snapshots.  hist_out: a pair
        of arrays the snapshots are written into in place (implies history)."""
        pos = np.array(pos, dtype=np.float64, order="C", copy=True)
        vel = np.array(vel, dtype=np.float64, order="C", copy=True)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        n = len(mass)
        history = history or hist_out is not None
        hp, hv = _history_arrays(nsteps, n, history, hist_out)
        self.check(self.lib.gh_run(self.handle, _ptr(pos), _ptr(vel), _ptr(mass), n,
                                   float(dt), float(eps), int(algorithm), float(theta),
                                   int(precision), int(nsteps), _ptr(hp), _ptr(hv)))
        return (pos, vel, hp, hv) if history else (pos, vel)

    def run_ext(self, pos, vel, mass, dt, eps, nsteps, external, t0=0.0, pass_velocity=False,
                algorithm=GH_DIRECT, theta=0.7, precision=GH_FP64, history=False,
                hist_out=None):
        """gh_run with external accelerations: external(x_half, v, t) -> (n, 3) array in
        km/s/Myr, called once per step with copies of the half-step positions and (when
        pass_velocity) the pre-kick velocities; v is None otherwise.  An exception raised
        by `external` stops the run and is re-raised here.  hist_out as for run."""
        pos = np.array(pos, dtype=np.float64, order="C", copy=True)
        vel = np.array(vel, dtype=np.float64, order="C", copy=True)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        n = len(mass)
        history = history or hist_out is not None
        hp, hv = _history_arrays(nsteps, n, history, hist_out)
        failure = []

        def trampoline(user, n_, x_ptr, v_ptr, t, acc_ptr):
            try:
                xh = np.ctypeslib.as_array(x_ptr, shape=(n_, 3)).copy()
                vv = np.ctypeslib.as_array(v_ptr, shape=(n_, 3)).copy() if v_ptr else None
                a = np.asarray(external(xh, vv, t), dtype=np.float64)
                np.ctypeslib.as_array(acc_ptr, shape=(n_, 3))[...] = a
                return 0
            except BaseException as e:  # re-raised after the C call returns
                failure.append(e)
                return 1

        cb = EXTERNAL_FN(trampoline)
        rc = self.lib.gh_run_ext(self.handle, _ptr(pos), _ptr(vel), _ptr(mass), n, float(dt),
                                 float(eps), int(algorithm), float(theta), int(precision),
                                 int(nsteps), float(t0), cb, None, 1 if pass_velocity else 0,
                                 _ptr(hp), _ptr(hv))
        if failure:
            raise failure[0]
        self.check(rc)
        return (pos, vel, hp, hv) if history else (pos, vel)

    def run_dev(self, d_pos, d_vel, d_mass, n, dt, eps, nsteps, algorithm=GH_DIRECT,
                theta=0.7, precision=GH_FP64, d_hist_pos=0, d_hist_vel=0, stream=0):
        """gh_run_dev on raw device addresses (ints), e.g. torch tensors' data_ptr()."""
        self.check(self.lib.gh_run_dev(
            self.handle, ctypes.c_void_p(d_pos), ctypes.c_void_p(d_vel),
            ctypes.c_void_p(d_mass), int(n), float(dt), float(eps), int(algorithm),
            float(theta), int(precision), int(nsteps), ctypes.c_void_p(d_hist_pos or None),
            ctypes.c_void_p(d_hist_vel or None), ctypes.c_void_p(stream or None)))

    def direct_dev(self, d_pos, d_mass, n, eps, precision, d_acc, d_force_pos=0, nf=0,
                   stream=0):
        self.check(self.lib.gh_direct_summation_dev(
            self.handle, ctypes.c_void_p(d_pos), ctypes.c_void_p(d_mass), int(n),
            ctypes.c_void_p(d_force_pos or None), int(nf), float(eps), int(precision),
            ctypes.c_void_p(d_acc), ctypes.c_void_p(stream or None)))

    def tree_dev(self, d_pos, d_mass, n, eps, theta, precision, d_acc, d_force_pos=0, nf=0,
                 stream=0):
        self.check(self.lib.gh_tree_force_dev(
            self.handle, ctypes.c_void_p(d_pos), ctypes.c_void_p(d_mass), int(n),
            ctypes.c_void_p(d_force_pos or None), int(nf), float(eps), float(theta),
            int(precision), ctypes.c_void_p(d_acc), ctypes.c_void_p(stream or None)))

    def set_tuning(self, key, value):
        self.check(self.lib.gh_set_tuning(self.handle, int(key), int(value)))
        self._tuning[int(key)] = int(value)

    def get_tuning(self, key):
        """The value last set for a tuning key on this context (0: automatic)."""
        return self._tuning.get(int(key), 0)

    def timing_enable(self, on=True):
        self.check(self.lib.gh_timing_enable(self.handle, 1 if on else 0))

    def timing_read(self, cls):
        """(total milliseconds, launches) of one kernel class since timing_enable."""
        t = ctypes.c_double()
        n = ctypes.c_int64()
        self.check(self.lib.gh_timing_read(self.handle, int(cls), ctypes.byref(t),
                                           ctypes.byref(n)))
        return t.value, n.value

    def energy(self, pos, vel, mass, eps):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        ke = ctypes.c_double()
        pe = ctypes.c_double()
        mom = np.zeros(3)
        self.check(self.lib.gh_energy(self.handle, _ptr(pos), _ptr(vel), _ptr(mass),
                                      len(mass), float(eps), ctypes.byref(ke),
                                      ctypes.byref(pe), _ptr(mom)))
        return ke.value, pe.value, mom


class Ensemble:
    """gh_ensemble: independent direct-sum simulations advanced together.

    sims: list of dicts with 'pos' (n,3), 'vel' (n,3), 'mass' (n,), 'eps', 'dt' in
    code units (kpc, km/s, Msun, Myr)."""

    def __init__(self, ctx, sims, ipt=4, persistent=False):
        self.ctx = ctx
        self.counts = np.array([len(s["mass"]) for s in sims], dtype=np.int64)
        pos = np.ascontiguousarray(np.concatenate([s["pos"] for s in sims]), dtype=np.float64)
        vel = np.ascontiguousarray(np.concatenate([s["vel"] for s in sims]), dtype=np.float64)
        mass = np.ascontiguousarray(np.concatenate([s["mass"] for s in sims]), dtype=np.float64)
        eps = np.array([s["eps"] for s in sims], dtype=np.float64)
        dt = np.array([s["dt"] for s in sims], dtype=np.float64)
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.gh_ensemble_create(ctx.handle, len(sims), _ptr(self.counts), _ptr(pos),
                                             _ptr(vel), _ptr(mass), _ptr(eps), _ptr(dt),
                                             int(ipt), 1 if persistent else 0,
                                             ctypes.byref(h)))
        self.handle = h
        self.ipt = ipt
        self.persistent = persistent

    def run(self, nsteps, graph=True):
        self.ctx.check(self.ctx.lib.gh_ensemble_run(self.handle, int(nsteps), 1 if graph else 0))

    def state(self):
        n = int(self.counts.sum())
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        self.ctx.check(self.ctx.lib.gh_ensemble_get(self.handle, _ptr(pos), _ptr(vel)))
        splits = np.cumsum(self.counts)[:-1]
        return np.split(pos, splits), np.split(vel, splits)

    def info(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.ctx.lib.gh_ensemble_info(self.handle, ctypes.byref(a),
                                                     ctypes.byref(b), ctypes.byref(c)))
        return {"n_padded": a.value, "block_pairs": b.value, "pairs_per_step": c.value}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.ctx.lib.gh_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    """The process-wide context on device $GRAVHOP_DEVICE (default 0)."""
    global _default
    with _lock:
        if _default is None:
            _default = Context(int(os.environ.get("GRAVHOP_DEVICE", "0")))
        return _default
