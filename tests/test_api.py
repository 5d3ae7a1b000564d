"""Host-side behaviour of the drop-in (no GPU needed): the C ABI library loads and
exports every declared entry point, argument validation raises the reference's
errors, and Simulation's parameter/IC handling follows gravhopper.py."""
import os
import re

import numpy as np
import pytest

import gravhopper_amd as gh
from gravhopper_amd import _jbgrav, _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gravhop.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gh_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    # and the ctypes prototypes cover exactly the declared set
    assert sorted(_native.SIGNATURES) == syms


def test_python_constants_match_header():
    # every GH_TUNE_*, GH_T_* and precision / algorithm constant the Python layer names
    # has the header's value
    txt = open(os.path.join(ROOT, "include", "gravhop.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define\s+(GH_[A-Z0-9_]+)\s+(-?\d+)\b", txt)}
    tune = {k: v for k, v in defs.items() if k.startswith("GH_TUNE_") and k != "GH_TUNE_COUNT"}
    assert len(tune) == defs["GH_TUNE_COUNT"]
    for k, v in tune.items():
        assert getattr(_native, k[3:]) == v, k
    for k in ("GH_FP64", "GH_FP32", "GH_DIRECT", "GH_TREE"):
        assert getattr(_native, k) == defs[k], k


def test_library_built_for_gfx950():
    data = open(_native.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_shim_shape_errors_match_reference():
    # messages of gravhopper/_jbgrav.c:89-108
    with pytest.raises(RuntimeError, match=r"^Position array does not have 2 dimensions\.$"):
        _jbgrav.direct_summation(np.zeros(3), np.zeros(1), 0.1)
    with pytest.raises(RuntimeError, match=r"^Position array is not Nx3\.$"):
        _jbgrav.direct_summation(np.zeros((4, 2)), np.zeros(4), 0.1)
    with pytest.raises(RuntimeError, match="Mass array and position array contain different"):
        _jbgrav.direct_summation(np.zeros((4, 3)), np.zeros(5), 0.1)
    with pytest.raises(RuntimeError, match=r"^Position array is not Nx3\.$"):
        _jbgrav.tree_force(np.zeros((4, 4)), np.zeros(4), 0.1, 0.7)
    # _position variants, :224-261 / :658-697
    with pytest.raises(RuntimeError, match=r"^Particle position array does not have 2"):
        _jbgrav.direct_summation_position(np.zeros(3), np.zeros(1), np.zeros((1, 3)), 0.1)
    with pytest.raises(RuntimeError, match="Mass array and particle position array"):
        _jbgrav.tree_force_position(np.zeros((2, 3)), np.zeros(3), np.zeros((1, 3)), 0.1, 0.7)
    with pytest.raises(RuntimeError, match=r"^Force position array does not have 2"):
        _jbgrav.direct_summation_position(np.zeros((2, 3)), np.zeros(2), np.zeros(3), 0.1)
    with pytest.raises(RuntimeError, match=r"^Force position array is not Nx3\.$"):
        _jbgrav.tree_force_position(np.zeros((2, 3)), np.zeros(2), np.zeros((1, 2)), 0.1, 0.7)
    with pytest.raises(TypeError):
        _jbgrav.direct_summation(np.zeros((2, 3)), np.zeros(2), "x")
    with pytest.raises(ValueError):
        _jbgrav.direct_summation(np.zeros((2, 3)), np.zeros(2), 0.1, precision="fp16")


def test_simulation_parameters():
    sim = gh.Simulation(dt=0.01, eps=0.001, algorithm='direct')
    assert sim.get_dt() == 0.01 and sim.get_eps() == 0.001
    assert sim.get_algorithm() == 'direct'
    with pytest.raises(ValueError, match="algorithm must be 'tree' or 'direct'."):
        sim.set_algorithm('fmm')
    with pytest.raises(ValueError, match="dt must have dimensions of time."):
        sim.set_dt("soon")
    with pytest.raises(ValueError, match="eps must have dimensions of length."):
        sim.set_eps(object())
    with pytest.raises(ValueError):
        sim.set_precision('fp16')
    # defaults of gravhopper.py:116: 1 Myr, 100 pc, tree
    d = gh.Simulation()
    assert d.get_dt() == 1.0 and d.get_eps() == 0.1 and d.get_algorithm() == 'tree'


def test_simulation_ic_handling():
    sim = gh.Simulation()
    with pytest.raises(gh.UninitializedSimulationException):
        sim.run(1)
    with pytest.raises(gh.ICException):
        sim.add_IC({'pos': np.zeros((2, 3)), 'vel': np.zeros((2, 3))})
    with pytest.raises(gh.ICException):
        sim.add_IC({'pos': np.zeros((2, 3)), 'vel': np.zeros((3, 3)), 'mass': np.ones(2)})
    sim.add_IC({'pos': np.zeros((2, 3)), 'vel': np.zeros((2, 3)), 'mass': np.ones(2)})
    sim.add_IC({'pos': np.ones(3), 'vel': np.ones(3), 'mass': np.ones(1)})
    assert sim.ICarrays['pos'].shape == (3, 3) and sim.ICarrays['mass'].shape == (3,)
    sim.init_run(5)
    assert sim.positions.shape == (6, 3, 3) and sim.times.shape == (6,)
    assert np.array_equal(sim.snap(0)['pos'][2], np.ones(3))
    sim.reset()
    assert sim.running is False and sim.positions is None


def test_external_force_registration_rejects_non_callables():
    sim = gh.Simulation()
    with pytest.raises(gh.ExternalPackageException):
        sim.add_external_force(42)
    sim.add_external_force([lambda p, a: 0 * p, lambda p, a: 0 * p])
    assert len(sim.extra_force_functions) == 2


def test_workload_generators_shapes_and_centring():
    from gravhopper_amd import workloads as W
    ic = W.plummer(2000, b=1e-3, mass=1e6, seed=1)
    assert ic['pos'].shape == (2000, 3) and ic['mass'].sum() == pytest.approx(1e6)
    assert np.abs(ic['pos'].mean(0)).max() < 1e-15
    assert np.abs(ic['vel'].mean(0)).max() < 1e-12
    # Plummer half-mass radius ~ 1.305 b
    r = np.sort(np.linalg.norm(ic['pos'], axis=1))
    assert 1.1e-3 < r[1000] < 1.5e-3
    t = W.plummer(5000, seed=2, velocities="tabulated")
    assert np.all(np.isfinite(t['vel']))
    g = W.galaxy(n_halo=4096, n_disk=4096, seed=5)
    assert g['pos'].shape == (8192, 3) and np.all(np.isfinite(g['vel']))


def test_run_history_output_arrays_are_checked():
    # snapshots can be written in place into caller arrays, e.g. rows of Simulation's
    # history; anything the library could not write through a plain pointer is refused
    hist = np.zeros((10, 5, 3))
    hp, hv = _native._history_arrays(4, 5, False, (hist[1:5], hist[5:9]))
    assert hp.base is hist and hv.base is hist
    assert _native._history_arrays(4, 5, False, None) == (None, None)
    a, b = _native._history_arrays(4, 5, True, None)
    assert a.shape == b.shape == (4, 5, 3)
    bad = [np.zeros((4, 5, 3), dtype=np.float32), np.zeros((3, 5, 3)),
           np.zeros((4, 5, 6))[:, :, :3], np.zeros((4, 5, 3)).view()]
    bad[3].flags.writeable = False
    for arr in bad:
        with pytest.raises(ValueError, match="hist_out"):
            _native._history_arrays(4, 5, True, (arr, np.zeros((4, 5, 3))))
