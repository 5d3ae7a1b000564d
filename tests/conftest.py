import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch ships its own HIP runtime; loaded after libgravhop.so has initialised the system
# one, it finds no GPU ("No HIP GPUs are available").  Tests that hand torch tensors to
# the library therefore need torch loaded first, whatever order the tests run in.
try:
    import torch  # noqa: F401
except ImportError:
    pass

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libgravhop.so kernels)")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def golden():
    z = np.load(GOLDEN, allow_pickle=False)
    names = [str(n) for n in z["names"]]
    cases = {}
    for n in names:
        cases[n] = {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(n + "/")}
    return cases


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx():
    """The GPU context.  GPU tests fail (not skip) when the library or the GPU is
    missing: the HIP path is the one under test."""
    from gravhopper_amd import _native
    return _native.default_context()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bitwise(a, b, what=""):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    same = (bits(a) == bits(b))
    if not same.all():
        idx = np.argwhere(~same)[:5]
        raise AssertionError("%s: %d/%d values differ, first at %s: %r vs %r" % (
            what, int((~same).sum()), same.size, idx.tolist(),
            [a[tuple(i)] for i in idx], [b[tuple(i)] for i in idx]))


def rel_err(a, ref):
    """|a - ref| / |ref| per row (vector relative error)."""
    num = np.linalg.norm(np.asarray(a) - np.asarray(ref), axis=1)
    den = np.linalg.norm(np.asarray(ref), axis=1)
    return num / np.where(den > 0, den, 1.0)
