"""gh_energy_dev_noalloc: the diagnostics of gh_energy without any device allocation.

Its per-block partial sums live in the library's static device array; the inputs and
the 5-double output are the caller's device buffers.  Checked bit for bit against
gh_energy (same arithmetic and summation order) on two calls in a row, the second with
fewer particles, so that the partials left by the first call must not leak into the
second.
"""
import ctypes

import numpy as np
import pytest

from gravhopper_amd import _native, workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def noalloc_energy(lib, pos, vel, mass, eps):
    dev = torch.device("cuda")
    tp = torch.tensor(pos, dtype=torch.float64, device=dev).contiguous()
    tv = torch.tensor(vel, dtype=torch.float64, device=dev).contiguous()
    tm = torch.tensor(mass, dtype=torch.float64, device=dev).contiguous()
    out = torch.full((5,), float("nan"), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    rc = lib.gh_energy_dev_noalloc(ctypes.c_void_p(tp.data_ptr()), ctypes.c_void_p(tv.data_ptr()),
                                   ctypes.c_void_p(tm.data_ptr()), len(mass), float(eps),
                                   ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    assert rc == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_energy_matches_gh_energy(ctx):
    lib = ctx.lib
    for n, seed in ((5000, 31), (1200, 32)):
        print("n", n, flush=True)
        ic = workloads.plummer(n, seed=seed)
        got = noalloc_energy(lib, ic["pos"], ic["vel"], ic["mass"], 5e-5)
        ke, pe, mom = ctx.energy(ic["pos"], ic["vel"], ic["mass"], 5e-5)
        want = np.array([ke, pe, mom[0], mom[1], mom[2]])
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (n, got, want)


def test_noalloc_energy_rejects_oversize(ctx):
    # the limit is checked on the host before any GPU work
    rc = ctx.lib.gh_energy_dev_noalloc(None, None, None, (1 << 22) + 1, 0.1,
                                       ctypes.c_void_p(8), None)
    assert rc == _native.GH_EARG
