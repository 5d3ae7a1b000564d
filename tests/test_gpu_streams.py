"""Calls on one context in call order across streams.

The device-pointer entry points run on the caller's stream and the host-pointer ones on
the context's stream, and all of them use the context's arenas (the FP32 paths share the
spatial order, the local-frame tiles and the block-pair list).  A call on a stream other
than the previous call's waits for that call first, so the results below must be
bitwise those of the same calls made one at a time with synchronisation between them.
"""
import numpy as np
import pytest

from conftest import assert_bitwise
from gravhopper_amd import _native, workloads

pytestmark = pytest.mark.gpu

FP32 = _native.GH_FP32


def inputs():
    a = workloads.plummer(3000, b=1e-3, mass=1e6, seed=51)
    b = workloads.plummer(4500, b=2e-3, mass=3e6, seed=52)
    return a, b


def test_leapfrog_on_side_stream_then_host_direct(ctx):
    torch = pytest.importorskip("torch")
    from gravhopper_amd.autograd import leapfrog
    a, b = inputs()
    args = (0.01, 2e-5, 20)

    def run_leapfrog():
        x0 = torch.tensor(a["pos"], device="cuda")
        v0 = torch.tensor(a["vel"], device="cuda")
        m = torch.tensor(a["mass"], device="cuda")
        return leapfrog(x0, v0, m, *args, checkpoint_every=5, precision="fp32")

    # one at a time
    x_sep, v_sep = (t.cpu().numpy() for t in run_leapfrog())
    torch.cuda.synchronize()
    acc_sep = ctx.direct(b["pos"], b["mass"], 5e-5, FP32)

    # back to back: the leapfrog is queued on a side stream and the host call follows
    # at once, on the context's stream
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        x, v = run_leapfrog()
    acc = ctx.direct(b["pos"], b["mass"], 5e-5, FP32)
    side.synchronize()
    assert_bitwise(acc, acc_sep, "host direct after side-stream leapfrog")
    assert_bitwise(x.cpu().numpy(), x_sep, "leapfrog x")
    assert_bitwise(v.cpu().numpy(), v_sep, "leapfrog v")


def test_device_calls_on_two_streams(ctx):
    torch = pytest.importorskip("torch")
    a, b = inputs()

    def direct_dev(ic, stream):
        pos = torch.tensor(ic["pos"], device="cuda")
        mass = torch.tensor(ic["mass"], device="cuda")
        acc = torch.empty_like(pos)
        torch.cuda.synchronize()
        ctx.direct_dev(pos.data_ptr(), mass.data_ptr(), len(ic["mass"]), 5e-5, FP32,
                       acc.data_ptr(), stream=stream.cuda_stream)
        return acc, (pos, mass)

    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    want = []
    for ic in (a, b):
        acc, keep = direct_dev(ic, s1)
        s1.synchronize()
        want.append(acc.cpu().numpy())
    # the second call's inputs are made before the first call is queued, so the two
    # calls follow each other with nothing between them
    pos_b = torch.tensor(b["pos"], device="cuda")
    mass_b = torch.tensor(b["mass"], device="cuda")
    acc_b = torch.empty_like(pos_b)
    pos_a = torch.tensor(a["pos"], device="cuda")
    mass_a = torch.tensor(a["mass"], device="cuda")
    acc_a = torch.empty_like(pos_a)
    torch.cuda.synchronize()
    ctx.direct_dev(pos_a.data_ptr(), mass_a.data_ptr(), len(a["mass"]), 5e-5, FP32,
                   acc_a.data_ptr(), stream=s1.cuda_stream)
    ctx.direct_dev(pos_b.data_ptr(), mass_b.data_ptr(), len(b["mass"]), 5e-5, FP32,
                   acc_b.data_ptr(), stream=s2.cuda_stream)
    s1.synchronize()
    s2.synchronize()
    assert_bitwise(acc_a.cpu().numpy(), want[0], "first call, stream 1")
    assert_bitwise(acc_b.cpu().numpy(), want[1], "second call, stream 2")
