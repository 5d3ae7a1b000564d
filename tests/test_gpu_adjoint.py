"""Adjoint of the direct-sum leapfrog (gravhopper_amd.autograd).

Checked against (1) central finite differences of the FP64 forward at N=256 along
random directions of (x0, v0, m), and (2) an FP64 torch-autograd restatement of the
same leapfrog at N=2048.  Horizons are short (tens of steps): the N-body map is chaotic,
so over long horizons the gradient is dominated by exponentially growing sensitivities
and no two implementations agree.  Tolerances are stated per test.
"""
import numpy as np
import pytest

from gravhopper_amd import _native, workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
ACCEL, DRIFT = 4.398600412921223e-09, 1.022712165045695e-03


def loss_np(x, target):
    return float(np.mean(np.sum((x - target) ** 2, axis=1)))


def config(n, seed):
    ic = workloads.plummer(n, b=1e-3, mass=1e6, seed=seed)
    rng = np.random.default_rng(seed + 1)
    m_t = ic["mass"] * (1 + 0.05 * rng.normal(size=n))
    return ic, m_t


def test_adjoint_vs_finite_differences(ctx):
    from gravhopper_amd.autograd import leapfrog
    n, steps, every, dt, eps = 256, 24, 5, 0.005, 5e-5
    ic, m_t = config(n, 300)
    target, _ = ctx.run(ic["pos"], ic["vel"], m_t, dt, eps, steps, precision=_native.GH_FP64)
    dev = "cuda"
    x0 = torch.tensor(ic["pos"], device=dev, requires_grad=True)
    v0 = torch.tensor(ic["vel"], device=dev, requires_grad=True)
    m = torch.tensor(ic["mass"], device=dev, requires_grad=True)
    tgt = torch.tensor(target, device=dev)
    x, _ = leapfrog(x0, v0, m, dt, eps, steps, checkpoint_every=every, precision="fp64")
    loss = ((x - tgt) ** 2).sum(-1).mean()
    loss.backward()
    gx, gv, gm = x0.grad.cpu().numpy(), v0.grad.cpu().numpy(), m.grad.cpu().numpy()
    rng = np.random.default_rng(7)
    scales = (1e-3, 1.0, ic["mass"][0])
    for trial in range(4):
        rx, rv, rm = (rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=n))
        # relative step 1e-8: at 1e-6 the loss is already visibly nonlinear over 24 steps
        # (central differences then differ from the exact derivative by ~3e-3)
        h = 1e-8
        dx, dv, dm = h * scales[0] * rx, h * scales[1] * rv, h * scales[2] * rm

        def L(sign):
            xx, _ = ctx.run(ic["pos"] + sign * dx, ic["vel"] + sign * dv, ic["mass"] + sign * dm,
                            dt, eps, steps, precision=_native.GH_FP64)
            return loss_np(xx, target)

        fd = (L(1) - L(-1)) / 2
        ad = (gx * dx).sum() + (gv * dv).sum() + (gm * dm).sum()
        print("trial %d fd %.6e adjoint %.6e rel %.2e" % (trial, fd, ad, abs(fd - ad) / abs(fd)))
        assert abs(fd - ad) <= 2e-4 * abs(fd)


def torch_leapfrog(x, v, m, dt, eps, nsteps):
    """FP64 torch restatement of the drift-kick-drift direct-sum leapfrog."""
    c = 0.5 * dt * DRIFT
    n = x.shape[0]
    off = 1.0 - torch.eye(n, dtype=x.dtype, device=x.device)
    for _ in range(nsteps):
        xh = x + c * v
        d = xh[None, :, :] - xh[:, None, :]
        r2 = (d * d).sum(-1) + eps * eps
        w = off * r2 ** -1.5
        a = ((m[None, :] * w)[:, :, None] * d).sum(1) * ACCEL
        v = v + dt * a
        x = xh + c * v
    return x, v


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-4), ("fp32", 2e-3)])
def test_adjoint_vs_torch_fp64(ctx, precision, tol):
    from gravhopper_amd.autograd import leapfrog
    n, steps, every, dt, eps = 2048, 20, 5, 0.005, 5e-5
    ic, m_t = config(n, 400)
    target, _ = ctx.run(ic["pos"], ic["vel"], m_t, dt, eps, steps, precision=_native.GH_FP64)
    dev = "cuda"
    tgt = torch.tensor(target, device=dev)
    grads = []
    for impl in ("ours", "torch"):
        x0 = torch.tensor(ic["pos"], device=dev, requires_grad=True)
        v0 = torch.tensor(ic["vel"], device=dev, requires_grad=True)
        m = torch.tensor(ic["mass"], device=dev, requires_grad=True)
        if impl == "ours":
            x, _ = leapfrog(x0, v0, m, dt, eps, steps, checkpoint_every=every,
                            precision=precision)
        else:
            x, _ = torch_leapfrog(x0, v0, m, dt, eps, steps)
        loss = ((x - tgt) ** 2).sum(-1).mean()
        loss.backward()
        grads.append([t.grad.cpu().numpy() for t in (x0, v0, m)])
        torch.cuda.synchronize()
    for k, name in enumerate(("x0", "v0", "m")):
        a, b = grads[0][k], grads[1][k]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print("%s %s: relative difference %.3e" % (precision, name, rel))
        assert rel <= tol, name


def test_adjoint_checkpoint_interval_invariance(ctx):
    from gravhopper_amd.autograd import leapfrog
    n, steps, dt, eps = 512, 12, 0.005, 5e-5
    ic, _ = config(n, 500)
    out = []
    for every in (1, 5, 12, 40):
        x0 = torch.tensor(ic["pos"], device="cuda", requires_grad=True)
        v0 = torch.tensor(ic["vel"], device="cuda", requires_grad=True)
        m = torch.tensor(ic["mass"], device="cuda", requires_grad=True)
        x, v = leapfrog(x0, v0, m, dt, eps, steps, checkpoint_every=every)
        (x.square().sum() + 1e-3 * v.square().sum()).backward()
        out.append([t.grad.cpu().numpy() for t in (x0, v0, m)])
    for o in out[1:]:
        for a, b in zip(o, out[0]):
            assert np.array_equal(a, b)
