"""Writes tests/golden/reference_vectors.npz: input/output vectors of the reference's
own C engine (gravhopper/_jbgrav.c compiled from its source into oracle/_ref by
oracle/Makefile).  Run in a checkout where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py

Only inputs and outputs are stored (float64 arrays), so the checks in tests/ work
where the reference source is absent.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def cases():
    rng = np.random.default_rng(20260001)
    out = []
    # (name, n, nf, eps, thetas, generator)
    out.append(("two_body", np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), np.array([1.0, 3.0]),
                0.0))
    out.append(("equilateral", np.array([[1.0, 0.0, 0.0], [-0.5, np.sqrt(3) / 2, 0.0],
                                         [-0.5, -np.sqrt(3) / 2, 0.0]]), np.ones(3), 0.01))
    for n, eps in ((17, 0.05), (64, 0.0), (256, 1e-3), (1000, 0.02)):
        pos = rng.normal(size=(n, 3)) * rng.uniform(0.1, 10.0)
        mass = rng.uniform(0.5, 2.0, size=n) * 1e3
        out.append(("gauss_%d" % n, pos, mass, eps))
    # a Plummer-like core/halo with very different scales (deep tree)
    n = 512
    xi = rng.uniform(0, 1, n)
    r = 1e-3 / np.sqrt(xi ** (-2.0 / 3.0) - 1.0)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    out.append(("plummer_512", r[:, None] * d, np.full(n, 1e6 / n), 5e-5))
    # near-wall positions: coordinates exactly on dyadic cell boundaries of the root
    n = 200
    pos = rng.integers(-8, 9, size=(n, 3)).astype(np.float64) / 8.0
    pos = np.unique(pos, axis=0)
    out.append(("grid_walls", pos, np.ones(len(pos)), 0.01))
    return out


def main():
    ref = O.ref_module()
    if ref is None:
        raise SystemExit("oracle/_ref is not built; run `make -C oracle ref` first")
    rng = np.random.default_rng(20260002)
    store = {}
    names = []
    for name, pos, mass, eps in cases():
        names.append(name)
        nf = max(1, len(pos) // 3)
        fpos = pos[rng.integers(0, len(pos), size=nf)] + rng.normal(size=(nf, 3)) * 0.1
        fpos[0] = pos[0]  # one target exactly on a particle
        store[name + "/pos"] = pos
        store[name + "/mass"] = mass
        store[name + "/eps"] = np.array(eps)
        store[name + "/force_pos"] = fpos
        store[name + "/direct"] = ref.direct_summation(pos, mass, eps)
        store[name + "/direct_position"] = ref.direct_summation_position(pos, mass, fpos, eps)
        for th in (0.0, 0.3, 0.7, 1e30):
            store[name + "/tree_%g" % th] = ref.tree_force(pos, mass, eps, th)
            store[name + "/tree_position_%g" % th] = ref.tree_force_position(pos, mass, fpos,
                                                                             eps, th)
    store["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **store)
    print("wrote", len(names), "cases")


if __name__ == "__main__":
    main()
