#!/usr/bin/env python3
"""Writes tests/golden/ref_fixtures.npz: inputs and outputs of the reference's own
DG classes (oracle/_ref/ref_driver, built from the reference sources by
oracle/Makefile.ref) on small jittered meshes of [-1,1]^2.  Each case stores the
mesh (topological ids, corner coordinates, boundary tags), law, order, the input
state and the reference residual (and, for step cases, the state after 3 steps).

    python tests/golden/make_ref_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.dirname(os.path.dirname(HERE))]
import ref_harness as rh  # noqa: E402
from helpers import np_tri  # noqa: E402

FF = (1.0, 0.2, -0.1, 2.6)


def state(law, p, E, seed):
    rng = np.random.default_rng(seed)
    if law == "scalar":
        return rng.standard_normal(E * np_tri(p))
    U = np.zeros((E, 4, np_tri(p)))
    U[:, :, 0] = np.array(FF) / np.sqrt(2.0)
    U += 0.01 * rng.standard_normal(U.shape)
    return U.ravel()


def main():
    if not rh.available():
        print("oracle/_ref/ref_driver is not built (make -C oracle -f Makefile.ref)")
        return 1
    out = {}
    k = 0
    cases = []
    for p in (1, 2, 3, 4):
        for law, flux, periodic in (("scalar", "upwind", True), ("euler", "lf", True), ("euler", "roe", False)):
            cases.append((law, flux, periodic, p, "residual", 0))
    cases += [("euler", "lf", True, 3, "rk4", 3), ("scalar", "upwind", True, 2, "tvd3", 3),
              ("euler", "roe", False, 2, "tvd2", 3), ("scalar", "upwind", False, 1, "fe", 3)]
    # mixed orders P1-P4 (p = 0: a seeded per-element order map), bounded meshes
    cases += [("euler", "roe", False, 0, "residual", 0), ("scalar", "upwind", False, 0, "residual", 0),
              ("euler", "roe", False, 0, "tvd3", 3)]
    for law, flux, periodic, p, op, nsteps in cases:
        m = rh.square_mesh(4, jitter=0.12, seed=100 + k, periodic=periodic, side_tags=(40000, 7, 40000, 7))
        E = len(m["lat"])
        if p == 0:
            orders = np.random.default_rng(500 + k).integers(1, 5, E).astype(np.int32)
            U = np.concatenate([state(law, int(q), 1, 1000 * k + e) for e, q in enumerate(orders)])
        else:
            U = state(law, p, E, k)
        dt = 0.002 if op != "residual" else 0.0
        ff = (0.3, 0, 0, 0) if law == "scalar" else FF
        R, vorder, resid = rh.run(m, law, flux, orders if p == 0 else p, op, U, dt=dt, nsteps=max(nsteps, 1),
                                  vel=(1.0, 0.5), farfield=ff)
        assert np.array_equal(vorder, m["lat"])
        pre = "c%02d_" % k
        out[pre + "meta"] = np.array([law == "euler", {"upwind": 0, "lf": 1, "roe": 2}[flux], p,
                                      ["residual", "fe", "tvd2", "tvd3", "rk4"].index(op), nsteps])
        out[pre + "vids"] = m["top"]
        out[pre + "xyz"] = m["xyz"]
        out[pre + "bc"] = m["bc"]
        out[pre + "U"] = U
        out[pre + "out"] = R
        out[pre + "dt"] = np.array([dt])
        out[pre + "farfield"] = np.array(ff)
        out[pre + "resid"] = np.array(resid if resid else [0.0])
        if p == 0:
            out[pre + "orders"] = orders
        k += 1
    np.savez_compressed(os.path.join(HERE, "ref_fixtures.npz"), **out)
    print("wrote %d cases" % k)
    return 0


if __name__ == "__main__":
    sys.exit(main())
