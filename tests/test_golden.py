"""The oracle (with its own exact tables) against golden fixtures produced by the
reference's own DG classes (tests/golden/make_ref_fixtures.py).  Tolerances per
order reflect the reference's printed basis/quadrature tables (see
test_oracle_vs_reference.py)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "ref_fixtures.npz")
OWN_TOL = {1: 5e-13, 2: 5e-13, 3: 2e-11, 4: 5e-9}
OPS = ["residual", "fe", "tvd2", "tvd3", "rk4"]
SCHEME = {"fe": po.RK_FE, "tvd2": po.RK_TVD2, "tvd3": po.RK_TVD3, "rk4": po.RK4}


def cases():
    z = np.load(FIX, allow_pickle=False)
    ks = sorted({k.split("_")[0] for k in z.files})
    return [(k, {n[len(k) + 1:]: z[n] for n in z.files if n.startswith(k + "_")}) for k in ks]


CASES = cases()


@pytest.mark.parametrize("name,c", CASES, ids=[k for k, _ in CASES])
def test_oracle_matches_reference_fixture(name, c):
    euler, flux, p, op, nsteps = (int(x) for x in c["meta"])
    orders = c["orders"] if "orders" in c else None  # mixed orders P1-P4
    tol_p = int(orders.max()) if orders is not None else p
    if orders is not None:
        p = orders
    if euler:
        o = po.Oracle(2, c["vids"], c["xyz"], p, po.LAW_EULER2D, [0, po.FLUX_LF, po.FLUX_ROE][flux], gamma=1.4,
                      farfield=tuple(c["farfield"]), bc=c["bc"])
    else:
        o = po.Oracle(2, c["vids"], c["xyz"], p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5),
                      farfield=tuple(c["farfield"][:1]), bc=c["bc"])
    if OPS[op] == "residual":
        got, _ = o.residual(c["U"])
        tol = OWN_TOL[tol_p]
    else:
        got, t, dt = c["U"], 0.0, float(c["dt"][0])
        for s in range(nsteps):
            got, r = o.step(got, SCHEME[OPS[op]], t, dt)
            t += dt
        tol = OWN_TOL[tol_p]
        assert abs(r - c["resid"][-1]) <= 1e-6 * abs(c["resid"][-1])
    ref = c["out"]
    if OPS[op] != "residual":
        # compare the increments, which carry the discretisation differences
        got, ref = got - c["U"], ref - c["U"]
    assert np.abs(got - ref).max() / np.abs(ref).max() < tol
