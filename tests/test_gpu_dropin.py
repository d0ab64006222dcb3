"""The host C++ mirror classes (ConservationLaw / DGAnalysis / TimeIntegrator
over the C ABI) driven from a C++ program, checked against the CPU oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import ROOT, VORTEX, mesh_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_vortex")


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_matches_oracle():
    print("\n[gpu case] C++ drop-in vortex", flush=True)
    with tempfile.TemporaryDirectory() as d:
        pre = os.path.join(d, "run")
        r = subprocess.run([EXE, pre], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        U0 = np.fromfile(pre + ".U0")
        R = np.fromfile(pre + ".R")
        U2 = np.fromfile(pre + ".U2")
        lines = dict(l.split(None, 1) for l in r.stdout.splitlines())
        dt = float(lines["dt"])
    vids, xyz, bc = mesh_tri(10, 10, lx=10.0, ly=10.0, jitter=0.1, seed=77)
    o = po.Oracle(2, vids, xyz, 3, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4, farfield=(1.0, 1.0, 1.0, 3.5))
    Uo = o.project(1, VORTEX)
    assert np.abs(U0 - Uo).max() / np.abs(Uo).max() < 1e-13
    Ro, _ = o.residual(U0)
    assert np.abs(R - Ro).max() / np.abs(Ro).max() < 1e-12
    assert abs(dt - o.dt(U0, 0.1)) <= 1e-14 * dt
    U = U0
    for s in range(2):
        U, _ = o.step(U, po.RK4, s * dt, dt)
    assert np.abs(U2 - U).max() / np.abs(U).max() < 1e-12
