#!/usr/bin/env python3
"""Writes tests/golden/ref_restart_tri.msh and ref_restart_tri.dat: a restart
written by the reference's own DGAnalysis::write (DGPref.cc:712-740, run through
oracle/_ref/ref_driver, built from the reference sources by oracle/Makefile.ref)
for a P2 Euler state on a 6x4 lattice of [0,1]x[0,1] (walls at the bottom and
top, far field at the ends; every lattice vertex its own Gmsh node).  The mesh
file is this script's (tests/helpers.py lattice_msh); the data stream carries
the reference's vertex ids (the Gmsh node numbers) and its coefficients.

    python tests/golden/make_ref_restart.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.dirname(os.path.dirname(HERE))]
import ref_harness as rh  # noqa: E402
from helpers import lattice_msh  # noqa: E402

FF = (1.0, 0.2, -0.1, 2.6)
NX, NY, P = 6, 4, 2


def main():
    if not rh.available():
        print("oracle/_ref/ref_driver is not built (make -C oracle -f Makefile.ref)")
        return 1
    with tempfile.TemporaryDirectory() as d:
        mp, sp, op = os.path.join(d, "m.msh"), os.path.join(d, "s.txt"), os.path.join(d, "o.dat")
        lattice_msh(mp, NX, NY, side_tags=(40000, 1, 40000, 1))
        rng = np.random.default_rng(42)
        nid = lambda i, j: j * (NX + 1) + i + 1  # noqa: E731
        with open(sp, "w") as f:
            for j in range(NY):
                for i in range(NX):
                    for tri in ((nid(i, j), nid(i + 1, j), nid(i + 1, j + 1)),
                                (nid(i, j), nid(i + 1, j + 1), nid(i, j + 1))):
                        c = np.zeros((4, 6))
                        c[:, 0] = np.array(FF) / np.sqrt(2.0)
                        c += 0.01 * rng.standard_normal(c.shape)
                        f.write("%d %d %d " % tri + " ".join("%.17e" % x for x in c.ravel()) + "\n")
        env = dict(os.environ, REF_FARFIELD=" ".join("%r" % x for x in FF))
        r = subprocess.run([rh.DRIVER, mp, "euler", "roe", str(P), "write", sp, op], cwd=d, env=env,
                           capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            print(r.stdout[-2000:], r.stderr[-2000:])
            return 1
        shutil.copy(mp, os.path.join(HERE, "ref_restart_tri.msh"))
        shutil.copy(op, os.path.join(HERE, "ref_restart_tri.dat"))
    print("wrote tests/golden/ref_restart_tri.msh, ref_restart_tri.dat")
    return 0


if __name__ == "__main__":
    sys.exit(main())
