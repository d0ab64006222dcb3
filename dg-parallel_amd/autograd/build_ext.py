#!/usr/bin/env python3
"""Builds dgpar_torch (dg_torch.cpp: the discrete RK4 adjoint as a libtorch
autograd Function over libdgpar's C ABI) in place:
dg-parallel_amd/autograd/dgpar_torch.so.  Needs libdgpar.so built first.

    python dg-parallel_amd/autograd/build_ext.py
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)


def build(verbose=False):
    import torch.utils.cpp_extension as ce
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    so = os.path.join(bdir, "dgpar_torch.so")
    if os.path.exists(so):
        os.remove(so)
    try:
        # builds, then imports from the build folder, where the relative rpath
        # ($ORIGIN/.. = dg-parallel_amd once the module sits in autograd/) does not
        # resolve yet: the import error is expected, the built file is what counts
        ce.load(name="dgpar_torch", sources=[os.path.join(HERE, "dg_torch.cpp")],
                extra_include_paths=[os.path.join(ROOT, "include")],
                extra_cflags=["-O2"],
                extra_ldflags=["-L" + PKG, "-ldgpar", "-Wl,-rpath,'$$ORIGIN/..'"],  # ninja turns $$ into $, the quotes keep it from the shell
                build_directory=bdir, verbose=verbose, is_python_module=True)
    except ImportError:
        pass
    if not os.path.exists(so):
        raise RuntimeError("dgpar_torch build failed")
    shutil.copy2(so, os.path.join(HERE, "dgpar_torch.so"))


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
