"""Differentiable RK4 steps of the DG residual for PyTorch: the discrete adjoint
(dg_adjoint_forward / dg_adjoint_backward of libdgpar) wrapped as a
torch::autograd::Function in C++ (autograd/dg_torch.cpp, built in place as
autograd/dgpar_torch.so by autograd/build_ext.py).

    import torch, dgpar, dgpar_autograd
    g = dgpar.DGAnalysis(mesh, law, 3)
    U0 = torch.tensor(U, dtype=torch.float64, requires_grad=True)
    UT = dgpar_autograd.rk4(U0, g, dt, nsteps, ckpt_every)
    loss(UT).backward()          # U0.grad = dL/dU0 through every RK4 stage

Tensors are host FP64 in the caller layout; the steps run on the GPU.
"""
import importlib.machinery
import importlib.util
import os

import torch  # noqa: F401  (torch's runtime is loaded before the library's)

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "autograd", "dgpar_torch.so")
if not os.path.exists(_SO):
    raise ImportError("dgpar_torch.so is not built: python dg-parallel_amd/autograd/build_ext.py")
_loader = importlib.machinery.ExtensionFileLoader("dgpar_torch", _SO)
_spec = importlib.util.spec_from_loader("dgpar_torch", _loader)
_ext = importlib.util.module_from_spec(_spec)
_loader.exec_module(_ext)


def rk4(U0, analysis, dt, nsteps, ckpt_every=20):
    """U(T) after nsteps RK4 steps of dt from U0, differentiable in U0"""
    return _ext.rk4(U0, int(analysis._h.value), float(dt), int(nsteps), int(ckpt_every))
