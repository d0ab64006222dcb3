"""GPU: the allocation-free time step dg_compute_dt_noalloc, through the C ABI,
against dg_compute_dt on the same contexts: two calls in a row, the second on
a smaller mesh (fewer workgroup minima in the shared scratch array), then the
larger one again; uniform and mixed orders and tetrahedra.  The minimum of the
same per-element values is exact, so the results must be equal bit for bit."""
import ctypes
import sys

import numpy as np
import pytest

from helpers import VORTEX

pytestmark = pytest.mark.gpu


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


class DeviceDouble:
    """one double of device memory owned by the test (hipMalloc / hipFree)"""

    def __init__(self, hip):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(8)) == 0

    def read(self):
        out = ctypes.c_double()
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(ctypes.byref(out), self.p, ctypes.c_size_t(8), 2) == 0  # device to host
        return out.value

    def write(self, v):
        x = ctypes.c_double(v)
        assert self.hip.hipMemcpy(self.p, ctypes.byref(x), ctypes.c_size_t(8), 1) == 0  # host to device

    def free(self):
        self.hip.hipFree(self.p)


def noalloc_dt(dgpar, g, cfl, buf):
    f = dgpar._lib.dg_compute_dt_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    buf.write(-1.0)
    st = f(g._h, cfl, buf.p)
    if st != dgpar.DG_OK:
        raise dgpar.DGError(st, dgpar._last_error(g._h).decode())
    return buf.read()


def contexts(dgpar):
    big = dgpar.DGAnalysis(dgpar.mesh_tri_structured(256, 256, lx=10.0, ly=10.0, jitter=0.1, seed=2024),
                           dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 3)
    big.project(dgpar.FIELD_VORTEX, VORTEX)
    small = dgpar.DGAnalysis(dgpar.mesh_tri_structured(12, 10, jitter=0.1, seed=5),
                             dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,)), 2)
    small.project(dgpar.FIELD_GAUSSIAN)
    return big, small


def test_noalloc_dt_matches_dg_compute_dt_twice_smaller_second():
    import dgpar
    announce("noalloc dt: large mesh, then a smaller one, then the large one again")
    big, small = contexts(dgpar)
    buf = DeviceDouble(dgpar._hip)
    try:
        for g, cfl in ((big, 0.1), (small, 0.3), (big, 0.05)):
            want = g.computeDt(cfl)
            got = noalloc_dt(dgpar, g, cfl, buf)
            assert got == want, (got, want)
    finally:
        buf.free()
        big.close()
        small.close()


def test_noalloc_dt_mixed_orders_and_tetrahedra():
    import dgpar
    announce("noalloc dt: mixed orders and tetrahedra")
    m0 = dgpar.mesh_tri_structured(64, 16, lx=1.0, ly=0.25, jitter=0.1, seed=3, periodic_x=False, periodic_y=False,
                                   side_tags=(dgpar.TAG_WALL, dgpar.TAG_FARFIELD2, dgpar.TAG_WALL, dgpar.TAG_FARFIELD))
    orders = np.random.default_rng(1).integers(1, 5, m0.n_elems).astype(np.int32)
    mixed = dgpar.DGAnalysis(dgpar.Mesh(2, m0.vids, m0.xyz, m0.bc, order=orders),
                             dgpar.EulerLaw(dgpar.FLUX_ROE_EF, 1.4, farfield=(1.0, 0.0, 0.0, 2.5, 0.125, 0.0, 0.0,
                                                                              0.25)), 1)
    mixed.project(dgpar.FIELD_SOD, (0.5,))
    tets = dgpar.DGAnalysis(dgpar.mesh_tet_structured(3, 3, 3), dgpar.MaxwellLaw(), 2)
    tets.project(dgpar.FIELD_CAVITY)
    small = dgpar.DGAnalysis(dgpar.mesh_tri_structured(4, 4), dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 1)
    small.project(dgpar.FIELD_VORTEX, (5.0, 0.5, 0.5, 1.0, 1.0, 1.0))
    buf = DeviceDouble(dgpar._hip)
    try:
        for g in (mixed, tets, small, mixed):
            want = g.computeDt(0.1)
            assert noalloc_dt(dgpar, g, 0.1, buf) == want
    finally:
        buf.free()
        for g in (mixed, tets, small):
            g.close()


def test_noalloc_dt_rejects_bad_arguments():
    import dgpar
    announce("noalloc dt: argument errors")
    g = dgpar.DGAnalysis(dgpar.mesh_tri_structured(4, 4), dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 1)
    f = dgpar._lib.dg_compute_dt_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    assert f(g._h, 0.1, None) == dgpar.DG_ERR_ARG
    assert f(None, 0.1, None) == dgpar.DG_ERR_ARG
    g.close()
