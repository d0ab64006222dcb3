"""CPU emulation of the non-physical recovery kernel's element function
(recover_element_p of dg_mixed.h, run element by element on the host through
tests/cpp/libemu.so) against the oracle's recovery: flags bit for bit and the
recovered state, for uniform and mixed orders; and the check-only form without a
flag array (dg_advance's check after its steps), which must only report."""
import ctypes
import os

import numpy as np
import pytest

import dgpar
from helpers import ROOT, mesh_tri, np_tri
from oracle import pyoracle as po

LIB = os.path.join(ROOT, "tests", "cpp", "libemu.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="emulator not built")
FF = (1.0, 0.2, -0.1, 2.6)


def emu():
    lib = ctypes.CDLL(LIB)
    lib.emu_recover.argtypes = [ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc), ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint8),
                                ctypes.POINTER(ctypes.c_uint)]
    return lib


def case(seed):
    vids, xyz, bc = mesh_tri(10, 10, periodic_x=False, periodic_y=False, jitter=0.1, seed=4,
                             side_tags=(40000, 1, 40000, 1))
    E = vids.shape[0]
    orders = (np.random.default_rng(seed).integers(1, 5, E).astype(np.int32) if seed is not None
              else np.full(E, 3, np.int32))
    rng = np.random.default_rng(7)
    bad = set(rng.choice(E, 8, replace=False).tolist())
    U = []
    for e, p in enumerate(orders):
        b = np.zeros((4, np_tri(p)))
        b[:, 0] = np.array(FF) / np.sqrt(2.0)
        b[:, 1:] = 0.01 * rng.standard_normal((4, np_tri(p) - 1))
        if e in bad:
            b[0, 1:] = 1.5 * rng.standard_normal(np_tri(p) - 1)
        U.append(b.ravel())
    U = np.concatenate(U)
    U[np.argmax(U > 1.8)] = np.nan  # one energy coefficient
    mesh = dgpar.Mesh(2, vids, xyz, bc, order=None if seed is None else orders)
    law = dgpar.Law(dgpar.LAW_EULER2D, dgpar.FLUX_ROE_EF, gamma=1.4, farfield=FF)
    o = po.Oracle(2, vids, xyz, orders, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=FF, bc=bc)
    return mesh, law, o, U


@pytest.mark.parametrize("seed", [None, 3])
@pytest.mark.parametrize("apply", [0, 1])
def test_emulated_recovery_matches_the_oracle(seed, apply):
    lib = emu()
    mesh, law, o, U = case(seed)
    md, ld = mesh.desc(), law.desc()
    V = U.copy()
    flags = np.zeros(mesh.n_elems, np.uint8)
    st = ctypes.c_uint(0)
    assert lib.emu_recover(ctypes.byref(md), ctypes.byref(ld), 3, apply, V.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(st)) == 0
    Uo, fo = o.recover(U)
    assert np.array_equal(flags, fo)
    assert (fo & 1).sum() >= 4 and (fo & 4).sum() == 4
    want = Uo if apply else U
    assert np.array_equal(np.isnan(V), np.isnan(want))
    assert np.array_equal(np.nan_to_num(V), np.nan_to_num(want))
    assert st.value == int(np.bitwise_or.reduce(fo))


def test_emulated_check_without_flag_array_only_reports():
    """dg_advance's check after its steps passes no flag array: the element
    function must record its bits in the status and write nothing"""
    lib = emu()
    mesh, law, o, U = case(None)
    md, ld = mesh.desc(), law.desc()
    V = U.copy()
    st = ctypes.c_uint(0)
    assert lib.emu_recover(ctypes.byref(md), ctypes.byref(ld), 3, 0, V.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           None, ctypes.byref(st)) == 0
    _, fo = o.recover(U)
    assert st.value == int(np.bitwise_or.reduce(fo)) and st.value & 1
    assert np.array_equal(np.nan_to_num(V), np.nan_to_num(U))
