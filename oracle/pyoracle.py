"""ctypes access to the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY:
used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")

LAW_SCALAR, LAW_EULER2D, LAW_MAXWELL3D, LAW_MAXWELL2D, LAW_BURGERS2D = 0, 1, 2, 3, 4
FLUX_UPWIND, FLUX_LF, FLUX_ROE, FLUX_CENTRAL, FLUX_ROE_EF, FLUX_EXACT = 0, 1, 2, 3, 4, 5
RK_FE, RK_TVD2, RK_TVD3, RK4 = 0, 1, 2, 3

_p_i32 = ctypes.POINTER(ctypes.c_int32)
_p_f64 = ctypes.POINTER(ctypes.c_double)


class _Mesh(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n_elems", ctypes.c_int), ("vids", _p_i32), ("xyz", _p_f64),
                ("bc", _p_i32), ("order", _p_i32)]


class _Law(ctypes.Structure):
    _fields_ = [("law", ctypes.c_int), ("flux", ctypes.c_int), ("gamma", ctypes.c_double),
                ("vel", ctypes.c_double * 3), ("farfield", ctypes.c_double * 8)]


# a FieldEvaluator: fn(x[3], t, val[3], user)
FIELD_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                            ctypes.c_void_p)


def field_fn(py):
    """a FieldEvaluator from a Python function (x, y) -> (vx, vy)"""
    def cb(x, t, val, user):
        v = py(x[0], x[1])
        val[0], val[1], val[2] = v[0], v[1], 0.0
    return FIELD_FN(cb)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    lib.or_create.restype = vp
    lib.or_create.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Law)]
    lib.or_destroy.argtypes = [vp]
    lib.or_nfields.argtypes = [vp]
    lib.or_ndof.restype = ctypes.c_long
    lib.or_ndof.argtypes = [vp]
    lib.or_nfaces.argtypes = [vp]
    lib.or_faces.argtypes = [vp, _p_i32]
    lib.or_residual.argtypes = [vp, _p_f64, ctypes.c_double, _p_f64, ctypes.c_int]
    lib.or_residual_abs.argtypes = [vp, _p_f64, ctypes.c_double, _p_f64, ctypes.c_int]
    lib.or_recover.argtypes = [vp, _p_f64, ctypes.POINTER(ctypes.c_uint8)]
    lib.or_set_recovery.argtypes = [vp, ctypes.c_int]
    lib.or_step.restype = ctypes.c_double
    lib.or_step.argtypes = [vp, ctypes.c_int, _p_f64, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.or_dt.restype = ctypes.c_double
    lib.or_dt.argtypes = [vp, _p_f64, ctypes.c_double]
    lib.or_detj.argtypes = [vp, _p_f64]
    lib.or_basis.argtypes = [ctypes.c_int, ctypes.c_int, _p_f64, _p_f64, _p_f64]
    lib.or_project.argtypes = [vp, ctypes.c_int, _p_f64, ctypes.c_double, _p_f64]
    lib.or_exact.argtypes = [ctypes.c_int, _p_f64, ctypes.POINTER(_Law), _p_f64, ctypes.c_double, _p_f64]
    lib.or_l2_error.argtypes = [vp, _p_f64, ctypes.c_int, _p_f64, ctypes.c_double, _p_f64]
    lib.or_limit.argtypes = [vp, _p_f64, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_uint8)]
    lib.or_set_limiter.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.or_sensor.argtypes = [vp, _p_f64, _p_f64, _p_f64, ctypes.c_int, ctypes.c_int, _p_i32]
    lib.or_set_velocity_field.argtypes = [vp, FIELD_FN, vp]
    lib.or_project_fn.argtypes = [vp, FIELD_FN, vp, ctypes.c_double, _p_f64]
    lib.or_riemann.argtypes = [ctypes.POINTER(_Law), _p_f64, _p_f64, _p_f64, _p_f64]
    lib.or_riemann_exact.argtypes = [ctypes.c_double, _p_f64, _p_f64, ctypes.c_double, _p_f64]
    lib.or_lts_levels.argtypes = [vp, ctypes.c_int, _p_i32]
    lib.or_lts_dt.restype = ctypes.c_double
    lib.or_lts_dt.argtypes = [vp, _p_f64, ctypes.c_double, _p_i32]
    lib.or_lts_step.argtypes = [vp, _p_i32, _p_f64, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.or_set_tri_basis.argtypes = [ctypes.c_int, _p_f64]
    lib.or_set_tri_rule.argtypes = [ctypes.c_int, ctypes.c_int, _p_f64]
    return lib


_lib = _load()


def set_tri_basis(C):
    C = np.ascontiguousarray(C, np.float64)
    assert _lib.or_set_tri_basis(C.shape[0], _p(C, _p_f64)) == 0


def set_tri_rule(degree, pts):
    pts = np.ascontiguousarray(pts, np.float64)
    assert _lib.or_set_tri_rule(degree, pts.shape[0], _p(pts, _p_f64)) == 0


def reset_tables():
    _lib.or_reset_tables()


def make_law(law, flux, gamma=1.4, vel=(0, 0, 0), farfield=()):
    ff = list(farfield) + [0.0] * (8 - len(farfield))
    v = list(vel) + [0.0] * (3 - len(vel))
    return _Law(law, flux, gamma, (ctypes.c_double * 3)(*v), (ctypes.c_double * 8)(*ff[:8]))


def riemann(law, n, uL, uR):
    """the oracle's numerical flux Fhat(uL, uR).n at one face point; law from make_law"""
    n = np.ascontiguousarray(np.resize(np.asarray(n, np.float64), 3))
    uL = np.ascontiguousarray(uL, np.float64)
    uR = np.ascontiguousarray(uR, np.float64)
    out = np.zeros(max(len(uL), 6))
    _lib.or_riemann(ctypes.byref(law), _p(n, _p_f64), _p(uL, _p_f64), _p(uR, _p_f64), _p(out, _p_f64))
    return out[:len(uL)]


def riemann_exact(gamma, wl, wr, s):
    """the exact 1D Riemann solution (rho, u, p) at x/t = s for primitive states wl, wr"""
    wl = np.ascontiguousarray(wl, np.float64)
    wr = np.ascontiguousarray(wr, np.float64)
    out = np.zeros(3)
    _lib.or_riemann_exact(gamma, _p(wl, _p_f64), _p(wr, _p_f64), s, _p(out, _p_f64))
    return out


def exact(field, x, t, par=(), law=None):
    """an exact field's conserved values at point x and time t"""
    law = law if law is not None else make_law(LAW_EULER2D, FLUX_LF)
    par = np.ascontiguousarray(list(par) + [0.0] * (8 - len(par)), np.float64)
    x = np.ascontiguousarray(np.resize(np.asarray(x, np.float64), 3))
    out = np.zeros(6)
    _lib.or_exact(field, _p(par, _p_f64), ctypes.byref(law), _p(x, _p_f64), t, _p(out, _p_f64))
    return out


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def basis(dim, p, x):
    n = (p + 1) * (p + 2) // 2 if dim == 2 else (p + 1) * (p + 2) * (p + 3) // 6
    x = np.ascontiguousarray(np.resize(np.asarray(x, np.float64), 3))
    phi = np.zeros(n)
    dphi = np.zeros(n * dim)
    if _lib.or_basis(dim, p, _p(x, _p_f64), _p(phi, _p_f64), _p(dphi, _p_f64)) != 0:
        raise ValueError("bad basis request")
    return phi, dphi.reshape(n, dim)


class Oracle:
    def __init__(self, dim, vids, xyz, order, law, flux, gamma=1.4, vel=(0, 0, 0), farfield=(), bc=None):
        self.vids = np.ascontiguousarray(vids, np.int32)
        self.xyz = np.ascontiguousarray(xyz, np.float64)
        E = self.vids.shape[0]
        if np.isscalar(order):
            order = np.full(E, order, np.int32)
        self.order = np.ascontiguousarray(order, np.int32)
        self.bc = None if bc is None else np.ascontiguousarray(bc, np.int32)
        m = _Mesh(dim, E, _p(self.vids, _p_i32), _p(self.xyz, _p_f64), _p(self.bc, _p_i32), _p(self.order, _p_i32))
        ff = list(farfield) + [0.0] * (8 - len(farfield))
        v = list(vel) + [0.0] * (3 - len(vel))
        self._law = _Law(law, flux, gamma, (ctypes.c_double * 3)(*v), (ctypes.c_double * 8)(*ff[:8]))
        self._h = _lib.or_create(ctypes.byref(m), ctypes.byref(self._law))
        if not self._h:
            raise ValueError("oracle rejected the mesh")
        self.ndof = _lib.or_ndof(self._h)
        self.nfields = _lib.or_nfields(self._h)
        self.E = E

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.or_destroy(self._h)
            self._h = None

    def faces(self):
        n = _lib.or_nfaces(self._h)
        out = np.zeros((n, 4), np.int32)
        _lib.or_faces(self._h, _p(out, _p_i32))
        return out

    def detj(self):
        out = np.zeros(self.E)
        _lib.or_detj(self._h, _p(out, _p_f64))
        return out

    def residual(self, U, t=0.0, nthreads=1):
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros(self.ndof)
        flag = _lib.or_residual(self._h, _p(U, _p_f64), t, _p(R, _p_f64), nthreads)
        return R, flag

    def residual_scale(self, U, t=0.0, nthreads=1):
        """per DOF the sum of the magnitudes of the residual's terms"""
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros(self.ndof)
        _lib.or_residual_abs(self._h, _p(U, _p_f64), t, _p(R, _p_f64), nthreads)
        return R

    def recover(self, U):
        """non-physical recovery: reduced copy of U and per-element flags
        (1 reduced, 2 reduced with a non-physical mean)"""
        U = np.array(U, np.float64, copy=True)
        flags = np.zeros(self.E, np.uint8)
        _lib.or_recover(self._h, _p(U, _p_f64), flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return U, flags

    def set_velocity_field(self, py):
        """ScalarLaw's velocity FieldEvaluator: py(x, y) -> (vx, vy)"""
        self._velfn = field_fn(py)  # kept alive with the context
        if _lib.or_set_velocity_field(self._h, self._velfn, None) != 0:
            raise ValueError("a velocity field serves the scalar law")

    def set_recovery(self, on):
        """recovery before every residual of step()"""
        _lib.or_set_recovery(self._h, 1 if on else 0)

    def step(self, U, scheme, t, dt, nthreads=1):
        U = np.array(U, np.float64, copy=True)
        r = _lib.or_step(self._h, scheme, _p(U, _p_f64), t, dt, nthreads)
        return U, r

    def dt(self, U, cfl):
        U = np.ascontiguousarray(U, np.float64)
        return _lib.or_dt(self._h, _p(U, _p_f64), cfl)

    def lts_levels(self, max_levels):
        """local time stepping: (levels per element, number of levels)"""
        lev = np.zeros(self.E, np.int32)
        n = _lib.or_lts_levels(self._h, max_levels, _p(lev, _p_i32))
        if n < 1:
            raise ValueError("local time stepping serves triangles")
        return lev, n

    def lts_dt(self, U, cfl, levels):
        U = np.ascontiguousarray(U, np.float64)
        levels = np.ascontiguousarray(levels, np.int32)
        return _lib.or_lts_dt(self._h, _p(U, _p_f64), cfl, _p(levels, _p_i32))

    def lts_step(self, U, levels, h, t=0.0, nthreads=1):
        """one local-time-stepping step of size h (a new copy of U)"""
        U = np.array(U, np.float64, copy=True)
        levels = np.ascontiguousarray(levels, np.int32)
        if _lib.or_lts_step(self._h, _p(levels, _p_i32), _p(U, _p_f64), t, h, nthreads) != 0:
            raise ValueError("local time stepping serves triangles")
        return U

    def project(self, field, par=(), t=0.0):
        par = np.ascontiguousarray(list(par) + [0.0] * (8 - len(par)), np.float64)
        U = np.zeros(self.ndof)
        if _lib.or_project(self._h, field, _p(par, _p_f64), t, _p(U, _p_f64)) != 0:
            raise ValueError("bad field")
        return U

    def project_fn(self, py, t=0.0):
        """L2 projection of py(x, y, z, t) -> Nf values"""
        def cb(x, tt, val, user):
            v = py(x[0], x[1], x[2], tt)
            for k in range(len(v)):
                val[k] = v[k]
        f = FIELD_FN(cb)
        U = np.zeros(self.ndof)
        _lib.or_project_fn(self._h, f, None, t, _p(U, _p_f64))
        return U

    def limit(self, U, tvb_m=0.0, nu=1.0):
        """limited copy of U and the per-element flags (bit f: field f limited)"""
        U = np.array(U, np.float64, copy=True)
        flags = np.zeros(self.E, np.uint8)
        if _lib.or_limit(self._h, _p(U, _p_f64), tvb_m, nu, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))):
            raise ValueError("limiter is built for triangles")
        return U, flags

    def set_limiter(self, on, tvb_m=0.0, nu=1.0):
        """limit after every stage of step()"""
        _lib.or_set_limiter(self._h, 1 if on else 0, tvb_m, nu)

    def sensor(self, U, s_hi, s_lo, pmin=1, pmax=4):
        U = np.ascontiguousarray(U, np.float64)
        hi = np.ascontiguousarray(s_hi, np.float64)
        lo = np.ascontiguousarray(s_lo, np.float64)
        out = np.zeros(self.E, np.int32)
        _lib.or_sensor(self._h, _p(U, _p_f64), _p(hi, _p_f64), _p(lo, _p_f64), pmin, pmax, _p(out, _p_i32))
        return out

    def l2_error(self, U, field, par=(), t=0.0):
        par = np.ascontiguousarray(list(par) + [0.0] * (8 - len(par)), np.float64)
        U = np.ascontiguousarray(U, np.float64)
        err = np.zeros(self.nfields)
        _lib.or_l2_error(self._h, _p(U, _p_f64), field, _p(par, _p_f64), t, _p(err, _p_f64))
        return err
