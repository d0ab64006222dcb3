"""Python binding of libdgpar.so (the C ABI declared in include/dg_capi.h).

A thin ctypes layer: numpy arrays in, numpy arrays out.  The GPU path is the
only path -- if the shared library is missing this module raises at import,
and every compute call goes through the HIP kernels.

The class names mirror the reference's driver surface:
    DGAnalysis(mesh, law, order, ...)      DGAnalysis.h:110-135
    .residual(t)                           TimeIntegrator::assembleVolume/Boundary
    .advanceInTime(t, dt, scheme, nsteps)  TimeIntegrator::advanceInTime (TimeIntegrator.h:28)
    .computeDt(cfl)                        DGCell::adaptTimeStep min (DGCell.cc:1155-1163)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdgpar.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libdgpar.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)

_lib = ctypes.CDLL(LIB_PATH)
# the HIP runtime libdgpar.so is linked against (already loaded by the line above)
_hip = ctypes.CDLL("libamdhip64.so.7")
_hip.hipDeviceSynchronize.restype = ctypes.c_int


def device_synchronize():
    """hipDeviceSynchronize on the runtime this library uses."""
    st = _hip.hipDeviceSynchronize()
    if st != 0:
        raise RuntimeError("hipDeviceSynchronize failed with %d" % st)

DG_OK, DG_ERR_ARG, DG_ERR_OOM, DG_ERR_HIP, DG_ERR_NONFINITE, DG_ERR_NONPHYSICAL, DG_ERR_UNSUPPORTED = range(7)
LAW_SCALAR, LAW_EULER2D, LAW_MAXWELL3D, LAW_MAXWELL2D, LAW_BURGERS2D = 0, 1, 2, 3, 4
FLUX_UPWIND, FLUX_LF, FLUX_ROE, FLUX_CENTRAL, FLUX_ROE_EF, FLUX_EXACT = 0, 1, 2, 3, 4, 5
RK_FE, RK_TVD2, RK_TVD3, RK4, RK_TVD2_LTS = 0, 1, 2, 3, 4
PREC_FP64, PREC_FP32 = 0, 1
TAG_FARFIELD, TAG_FARFIELD2, TAG_CURVED_WALL, TAG_CHARACTERISTIC, TAG_WALL = 1, 2, 20000, 30000, 40000
FIELD_GAUSSIAN, FIELD_VORTEX, FIELD_SOD, FIELD_CAVITY = 0, 1, 2, 3

STATUS_NAMES = {0: "DG_OK", 1: "DG_ERR_ARG", 2: "DG_ERR_OOM", 3: "DG_ERR_HIP", 4: "DG_ERR_NONFINITE",
                5: "DG_ERR_NONPHYSICAL", 6: "DG_ERR_UNSUPPORTED"}

_p_i32 = ctypes.POINTER(ctypes.c_int32)
_p_f64 = ctypes.POINTER(ctypes.c_double)


class MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n_elems", ctypes.c_int), ("vids", _p_i32), ("xyz", _p_f64),
                ("bc", _p_i32), ("order", _p_i32)]


class LawDesc(ctypes.Structure):
    _fields_ = [("law", ctypes.c_int), ("flux", ctypes.c_int), ("gamma", ctypes.c_double),
                ("vel", ctypes.c_double * 3), ("farfield", ctypes.c_double * 8)]


class Options(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("precision", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("device", ctypes.c_int), ("reserved", ctypes.c_int * 8)]


class DGError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_vp = ctypes.c_void_p
_create = _sig("dg_create", ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(LawDesc),
                                            ctypes.POINTER(Options), ctypes.POINTER(_vp)])
_destroy = _sig("dg_destroy", ctypes.c_int, [_vp])
_last_error = _sig("dg_last_error", ctypes.c_char_p, [_vp])
_ndof = _sig("dg_ndof", ctypes.c_size_t, [_vp])
_nfaces = _sig("dg_nfaces", ctypes.c_int, [_vp])
_faces = _sig("dg_faces", ctypes.c_int, [_vp, _p_i32, ctypes.c_size_t])
_set_state = _sig("dg_set_state", ctypes.c_int, [_vp, _p_f64, ctypes.c_size_t])
_get_state = _sig("dg_get_state", ctypes.c_int, [_vp, _p_f64, ctypes.c_size_t])
_residual = _sig("dg_residual", ctypes.c_int, [_vp, ctypes.c_double, _p_f64])
_compute_dt = _sig("dg_compute_dt", ctypes.c_int, [_vp, ctypes.c_double, _p_f64])
_advance = _sig("dg_advance", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _p_f64])
_time_advance = _sig("dg_time_advance", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                       ctypes.POINTER(ctypes.c_float)])
_lts_setup = _sig("dg_lts_setup", ctypes.c_int, [_vp, ctypes.c_int, _p_i32])
_adapt_h = _sig("dg_adapt_h", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int8), ctypes.c_int, _p_i32])
_get_mesh = _sig("dg_get_mesh", ctypes.c_int, [_vp, _p_i32, _p_i32, _p_f64, _p_i32, _p_i32])
class SensorOpts(ctypes.Structure):
    _fields_ = [("s_hi", ctypes.c_double * 4), ("s_lo", ctypes.c_double * 4), ("pmin", ctypes.c_int),
                ("pmax", ctypes.c_int)]


_p_u8 = ctypes.POINTER(ctypes.c_uint8)
_set_limiter = _sig("dg_set_limiter", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double])
_limit = _sig("dg_limit", ctypes.c_int, [_vp, _p_u8])
_limiter_flags = _sig("dg_limiter_flags", ctypes.c_int, [_vp, _p_u8])
# a FieldEvaluator (dg_field_fn): fn(x[3], t, val, user)
FIELD_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                            ctypes.c_void_p)
_set_velocity_field = _sig("dg_set_velocity_field", ctypes.c_int, [_vp, FIELD_FN, _vp])
_set_recovery = _sig("dg_set_recovery", ctypes.c_int, [_vp, ctypes.c_int])
_recover = _sig("dg_recover", ctypes.c_int, [_vp, ctypes.c_int, _p_u8])
_recovery_flags = _sig("dg_recovery_flags", ctypes.c_int, [_vp, _p_u8])
_select_orders = _sig("dg_select_orders", ctypes.c_int, [_vp, ctypes.POINTER(SensorOpts), ctypes.c_int, _p_i32])
_get_orders = _sig("dg_get_orders", ctypes.c_int, [_vp, _p_i32])
_write_restart = _sig("dg_write_restart", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_int])
_read_restart = _sig("dg_read_restart", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                                       ctypes.POINTER(ctypes.c_int)])
_sample_lattice = _sig("dg_sample_lattice", ctypes.c_int, [_vp, ctypes.c_int, _p_f64, ctypes.c_size_t])
_norms = _sig("dg_norms", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _p_f64, _p_f64, _p_f64, _p_f64, _p_f64,
                                         _p_f64])
_export_pos = _sig("dg_export_pos", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_int, _p_f64])
_jvp = _sig("dg_residual_jvp", ctypes.c_int, [_vp, _p_f64, _p_f64, _p_f64])
_vjp = _sig("dg_residual_vjp", ctypes.c_int, [_vp, _p_f64, _p_f64, _p_f64])
_adj_fwd = _sig("dg_adjoint_forward", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _p_f64, _p_f64])
_adj_bwd = _sig("dg_adjoint_backward", ctypes.c_int, [_vp, _p_f64, _p_f64])
_adj_timing = _sig("dg_adjoint_timing", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float),
                                                       ctypes.POINTER(ctypes.c_float)])
class Schedule(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int), ("grid", ctypes.c_int), ("reserved", ctypes.c_int * 6)]


TUNE_FORCE = 1
TUNE_MAX_CAND = 16


class TuneReport(ctypes.Structure):
    _fields_ = [("best", Schedule), ("ncand", ctypes.c_int), ("cand", Schedule * TUNE_MAX_CAND),
                ("ms_per_step", ctypes.c_float * TUNE_MAX_CAND), ("from_cache", ctypes.c_int),
                ("key", ctypes.c_char * 256)]


_get_schedule = _sig("dg_get_schedule", ctypes.c_int, [_vp, ctypes.POINTER(Schedule)])
_set_schedule = _sig("dg_set_schedule", ctypes.c_int, [_vp, ctypes.POINTER(Schedule)])
_autotune = _sig("dg_autotune", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_char_p,
                                                ctypes.c_int, ctypes.POINTER(TuneReport)])
_copy_probe = _sig("dg_copy_probe", ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_float)])
_project = _sig("dg_project", ctypes.c_int, [_vp, ctypes.c_int, _p_f64, ctypes.c_double])
_project_fn = _sig("dg_project_fn", ctypes.c_int, [_vp, FIELD_FN, _vp, ctypes.c_double])
_l2_error = _sig("dg_l2_error", ctypes.c_int, [_vp, ctypes.c_int, _p_f64, ctypes.c_double, _p_f64])
_mesh_tri = _sig("dg_mesh_tri_structured", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                  ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _p_i32, _p_i32, _p_f64, _p_i32])


def copy_probe(n_doubles, device=0):
    """Time one streaming device copy of n_doubles (HBM counter reference point)."""
    ms = ctypes.c_float()
    st = _copy_probe(device, n_doubles, ctypes.byref(ms))
    if st != DG_OK:
        raise DGError(st, "dg_copy_probe failed")
    return ms.value


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Mesh:
    """Element-wise mesh description handed to the library (dg_mesh_desc)."""

    def __init__(self, dim, vids, xyz, bc=None, order=None):
        self.dim = int(dim)
        self.vids = np.ascontiguousarray(vids, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.bc = None if bc is None else np.ascontiguousarray(bc, dtype=np.int32)
        self.order = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        self.n_elems = self.vids.shape[0]

    def desc(self):
        return MeshDesc(self.dim, self.n_elems, _ptr(self.vids, _p_i32), _ptr(self.xyz, _p_f64),
                        _ptr(self.bc, _p_i32), _ptr(self.order, _p_i32))


def mesh_tri_structured(nx, ny, x0=0.0, y0=0.0, lx=1.0, ly=1.0, periodic_x=True, periodic_y=True,
                        jitter=0.0, seed=0, side_tags=(0, 0, 0, 0)):
    """Structured, optionally jittered, triangle mesh (dg_mesh_tri_structured)."""
    E = 2 * nx * ny
    vids = np.zeros((E, 3), np.int32)
    xyz = np.zeros((E, 3, 2), np.float64)
    bc = np.zeros((E, 3), np.int32)
    tags = np.asarray(side_tags, np.int32)
    st = _mesh_tri(nx, ny, x0, y0, lx, ly, int(periodic_x), int(periodic_y), jitter, seed, _ptr(tags, _p_i32),
                   _ptr(vids, _p_i32), _ptr(xyz, _p_f64), _ptr(bc, _p_i32))
    if st != DG_OK:
        raise DGError(st, "dg_mesh_tri_structured rejected its arguments")
    return Mesh(2, vids, xyz, bc)


_mesh_tet = _sig("dg_mesh_tet_structured", ctypes.c_int,
                 [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                  _p_i32, _p_f64])


def mesh_tet_structured(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    """Kuhn subdivision: 6 tetrahedra per box (dg_mesh_tet_structured)."""
    E = 6 * nx * ny * nz
    vids = np.zeros((E, 4), np.int32)
    xyz = np.zeros((E, 4, 3), np.float64)
    st = _mesh_tet(nx, ny, nz, lx, ly, lz, _ptr(vids, _p_i32), _ptr(xyz, _p_f64))
    if st != DG_OK:
        raise DGError(st, "dg_mesh_tet_structured rejected its arguments")
    return Mesh(3, vids, xyz)


class GmshInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n_elems", ctypes.c_int), ("n_nodes", ctypes.c_int),
                ("n_quads_split", ctypes.c_int), ("n_periodic_pairs", ctypes.c_int), ("n_tagged", ctypes.c_int)]


_read_gmsh = _sig("dg_mesh_read_gmsh", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(GmshInfo), _p_i32,
                                                       _p_f64, _p_i32, ctypes.c_char_p, ctypes.c_size_t])


def mesh_read_gmsh(path, split_quads=False):
    """A Gmsh version-1 file as a Mesh (dg_mesh_read_gmsh): straight triangles or
    tetrahedra, boundary tags in bc, 510/610 periodic sides joined.  Returns
    (mesh, info dict)."""
    info = GmshInfo()
    err = ctypes.create_string_buffer(512)
    st = _read_gmsh(str(path).encode(), 1 if split_quads else 0, ctypes.byref(info), None, None, None, err, 512)
    if st != DG_OK:
        raise DGError(st, err.value.decode())
    nv = info.dim + 1
    vids = np.zeros((info.n_elems, nv), np.int32)
    xyz = np.zeros((info.n_elems, nv, info.dim), np.float64)
    bc = np.zeros((info.n_elems, nv), np.int32)
    st = _read_gmsh(str(path).encode(), 1 if split_quads else 0, ctypes.byref(info), _ptr(vids, _p_i32),
                    _ptr(xyz, _p_f64), _ptr(bc, _p_i32), err, 512)
    if st != DG_OK:
        raise DGError(st, err.value.decode())
    return Mesh(info.dim, vids, xyz, bc), {f: getattr(info, f) for f, _ in GmshInfo._fields_}


_write_gmsh = _sig("dg_mesh_write_gmsh", ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.c_char_p, ctypes.c_char_p,
                                                         ctypes.c_size_t])


def mesh_write_gmsh(mesh, path):
    """The mesh as a Gmsh version-1 file that mesh_read_gmsh reads back
    (dg_mesh_write_gmsh: boundary tags, periodic sides as 510/610 lines)."""
    err = ctypes.create_string_buffer(512)
    md = mesh.desc()
    st = _write_gmsh(ctypes.byref(md), str(path).encode(), err, 512)
    if st != DG_OK:
        raise DGError(st, err.value.decode())


_lattice_points = _sig("dg_lattice_points", ctypes.c_size_t, [ctypes.c_int, ctypes.c_int])


def lattice_points(dim, nsub):
    """points of the lattice (i, j[, k]) / nsub of a reference simplex"""
    return int(_lattice_points(dim, nsub))


def MaxwellLaw(flux=FLUX_UPWIND):
    return Law(LAW_MAXWELL3D, flux)


class Law:
    """Conservation-law selection (dg_law_desc), mirroring the ConservationLaw subclasses."""

    def __init__(self, law, flux, gamma=1.4, vel=(0.0, 0.0, 0.0), farfield=()):
        self.law, self.flux, self.gamma = law, flux, gamma
        self.vel = tuple(vel) + (0.0,) * (3 - len(vel))
        ff = list(farfield) + [0.0] * (8 - len(farfield))
        self.farfield = tuple(ff[:8])

    def desc(self):
        return LawDesc(self.law, self.flux, self.gamma, (ctypes.c_double * 3)(*self.vel),
                       (ctypes.c_double * 8)(*self.farfield))

    @property
    def nfields(self):
        return {LAW_SCALAR: 1, LAW_EULER2D: 4, LAW_MAXWELL3D: 6, LAW_MAXWELL2D: 3, LAW_BURGERS2D: 1}[self.law]


def Maxwell2dLaw(farfield=(0.0, 0.0, 0.0)):
    """Maxwell2d (MaxwellLaw.cc:21-200): (Hx, Hy, Ez), its own upwind interface state"""
    return Law(LAW_MAXWELL2D, FLUX_UPWIND, farfield=farfield)


def Burgers2DLaw(farfield=(0.0,)):
    """Burgers2D (Burgers2D.cc:22-92): F = (u^2/2, u^2/2), central flux"""
    return Law(LAW_BURGERS2D, FLUX_CENTRAL, farfield=farfield)


def ScalarLaw(vel):
    return Law(LAW_SCALAR, FLUX_UPWIND, vel=vel)


def EulerLaw(flux=FLUX_LF, gamma=1.4, farfield=(1.0, 0.0, 0.0, 2.5)):
    return Law(LAW_EULER2D, flux, gamma=gamma, farfield=farfield)


class DGAnalysis:
    """One context: HBM-resident tables and state of a mesh + law + order."""

    def __init__(self, mesh, law, order, precision=PREC_FP64, use_graph=True, device=0, kernel=0, adj_nd=0):
        self.mesh, self.law, self.order = mesh, law, order
        md, ld = mesh.desc(), law.desc()
        opt = Options(order, precision, int(use_graph), device, (ctypes.c_int * 8)(kernel, adj_nd))
        h = _vp()
        st = _create(ctypes.byref(md), ctypes.byref(ld), ctypes.byref(opt), ctypes.byref(h))
        self._h = h
        if st != DG_OK:
            msg = _last_error(h).decode() if h.value else "dg_create failed"
            if h.value:
                _destroy(h)
            self._h = None
            raise DGError(st, msg)
        self.ndof = _ndof(h)
        self.mesh_E = md.n_elems

    def _check(self, st):
        if st != DG_OK:
            raise DGError(st, _last_error(self._h).decode())

    def close(self):
        if self._h:
            _destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def faces(self):
        n = _nfaces(self._h)
        out = np.zeros((n, 4), np.int32)
        self._check(_faces(self._h, _ptr(out, _p_i32), out.size))
        return out

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        self._check(_set_state(self._h, _ptr(U, _p_f64), U.size))

    def get_state(self):
        out = np.zeros(self.ndof, np.float64)
        self._check(_get_state(self._h, _ptr(out, _p_f64), out.size))
        return out

    def residual(self, t=0.0):
        out = np.zeros(self.ndof, np.float64)
        self._check(_residual(self._h, t, _ptr(out, _p_f64)))
        return out

    def computeDt(self, cfl):
        dt = ctypes.c_double()
        self._check(_compute_dt(self._h, cfl, ctypes.byref(dt)))
        return dt.value

    def advanceInTime(self, t, dt, scheme=RK4, nsteps=1, want_resid=True):
        r = ctypes.c_double()
        self._check(_advance(self._h, scheme, t, dt, nsteps, ctypes.byref(r) if want_resid else None))
        return r.value

    def project(self, field, par=(), t=0.0):
        par = np.ascontiguousarray(list(par) + [0.0] * (8 - len(par)), np.float64)
        self._check(_project(self._h, field, _ptr(par, _p_f64), t))

    def project_fn(self, fn, t=0.0):
        """L2 projection of fn(x, y, z, t) -> Nf values (dg_project_fn; the
        reference's DGAnalysis projecting its far-field FieldEvaluator)"""
        nf = self.law.nfields

        def cb(x, tt, val, user):
            v = fn(x[0], x[1], x[2], tt)
            for k in range(nf):
                val[k] = v[k]
        cfn = FIELD_FN(cb)
        self._check(_project_fn(self._h, cfn, None, t))

    def l2_error(self, field, par=(), t=0.0):
        par = np.ascontiguousarray(list(par) + [0.0] * (8 - len(par)), np.float64)
        err = np.zeros(self.law.nfields)
        self._check(_l2_error(self._h, field, _ptr(par, _p_f64), t, _ptr(err, _p_f64)))
        return err

    def sortCellsBySize(self, max_levels):
        """local time stepping (DGAnalysis::sortCellsBySize + RungeKuttaTVD2adaptive):
        levels by cell size; afterwards computeDt returns the coarsest level's step
        and RK_TVD2_LTS steps advance every level.  Returns the levels (caller order)."""
        lev = np.zeros(self.mesh_E, np.int32)
        self._check(_lts_setup(self._h, max_levels, _ptr(lev, _p_i32)))
        return lev

    lts_setup = sortCellsBySize

    def adaptH(self, marks, max_level=1):
        """h-adaptivity (DGAnalysis::adaptH's mesh operations): cells marked > 0 are
        split into four, families of four children marked < 0 merged; the state
        follows by L2 projection.  Returns the new cell count; refinedMesh() gives
        the new mesh (get_state/set_state use its cells)."""
        marks = np.ascontiguousarray(marks, np.int8)
        if marks.shape != (self.mesh_E,):
            raise ValueError("one mark per current cell (%d)" % self.mesh_E)
        n = ctypes.c_int32()
        self._check(_adapt_h(self._h, marks.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), max_level, ctypes.byref(n)))
        self.mesh_E = n.value
        self.ndof = _ndof(self._h)
        return n.value

    def refinedMesh(self):
        """the current mesh (after adaptH): (vids [E,3], xyz [E,3,2], bc [E,3], levels [E])"""
        n = ctypes.c_int32()
        self._check(_get_mesh(self._h, ctypes.byref(n), None, None, None, None))
        E = n.value
        vids = np.zeros((E, 3), np.int32)
        xyz = np.zeros((E, 3, 2))
        bc = np.zeros((E, 3), np.int32)
        lev = np.zeros(E, np.int32)
        self._check(_get_mesh(self._h, None, _ptr(vids, _p_i32), _ptr(xyz, _p_f64), _ptr(bc, _p_i32),
                              _ptr(lev, _p_i32)))
        return vids, xyz, bc, lev

    def time_advance(self, scheme, dt, nsteps):
        ms = ctypes.c_float()
        self._check(_time_advance(self._h, scheme, dt, nsteps, ctypes.byref(ms)))
        return ms.value

    # ---- launch schedule and tuner ----
    def schedule(self):
        """(shape, grid): tile shape of the tiled kernels (0: 4 waves per 64-element
        tile, 1: 8 waves), workgroups of the persistent mixed-order launch (0: default)"""
        sc = Schedule()
        self._check(_get_schedule(self._h, ctypes.byref(sc)))
        return sc.shape, sc.grid

    def set_schedule(self, shape=0, grid=0):
        self._check(_set_schedule(self._h, ctypes.byref(Schedule(shape, grid))))

    def autotune(self, scheme, dt, steps=5, cache=None, force=False):
        """time every candidate schedule on this mesh and state (restored after),
        keep the fastest; returns {"best", "candidates": [(shape, grid, ms/step)],
        "from_cache", "key"}"""
        rep = TuneReport()
        self._check(_autotune(self._h, scheme, dt, steps, cache.encode() if cache else None,
                              TUNE_FORCE if force else 0, ctypes.byref(rep)))
        return {"best": (rep.best.shape, rep.best.grid),
                "candidates": [(rep.cand[i].shape, rep.cand[i].grid, rep.ms_per_step[i]) for i in range(rep.ncand)],
                "from_cache": bool(rep.from_cache), "key": rep.key.decode()}

    # ---- derivatives (FP64) ----
    def residual_jvp(self, U, dU):
        """dR = (dR/dU)(U) dU, boundary layout"""
        U, dU = np.ascontiguousarray(U, np.float64), np.ascontiguousarray(dU, np.float64)
        out = np.zeros(self.ndof)
        self._check(_jvp(self._h, _ptr(U, _p_f64), _ptr(dU, _p_f64), _ptr(out, _p_f64)))
        return out

    def residual_vjp(self, U, z):
        """w = (dR/dU)(U)^T z, boundary layout"""
        U, z = np.ascontiguousarray(U, np.float64), np.ascontiguousarray(z, np.float64)
        out = np.zeros(self.ndof)
        self._check(_vjp(self._h, _ptr(U, _p_f64), _ptr(z, _p_f64), _ptr(out, _p_f64)))
        return out

    def adjoint_forward(self, dt, nsteps, ckpt_every, target=None):
        """nsteps RK4 steps from the current state with checkpoints; returns the
        density misfit J against target (0.0 without one)"""
        J = ctypes.c_double()
        t = None if target is None else np.ascontiguousarray(target, np.float64)
        self._check(_adj_fwd(self._h, dt, nsteps, ckpt_every, _ptr(t, _p_f64), ctypes.byref(J)))
        return J.value

    def adjoint_backward(self, seed=None):
        """dJ/dU(0) for the seed dJ/dU(T) (None: the gradient of the forward's misfit)"""
        s = None if seed is None else np.ascontiguousarray(seed, np.float64)
        out = np.zeros(self.ndof)
        self._check(_adj_bwd(self._h, _ptr(s, _p_f64), _ptr(out, _p_f64)))
        return out

    def adjoint_timing(self):
        """device milliseconds of the last adjoint_forward and adjoint_backward"""
        f, b = ctypes.c_float(), ctypes.c_float()
        self._check(_adj_timing(self._h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    # ---- limiter and order selection (FP64 form) ----
    def set_limiter(self, enabled, tvb_m=0.0, nu=1.0):
        """limit after every stage of advanceInTime / time_advance"""
        self._check(_set_limiter(self._h, 1 if enabled else 0, tvb_m, nu))

    def limit(self):
        """limit the current state once; per-element flags (bit f: field f limited)"""
        out = np.zeros(self.mesh_E, np.uint8)
        self._check(_limit(self._h, out.ctypes.data_as(_p_u8)))
        return out

    def limiter_flags(self):
        out = np.zeros(self.mesh_E, np.uint8)
        self._check(_limiter_flags(self._h, out.ctypes.data_as(_p_u8)))
        return out

    def set_velocity_field(self, fn):
        """ScalarLaw's velocity FieldEvaluator (ScalarLaw.cc:29-97): fn(x, y) ->
        (vx, vy), evaluated once at every quadrature point (dg_set_velocity_field)"""
        def cb(x, t, val, user):
            v = fn(x[0], x[1])
            val[0], val[1], val[2] = v[0], v[1], 0.0
        cfn = FIELD_FN(cb)
        self._check(_set_velocity_field(self._h, cfn, None))

    # ---- non-physical states (Euler, triangles) ----
    def set_recovery(self, enabled):
        """reduce non-physical elements to their means before every residual"""
        self._check(_set_recovery(self._h, 1 if enabled else 0))

    def recover(self, apply=True):
        """one recovery pass (apply) or check over the current state; per-element
        flags (bit 0 non-physical point, bit 1 non-physical mean, bit 2 NaN/Inf)"""
        out = np.zeros(self.mesh_E, np.uint8)
        self._check(_recover(self._h, 1 if apply else 0, out.ctypes.data_as(_p_u8)))
        return out

    def recovery_flags(self):
        out = np.zeros(self.mesh_E, np.uint8)
        self._check(_recovery_flags(self._h, out.ctypes.data_as(_p_u8)))
        return out

    def select_orders(self, s_hi, s_lo, pmin=1, pmax=4, apply=False):
        """sensor orders per element; with apply the context re-buckets (ndof changes)"""
        o = SensorOpts((ctypes.c_double * 4)(*s_hi), (ctypes.c_double * 4)(*s_lo), pmin, pmax)
        out = np.zeros(self.mesh_E, np.int32)
        self._check(_select_orders(self._h, ctypes.byref(o), 1 if apply else 0, _ptr(out, _p_i32)))
        if apply:
            self.ndof = _ndof(self._h)
        return out

    def orders(self):
        out = np.zeros(self.mesh_E, np.int32)
        self._check(_get_orders(self._h, _ptr(out, _p_i32)))
        return out

    def norms(self, t=0.0, exact_field=-1, par=()):
        """dg_norms: per-field totals (int u) and energies (int u^2), and with an
        exact field the L1, L2 and Linf error norms; a dict of arrays"""
        nf = self.law.nfields
        out = {k: np.zeros(nf) for k in ("totals", "energy")}
        if exact_field >= 0:
            out.update({k: np.zeros(nf) for k in ("l1", "l2", "linf")})
        pv = np.ascontiguousarray(list(par) + [0.0] * (8 - len(par)), np.float64)
        g = lambda k: _ptr(out[k], _p_f64) if k in out else None  # noqa: E731
        self._check(_norms(self._h, t, exact_field, _ptr(pv, _p_f64), g("l1"), g("l2"), g("linf"), g("totals"),
                           g("energy")))
        return out

    # ---- checkpoint / restart and state export ----
    def write_restart(self, path, t=0.0, iteration=0):
        """the data stream of DGAnalysis::write (DGPref.cc:712-740)"""
        self._check(_write_restart(self._h, str(path).encode(), t, iteration))

    def read_restart(self, path):
        """DGAnalysis::read (DGPref.cc:742-768); returns (t, iteration)"""
        t, it = ctypes.c_double(), ctypes.c_int()
        self._check(_read_restart(self._h, str(path).encode(), ctypes.byref(t), ctypes.byref(it)))
        self.ndof = _ndof(self._h)
        return t.value, it.value

    def write(self, mesh_path, data_path, t=0.0, iteration=0):
        """DGAnalysis::write(mesh, data): the mesh as a Gmsh file and the data stream"""
        mesh_write_gmsh(self.mesh, mesh_path)
        self.write_restart(data_path, t, iteration)

    def sample_lattice(self, nsub):
        """every field at the lattice points of every element: [E][Nf][points]"""
        npts = lattice_points(self.mesh.dim, nsub)
        out = np.zeros((self.mesh_E, self.law.nfields, max(npts, 1)), np.float64)
        self._check(_sample_lattice(self._h, nsub, _ptr(out, _p_f64), out.size if npts else 0))
        return out

    def export_pos(self, prefix, iteration=0, nsub=0, t=0.0, exact_field=-1, exact_par=()):
        """Gmsh .pos views of the fields of interest (DGAnalysis::exportGmsh)"""
        par = np.ascontiguousarray(list(exact_par) + [0.0] * (8 - len(exact_par)), np.float64)
        self._check(_export_pos(self._h, str(prefix).encode(), iteration, nsub, t, exact_field, _ptr(par, _p_f64)))
