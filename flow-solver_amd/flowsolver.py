"""Python host mirror of the residual-and-implicit-update path.

Thin ctypes layer over libflowsolver.so (the C-ABI in include/fs_capi.h).
It mirrors the reference's driver-level objects: ``Params`` is PARAMS
(include/params.h:19-76), ``Mesh`` is the preprocessed GRID/SYS_MEM pair
(mesh_io.C, maps.C, geometry.C, cv_calc.C, init_memory.C), ``Solver`` holds
the device-resident state and exposes the reference's per-step functions
(solve.C:44 solve, residual.C:40/359, linear_algebra.C:122, ...).

There is no CPU fallback: constructing a Solver without the built library or
without a HIP device raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflowsolver.so")

FS_PREC_F64 = 0
FS_PREC_MIXED = 1
FS_PREC_MIXED16 = 2

_MESH_ARRAYS = [
    "x", "tri", "nn_map", "ncsn", "csn", "nnsn", "nsn", "esn", "edges", "c2e", "sub_edge", "sub_elem", "sub_n",
    "sub_xm", "sub_xg", "el_cent", "bedge", "bedge_n", "bedge_xg", "cv_area", "moments", "edge_weight", "ia", "iau",
    "ja", "color", "color_off", "color_rows", "st_off", "st_nodes", "st_W", "bbc", "bseg_off", "bseg", "bedge_gn", "quad",
]
_FIELDS = ["Q", "nQ", "dQ", "R", "RHS", "grad", "phi", "hess", "pv", "ev_con", "dt", "LHS", "LHS32"]
_TUNING = ["gs_block", "gs_lanes", "gs32_block", "gs32_pre", "node_block", "face_block", "cell_block", "deriv_lds",
           "gs32_scaled", "node_fused", "gs32_dpp"]


def _set_tuning(fn, h, cur, kw):
    v = np.array([kw.get(k, cur[k]) for k in _TUNING], dtype=np.int32)
    _check(fn(h, v.ctypes.data))


def _tuning(fn, h):
    v = np.zeros(len(_TUNING), dtype=np.int32)
    fc = C.c_int()
    _check(fn(h, v.ctypes.data, C.byref(fc)))
    out = dict(zip(_TUNING, (int(x) for x in v)))
    out["from_cache"] = bool(fc.value)
    return out


class FlowSolverError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fs error %d: %s" % (code, msg))
        self.code = code


def _pt(v):
    a = (C.c_double * 2)(float(v[0]), float(v[1]))
    return a


def curved_midpoint(bc, xL, xR):
    """curved_boundary_midpoint (curved_boundaries.C:91)."""
    out = (C.c_double * 2)()
    _check(lib().fs_curved_midpoint(bc, _pt(xL), _pt(xR), out))
    return (out[0], out[1])


def curved_gauss_points(bc, xL, xR):
    """curved_boundary_gauss_points (curved_boundaries.C:345): 3 (x, y) points."""
    out = (C.c_double * 6)()
    _check(lib().fs_curved_gauss_points(bc, _pt(xL), _pt(xR), out))
    return [(out[2 * k], out[2 * k + 1]) for k in range(3)]


def curved_normal(bc, x):
    """curved_boundary_normal_vector (curved_boundaries.C:701)."""
    nx, ny = C.c_double(), C.c_double()
    _check(lib().fs_curved_normal(bc, _pt(x), C.byref(nx), C.byref(ny)))
    return (nx.value, ny.value)


def curved_arclength(bc, xL, xR):
    """curved_boundary_arclength (curved_boundaries.C:1009)."""
    ln = C.c_double()
    _check(lib().fs_curved_arclength(bc, _pt(xL), _pt(xR), C.byref(ln)))
    return ln.value


def hbm_probe(nbytes=4 << 30, reps=5):
    """Achievable HBM bandwidth (GB/s) of the current device: (streaming read,
    copy counting both directions)."""
    r, c = C.c_double(), C.c_double()
    _check(lib().fs_hbm_probe(nbytes, reps, C.byref(r), C.byref(c)))
    return r.value, c.value


class Params(C.Structure):
    """PARAMS: the 24 run controls of euler.params (params.C:44-217)."""
    _fields_ = [
        ("nsteps", C.c_int), ("iter", C.c_int), ("jufreq", C.c_int), ("CFL", C.c_double), ("ramp", C.c_int),
        ("gamma", C.c_double), ("alpha", C.c_double), ("mach", C.c_double), ("subits", C.c_int),
        ("unsteady", C.c_int), ("tacc", C.c_int), ("mintime", C.c_double), ("method", C.c_int), ("order", C.c_int),
        ("foits", C.c_int), ("soits", C.c_int), ("toits", C.c_int), ("isrestart", C.c_int), ("limit", C.c_int),
        ("reorder", C.c_int), ("ic", C.c_int), ("p_or_cv", C.c_int), ("start_post_pro", C.c_int),
        ("post_pro_freq", C.c_int),
    ]

    @classmethod
    def make(cls, **kw):
        p = cls()
        lib().fs_params_default(C.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        return p

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nn", "ng", "ne", "nsub", "nbe", "nnz", "ntri", "nquad", "nb", "ncolor",
                                       "nstencil", "prepared", "order_prepared")]


class BatchInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ninst", "nn", "ne", "nsub", "nbe", "nnz", "ncolor")] + [
        ("device_bytes", C.c_int64)]


INST_RUNNING, INST_CONVERGED, INST_MAX_STEPS, INST_FAILED, INST_PAUSED, INST_CANCELLED = range(6)
FALLBACK_LIMITER, FALLBACK_FIRST_ORDER, FALLBACK_STALL = 1, 2, 4


class ServerConfig(C.Structure):
    _fields_ = [("device_budget", C.c_int64), ("precision", C.c_int), ("max_steps", C.c_int), ("tol", C.c_double),
                ("prep_threads", C.c_int), ("use_graph", C.c_int), ("max_recoveries", C.c_int),
                ("checkpoint_every", C.c_int), ("max_steps_per_sqrt_node", C.c_double), ("limiter_freeze", C.c_int),
                ("limiter_fallback", C.c_int), ("first_order_fallback", C.c_int),
                ("stall_steps_per_sqrt_node", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("steps", C.c_int), ("drop", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("device_bytes", C.c_int64), ("nn", C.c_int), ("ng", C.c_int), ("forces", C.c_double * 2),
                ("recoveries", C.c_int), ("fallbacks", C.c_int), ("max_steps", C.c_int), ("r0", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("step_ms", C.c_double), ("edge_ms", C.c_double), ("node_ms", C.c_double), ("gs_ms", C.c_double),
                ("lu_ms", C.c_double), ("other_ms", C.c_double), ("flux_evals", C.c_int64)]


_lib = None


def lib():
    """Load libflowsolver.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlowSolverError(-7, "libflowsolver.so not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    sig = {
        "fs_last_error": (C.c_char_p, []),
        "fs_device_count": (i, [C.POINTER(i)]),
        "fs_hbm_probe": (i, [i64, i, C.POINTER(d), C.POINTER(d)]),
        "fs_params_default": (None, [vp]),
        "fs_params_read": (i, [C.c_char_p, vp]),
        "fs_params_write": (i, [C.c_char_p, vp]),
        "fs_mesh_read": (i, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
        "fs_mesh_write": (i, [vp, C.c_char_p, C.c_char_p]),
        "fs_mesh_create_bump": (i, [i, i, d, C.c_uint64, C.POINTER(vp)]),
        "fs_mesh_create_airfoil": (i, [i, i, d, d, d, C.c_uint64, C.POINTER(vp)]),
        "fs_mesh_create_airfoil_split": (i, [i, i, d, d, d, C.c_uint64, i, i, C.POINTER(vp)]),
        "fs_mesh_create_annulus": (i, [i, i, d, d, d, C.c_uint64, i, i, C.POINTER(vp)]),
        "fs_curved_midpoint": (i, [i, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
        "fs_curved_gauss_points": (i, [i, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
        "fs_curved_normal": (i, [i, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
        "fs_curved_arclength": (i, [i, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
        "fs_mesh_prepare": (i, [vp, vp]),
        "fs_mesh_get_info": (i, [vp, vp]),
        "fs_mesh_get_array": (i, [vp, i, C.POINTER(vp), C.POINTER(i64), C.POINTER(i)]),
        "fs_mesh_destroy": (None, [vp]),
        "fs_solver_create": (i, [vp, vp, i, C.POINTER(vp)]),
        "fs_solver_destroy": (None, [vp]),
        "fs_solver_init_freestream": (i, [vp]),
        "fs_solver_set": (i, [vp, i, vp, i64]),
        "fs_solver_get": (i, [vp, i, vp, i64]),
        "fs_build_matrix": (i, [vp, i]),
        "fs_gradient": (i, [vp]),
        "fs_derivatives": (i, [vp]),
        "fs_limiter": (i, [vp, i]),
        "fs_interior_residual": (i, [vp, i, i]),
        "fs_boundary_residual": (i, [vp, i]),
        "fs_linear_solve": (i, [vp, C.POINTER(d)]),
        "fs_solve": (i, [vp, i, C.POINTER(d), C.POINTER(d)]),
        "fs_body_forces": (i, [vp, i, vp]),
        "fs_surface_cp": (i, [vp, i, vp, vp, i64, C.POINTER(i64)]),
        "fs_write_surface_cp": (i, [vp, C.c_char_p]),
        "fs_batch_body_forces": (i, [vp, vp]),
        "fs_batch_set_compaction": (i, [vp, i]),
        "fs_batch_set_recovery": (i, [vp, i, i, i, i]),
        "fs_batch_inst_recovery": (i, [vp, i, C.POINTER(i), C.POINTER(i)]),
        "fs_batch_set_limiter_freeze": (i, [vp, i]),
        "fs_batch_set_stall": (i, [vp, d]),
        "fs_roe_flux_batch": (i, [i, d, vp, vp, vp, vp, vp]),
        "fs_solve_steady": (i, [vp, i, d, C.POINTER(i), C.POINTER(d)]),
        "fs_solver_time_steps": (i, [vp, i, i, vp]),
        "fs_solver_use_graph": (i, [vp, i]),
        "fs_batch_create": (i, [i, C.POINTER(vp), vp, i, C.POINTER(vp)]),
        "fs_batch_destroy": (None, [vp]),
        "fs_batch_get_info": (i, [vp, vp]),
        "fs_batch_init_freestream": (i, [vp]),
        "fs_batch_use_graph": (i, [vp, i]),
        "fs_batch_defer_k2": (i, [vp, i]),
        "fs_batch_set_active": (i, [vp, i, i]),
        "fs_batch_step": (i, [vp, vp, vp]),
        "fs_batch_solve_steady": (i, [vp, i, d, vp, vp, vp]),
        "fs_batch_inst_status": (i, [vp, i, C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
        "fs_batch_get": (i, [vp, i, i, vp, i64]),
        "fs_batch_set": (i, [vp, i, i, vp, i64]),
        "fs_batch_time_steps": (i, [vp, i, vp]),
        "fs_mesh_footprint": (i, [vp, i, C.POINTER(i64)]),
        "fs_server_config_default": (None, [vp]),
        "fs_server_create": (i, [vp, C.POINTER(vp)]),
        "fs_server_destroy": (None, [vp]),
        "fs_server_submit": (i, [vp, vp, vp, C.POINTER(i64)]),
        "fs_server_wait": (i, [vp, d, vp, C.POINTER(i)]),
        "fs_server_get_state": (i, [vp, i64, vp, i64]),
        "fs_server_release": (i, [vp, i64]),
        "fs_server_cancel": (i, [vp, i64]),
        "fs_server_counts": (i, [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "fs_server_clock": (d, [vp]),
        "fs_restart_write": (i, [vp, C.c_char_p, vp, i64, i, d, i]),
        "fs_mesh_restart_info": (i, [vp, C.POINTER(i), C.POINTER(i), C.POINTER(d), C.POINTER(i)]),
        "fs_solver_init_restart": (i, [vp]),
        "fs_solver_autotune": (i, [vp, C.POINTER(d)]),
        "fs_solver_get_tuning": (i, [vp, vp, C.POINTER(i)]),
        "fs_batch_autotune": (i, [vp, C.POINTER(d)]),
        "fs_batch_get_tuning": (i, [vp, vp, C.POINTER(i)]),
        "fs_solver_set_tuning": (i, [vp, vp]),
        "fs_batch_set_tuning": (i, [vp, vp]),
        "fs_mesh_create": (i, [i, vp, i, vp, i, vp, i, vp, vp, vp, C.POINTER(vp)]),
        "fs_roe_flux_batch_form": (i, [i, i, d, vp, vp, vp, vp, vp]),
        "fs_state_norms": (i, [i64, vp, vp, vp, vp]),
        "fs_eigenvalue_contribution": (i, [vp]),
        "fs_set_boundary_values": (i, [vp]),
        "fs_state_norms_noalloc": (i, [i64, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise FlowSolverError(rc, lib().fs_last_error().decode())


def device_count():
    n = C.c_int(0)
    _check(lib().fs_device_count(C.byref(n)))
    return n.value


def read_params(path):
    p = Params()
    _check(lib().fs_params_read(path.encode(), C.byref(p)))
    return p


def write_params(path, p):
    _check(lib().fs_params_write(path.encode(), C.byref(p)))


class Mesh:
    """A grid in the reference format plus its preprocessed maps/geometry."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def bump(cls, nx, ny, jitter=0.2, seed=1):
        h = C.c_void_p()
        _check(lib().fs_mesh_create_bump(nx, ny, jitter, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def airfoil(cls, nwall, nrad, h0=1e-3, rmax=30.0, jitter=0.1, seed=3):
        h = C.c_void_p()
        _check(lib().fs_mesh_create_airfoil(nwall, nrad, h0, rmax, jitter, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, x, tri, quad, bseg_off, bseg, bbc):
        """A mesh from 0-based arrays: node coordinates (nn x 2), triangles
        (ntri x 3), quads (nquad x 4), boundary segment offsets (nb + 1),
        segments (nseg x 2) and boundary codes (nb): the grid file's contents."""
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1)
        quad = np.ascontiguousarray(quad, dtype=np.int32).reshape(-1)
        bseg_off = np.ascontiguousarray(bseg_off, dtype=np.int32)
        bseg = np.ascontiguousarray(bseg, dtype=np.int32).reshape(-1)
        bbc = np.ascontiguousarray(bbc, dtype=np.int32)
        h = C.c_void_p()
        _check(lib().fs_mesh_create(len(x) // 2, x.ctypes.data, len(tri) // 3, tri.ctypes.data, len(quad) // 4,
                                    quad.ctypes.data, len(bbc), bseg_off.ctypes.data, bseg.ctypes.data,
                                    bbc.ctypes.data, C.byref(h)))
        return cls(h.value)

    @classmethod
    def airfoil_split(cls, nwall, nrad, upper_bc=17, lower_bc=19, h0=1e-3, rmax=30.0, jitter=0.1, seed=3):
        """O-mesh with separate upper/lower walls: codes 17/19 are the analytic
        NACA 0012 surfaces, 1/1 polyline walls."""
        h = C.c_void_p()
        _check(lib().fs_mesh_create_airfoil_split(nwall, nrad, h0, rmax, jitter, seed, upper_bc, lower_bc,
                                                  C.byref(h)))
        return cls(h.value)

    @classmethod
    def annulus(cls, ntheta, nrad, r_in=1.0, r_out=300.0, bc_in=11, bc_out=22, jitter=0.1, seed=5):
        """Annulus of geometrically spaced rings: by default the unit-circle
        wall (code 11) inside the R=300 farfield circle (code 22)."""
        h = C.c_void_p()
        _check(lib().fs_mesh_create_annulus(ntheta, nrad, r_in, r_out, jitter, seed, bc_in, bc_out, C.byref(h)))
        return cls(h.value)

    @classmethod
    def read(cls, grid_path, bc_path):
        h = C.c_void_p()
        _check(lib().fs_mesh_read(grid_path.encode(), bc_path.encode() if bc_path else None, C.byref(h)))
        return cls(h.value)

    def write(self, grid_path, bc_path):
        _check(lib().fs_mesh_write(self._h, grid_path.encode(), bc_path.encode() if bc_path else None))

    def prepare(self, params):
        _check(lib().fs_mesh_prepare(self._h, C.byref(params)))
        return self

    def info(self):
        mi = MeshInfo()
        _check(lib().fs_mesh_get_info(self._h, C.byref(mi)))
        return {k: getattr(mi, k) for k, _ in mi._fields_}

    def array(self, name):
        """Copy of a preprocessed array (0-based indices)."""
        which = _MESH_ARRAYS.index(name)
        data, cnt, dt = C.c_void_p(), C.c_int64(), C.c_int()
        _check(lib().fs_mesh_get_array(self._h, which, C.byref(data), C.byref(cnt), C.byref(dt)))
        n = cnt.value
        if n == 0:
            return np.zeros(0, dtype=np.int32 if dt.value == 0 else np.float64)
        ct = C.c_int32 if dt.value == 0 else C.c_double
        buf = C.cast(data, C.POINTER(ct * n)).contents
        return np.frombuffer(buf, dtype=np.int32 if dt.value == 0 else np.float64).copy()

    def restart_info(self):
        """(has_restart, last step, last CFL, has ghost states) of the grid file read."""
        h, it, cfl, g = C.c_int(), C.c_int(), C.c_double(), C.c_int()
        _check(lib().fs_mesh_restart_info(self._h, C.byref(h), C.byref(it), C.byref(cfl), C.byref(g)))
        return bool(h.value), it.value, cfl.value, bool(g.value)

    def write_restart(self, grid_path, Q, citer, cfl, ghosts=True):
        """Restart file of this prepared mesh (solver numbering) with state Q."""
        q = np.ascontiguousarray(Q, dtype=np.float64).ravel()
        _check(lib().fs_restart_write(self._h, grid_path.encode(), q.ctypes.data, q.size, citer, cfl,
                                      1 if ghosts else 0))

    def footprint(self, precision=FS_PREC_F64):
        """Device bytes this (prepared) mesh costs inside a solver or batch."""
        b = C.c_int64()
        _check(lib().fs_mesh_footprint(self._h, precision, C.byref(b)))
        return b.value

    def close(self):
        if self._h:
            lib().fs_mesh_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Device-resident solver for one mesh on the current HIP device."""

    def __init__(self, mesh, params, precision=FS_PREC_F64):
        self.mesh = mesh
        self.params = params
        self.info = mesh.info()
        h = C.c_void_p()
        _check(lib().fs_solver_create(mesh._h, C.byref(params), precision, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().fs_solver_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _count(self, field):
        return _field_count(self.info, field)

    def get(self, field):
        n = self._count(field)
        out = np.empty(n, dtype=np.float64)
        _check(lib().fs_solver_get(self._h, _FIELDS.index(field), out.ctypes.data, n))
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        _check(lib().fs_solver_set(self._h, _FIELDS.index(field), v.ctypes.data, v.size))

    def init_freestream(self):
        _check(lib().fs_solver_init_freestream(self._h))

    def init_restart(self):
        """State from the mesh's restart payload (ghost states when saved)."""
        _check(lib().fs_solver_init_restart(self._h))

    # reference-function mirrors
    def build_matrix(self, ts):
        _check(lib().fs_build_matrix(self._h, ts))

    def gradient(self):
        _check(lib().fs_gradient(self._h))

    def eigenvalue_contribution(self):
        """cv_calc_eigenvalue_contribution (cv_calc.C:790): ev_con of nQ."""
        _check(lib().fs_eigenvalue_contribution(self._h))

    def set_boundary_values(self):
        """Set_Boundary_Values (cvbc.C:37): the ghost states of nQ."""
        _check(lib().fs_set_boundary_values(self._h))

    def derivatives(self):
        _check(lib().fs_derivatives(self._h))

    def limiter(self, ts):
        _check(lib().fs_limiter(self._h, ts))

    def interior_residual(self, jupdate, ts):
        _check(lib().fs_interior_residual(self._h, jupdate, ts))

    def boundary_residual(self, jupdate):
        _check(lib().fs_boundary_residual(self._h, jupdate))

    def linear_solve(self):
        r = C.c_double()
        _check(lib().fs_linear_solve(self._h, C.byref(r)))
        return r.value

    def solve(self, ts):
        dq, r2 = C.c_double(), C.c_double()
        _check(lib().fs_solve(self._h, ts, C.byref(dq), C.byref(r2)))
        return dq.value, r2.value

    def body_forces(self, citer):
        """Compute_Body_Forces of the resident state after step `citer`: (lift, drag)."""
        f = np.zeros(2)
        _check(lib().fs_body_forces(self._h, citer, f.ctypes.data))
        return f

    def surface_cp(self, b):
        """write_surface_cp's nodal values for boundary b: (x, cp) of its unique nodes."""
        n = C.c_int64()
        _check(lib().fs_surface_cp(self._h, b, None, None, 0, C.byref(n)))
        x, cp = np.zeros(n.value), np.zeros(n.value)
        _check(lib().fs_surface_cp(self._h, b, x.ctypes.data, cp.ctypes.data, n.value, C.byref(n)))
        return x, cp

    def write_surface_cp(self, directory):
        _check(lib().fs_write_surface_cp(self._h, os.fsencode(directory)))

    def solve_steady(self, max_steps, tol):
        st, dr = C.c_int(), C.c_double()
        _check(lib().fs_solve_steady(self._h, max_steps, tol, C.byref(st), C.byref(dr)))
        return st.value, dr.value

    def time_steps(self, ts_first, nsteps):
        t = Timing()
        _check(lib().fs_solver_time_steps(self._h, ts_first, nsteps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def use_graph(self, on=True):
        """Run each pseudo-time step as one captured hipGraph launch."""
        _check(lib().fs_solver_use_graph(self._h, 1 if on else 0))

    def autotune(self):
        """Tune launch shapes on this solver's data (after a step); returns
        the chosen shapes and the seconds spent."""
        t = C.c_double()
        _check(lib().fs_solver_autotune(self._h, C.byref(t)))
        return dict(self.tuning(), seconds=t.value)

    def tuning(self):
        return _tuning(lib().fs_solver_get_tuning, self._h)

    def set_tuning(self, **kw):
        """Set launch shapes by name (the others keep their values)."""
        _set_tuning(lib().fs_solver_set_tuning, self._h, self.tuning(), kw)


def _field_count(info, field):
    return {"Q": 4 * (info["nn"] + info["ng"]), "nQ": 4 * (info["nn"] + info["ng"]), "dQ": 4 * info["nn"],
            "R": 4 * info["nn"], "RHS": 4 * info["nn"], "grad": 8 * info["nn"], "phi": info["nn"],
            "hess": 20 * info["nn"], "pv": 4 * info["nn"], "ev_con": info["nn"], "dt": info["nn"],
            "LHS": 16 * info["nnz"], "LHS32": 16 * info["nnz"]}[field]


class Batch:
    """Many independent steady solves packed into one device image: every
    kernel of a pseudo-time step is one launch over all instances."""

    def __init__(self, meshes, params, precision=FS_PREC_F64):
        if isinstance(params, Params):
            params = [params] * len(meshes)
        if len(params) != len(meshes):
            raise ValueError("one parameter set per mesh")
        self.meshes = list(meshes)          # kept alive: the batch references them
        self.params = list(params)
        self.infos = [m.info() for m in meshes]
        n = len(meshes)
        hs = (C.c_void_p * n)(*[m._h.value for m in meshes])
        ps = (Params * n)(*params)
        h = C.c_void_p()
        _check(lib().fs_batch_create(n, C.cast(hs, C.POINTER(C.c_void_p)), ps, precision, C.byref(h)))
        self._h = h
        self.n = n

    def close(self):
        if self._h:
            lib().fs_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        bi = BatchInfo()
        _check(lib().fs_batch_get_info(self._h, C.byref(bi)))
        return {k: getattr(bi, k) for k, _ in bi._fields_}

    def init_freestream(self):
        _check(lib().fs_batch_init_freestream(self._h))

    def use_graph(self, on=True):
        _check(lib().fs_batch_use_graph(self._h, 1 if on else 0))

    def defer_k2(self, on=True):
        """Deferred k=2 fields (default on; not in the reference): the k=2
        phase's derivatives and limiter, which no residual or Jacobian of
        that phase reads, are evaluated when an output reads them, from the
        state the last step started from (the same bits as every step)."""
        _check(lib().fs_batch_defer_k2(self._h, 1 if on else 0))

    def autotune(self):
        t = C.c_double()
        _check(lib().fs_batch_autotune(self._h, C.byref(t)))
        return dict(self.tuning(), seconds=t.value)

    def tuning(self):
        return _tuning(lib().fs_batch_get_tuning, self._h)

    def set_tuning(self, **kw):
        """Set launch shapes by name (the others keep their values)."""
        _set_tuning(lib().fs_batch_set_tuning, self._h, self.tuning(), kw)

    def set_active(self, i, on):
        _check(lib().fs_batch_set_active(self._h, i, 1 if on else 0))

    def set_compaction(self, on=True):
        """Re-pack running instances into a smaller engine during steady solves."""
        _check(lib().fs_batch_set_compaction(self._h, 1 if on else 0))

    def step(self):
        r2, dq = np.zeros(self.n), np.zeros(self.n)
        _check(lib().fs_batch_step(self._h, r2.ctypes.data, dq.ctypes.data))
        return r2, dq

    def solve_steady(self, max_steps, tol):
        steps = np.zeros(self.n, dtype=np.int32)
        drop = np.zeros(self.n)
        status = np.zeros(self.n, dtype=np.int32)
        _check(lib().fs_batch_solve_steady(self._h, max_steps, tol, steps.ctypes.data, drop.ctypes.data,
                                           status.ctypes.data))
        return steps, drop, status

    def set_recovery(self, max_recoveries, checkpoint_every=10, limiter_fallback=0, first_order_fallback=False):
        """Recovery from non-physical states (not in the reference; off with
        max_recoveries 0): restore the last checkpoint with half the CFL
        target and a restarted ramp; with limiter_fallback > 0 an unlimited
        solve's first failure restarts its phase with the Barth limiter,
        frozen after that many limited steps; with first_order_fallback the
        next failure restarts it at first order."""
        _check(lib().fs_batch_set_recovery(self._h, max_recoveries, checkpoint_every, limiter_fallback,
                                           1 if first_order_fallback else 0))

    def set_stall(self, steps_per_sqrt_node):
        """Stall test of the recovery policy (not in the reference; 0: off):
        a solve whose residual has not halved its lowest mark of the phase
        for max(100, ceil(steps_per_sqrt_node * sqrt(nodes))) steps starts
        the next fallback that applies, at its CFL."""
        _check(lib().fs_batch_set_stall(self._h, steps_per_sqrt_node))

    def set_limiter_freeze(self, steps):
        """Keep the limiter values once the limited phase has run `steps`
        steps (not in the reference; 0: off)."""
        _check(lib().fs_batch_set_limiter_freeze(self._h, steps))

    def recoveries(self, i):
        """(restores, FALLBACK_* bits of the fallbacks that acted) of instance i."""
        n, f = C.c_int(), C.c_int()
        _check(lib().fs_batch_inst_recovery(self._h, i, C.byref(n), C.byref(f)))
        return n.value, f.value

    def status(self, i):
        st, steps, drop = C.c_int(), C.c_int(), C.c_double()
        _check(lib().fs_batch_inst_status(self._h, i, C.byref(st), C.byref(steps), C.byref(drop)))
        return st.value, steps.value, drop.value

    def get(self, i, field):
        n = _field_count(self.infos[i], field)
        out = np.empty(n, dtype=np.float64)
        _check(lib().fs_batch_get(self._h, i, _FIELDS.index(field), out.ctypes.data, n))
        return out

    def set(self, i, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        _check(lib().fs_batch_set(self._h, i, _FIELDS.index(field), v.ctypes.data, v.size))

    def body_forces(self):
        """Compute_Body_Forces of every instance after its last step: (n, 2) lift, drag."""
        f = np.zeros(2 * self.n)
        _check(lib().fs_batch_body_forces(self._h, f.ctypes.data))
        return f.reshape(-1, 2)

    def time_steps(self, nsteps):
        t = Timing()
        _check(lib().fs_batch_time_steps(self._h, nsteps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}


class Server:
    """The serving request loop (C++ worker and device threads inside the
    library): submit steady-solve requests, wait for finished ones."""

    def __init__(self, precision=FS_PREC_MIXED, max_steps=2000, tol=1e-10, device_budget=0, prep_threads=8,
                 use_graph=True, max_recoveries=0, checkpoint_every=10, max_steps_per_sqrt_node=0.0,
                 limiter_freeze=0, limiter_fallback=0, first_order_fallback=False, stall_steps_per_sqrt_node=0.0):
        cfg = ServerConfig()
        lib().fs_server_config_default(C.byref(cfg))
        cfg.precision, cfg.max_steps, cfg.tol = precision, max_steps, tol
        cfg.device_budget, cfg.prep_threads, cfg.use_graph = device_budget, prep_threads, 1 if use_graph else 0
        cfg.max_recoveries, cfg.checkpoint_every = max_recoveries, checkpoint_every
        cfg.max_steps_per_sqrt_node = max_steps_per_sqrt_node
        cfg.limiter_freeze = limiter_freeze
        cfg.limiter_fallback = limiter_fallback
        cfg.first_order_fallback = 1 if first_order_fallback else 0
        cfg.stall_steps_per_sqrt_node = stall_steps_per_sqrt_node
        h = C.c_void_p()
        _check(lib().fs_server_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().fs_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, mesh, params):
        """Queue a solve; the server takes the mesh over (the Mesh object is
        emptied)."""
        rid = C.c_int64()
        _check(lib().fs_server_submit(self._h, mesh._h, C.byref(params), C.byref(rid)))
        mesh._h = C.c_void_p()
        return rid.value

    def wait(self, timeout=-1.0):
        """Next finished request as a dict, or None after `timeout` seconds."""
        info, got = RequestInfo(), C.c_int()
        _check(lib().fs_server_wait(self._h, timeout, C.byref(info), C.byref(got)))
        if not got.value:
            return None
        out = {k: getattr(info, k) for k, _ in info._fields_}
        out["forces"] = (info.forces[0], info.forces[1])
        if info.status == INST_FAILED:
            out["error"] = lib().fs_last_error().decode()
        return out

    def state(self, info):
        n = 4 * (info["nn"] + info["ng"])
        out = np.empty(n)
        _check(lib().fs_server_get_state(self._h, info["id"], out.ctypes.data, n))
        return out

    def release(self, rid):
        _check(lib().fs_server_release(self._h, rid))

    def cancel(self, rid):
        _check(lib().fs_server_cancel(self._h, rid))

    def counts(self):
        q, r, f = C.c_int(), C.c_int(), C.c_int()
        _check(lib().fs_server_counts(self._h, C.byref(q), C.byref(r), C.byref(f)))
        return q.value, r.value, f.value

    def clock(self):
        return lib().fs_server_clock(self._h)
