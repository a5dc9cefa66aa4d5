"""ctypes binding of ``libsimcoproj.so`` (the C ABI in include/simco_proj.h).

The binding loads the in-tree library built by ``__graft_entry__.build()``
and raises if it is missing: there is no CPU fallback.
"""

import ctypes
import os

import numpy as np

from ._paths import ENGINE_LIB

F64P = ctypes.POINTER(ctypes.c_double)
I32P = ctypes.POINTER(ctypes.c_int32)

SP_OK = 0
SP_ERR_INVALID_ARG = -1
SP_ERR_HIP = -2
SP_ERR_NOT_CONVERGED = -3
SP_ERR_NONFINITE = -4
SP_ERR_OOM = -5
SP_ERR_GRID = -6
SP_ERR_STATE = -7
SP_ERR_NO_DEVICE = -8
SP_ERR_CAPACITY = -9

SP_PREC_FP64 = 0
SP_PREC_MIXED = 1
SP_PRECOND_MG = 0
SP_PRECOND_JACOBI = 1
SP_FLAG_UNFUSED_SMOOTHER = 1
SP_FLAG_CHEBYSHEV = 2

EXPORTED_SYMBOLS = (
    "sp_params_default", "sp_create", "sp_destroy", "sp_set_params", "sp_set_grid",
    "sp_set_grid_1d", "sp_project", "sp_project_dev", "sp_adjoint", "sp_adjoint_dev",
    "sp_stream", "sp_kernel_times", "sp_set_profiling", "sp_set_profiling_filter",
    "sp_strerror", "sp_last_error", "sp_wall_step", "sp_wall_step_dev", "sp_write_vtr",
    "sp_write_vtk", "sp_grid_preprocess", "sp_grid_preprocess_dev", "sp_define_momentum_exchange_cell",
    "sp_number_external_cell", "sp_new_cell_face", "sp_compute_uvw_liquid_field",
    "sp_number_external_cell_noalloc", "sp_predict_uvw", "sp_interpolate_lvs",
    "sp_clsvof_scheme",
)

# members of sp_liquid_field, in order
LIQUID_FIELDS = ("dx", "dy", "dz", "phi_l", "nx_l", "ny_l", "nz_l", "vof_l")

# members of sp_cell_interp, in order
INTERP_FIELDS = ("moex", "msce", "sxe", "etae", "ale", "dale", "syn", "etan", "aln", "daln",
                 "szt", "etat", "alt", "dalt")

# members of sp_cell_geom, in order
GEOM_FIELDS = ("dx", "dy", "dz", "phi", "nx", "ny", "nz", "vof", "vof_l", "cell_type",
               "ee_area", "ne_area", "te_area", "fce", "fcn", "fct", "cell_cent", "wllh",
               "delh")
GEOM_INPUTS = GEOM_FIELDS[:7]
GEOM_UPDATED = ("vof", "vof_l", "cell_type", "fce", "fcn", "fct", "delh")
GEOM_WRITTEN = ("ee_area", "ne_area", "te_area", "cell_cent", "wllh")


# members a host may allocate with the halo (sp_cell_geom.halo_fields)
GEOM_HALO_FIELDS = ("phi", "nx", "ny", "nz", "vof", "vof_l")


def geom_shapes(I, J, K, halo_fields=False):
    """Fortran shapes of the sp_cell_geom arrays for an I x J x K grid
    (halo_fields: phi, nx, ny, nz, vof, vof_l with the halo)."""
    cell, halo = (I, J, K), (I + 2, J + 2, K + 2)
    sh = {n: cell for n in GEOM_FIELDS}
    if halo_fields:
        for n in GEOM_HALO_FIELDS:
            sh[n] = halo
    for n in ("ee_area", "ne_area", "te_area"):
        sh[n] = halo
    for n in ("fce", "fcn", "fct"):
        sh[n] = halo + (3,)
    sh["cell_cent"] = cell + (3,)
    return sh


class SpParams(ctypes.Structure):
    _fields_ = [
        ("imax", ctypes.c_int32), ("jmax", ctypes.c_int32), ("kmax", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("precond", ctypes.c_int32),
        ("mg_levels", ctypes.c_int32), ("mg_pre", ctypes.c_int32), ("mg_post", ctypes.c_int32),
        ("mg_coarse_sweeps", ctypes.c_int32), ("mg_coarse_scale", ctypes.c_double),
        ("max_iter", ctypes.c_int32), ("div_tol", ctypes.c_double),
        ("rho_l", ctypes.c_double), ("rho_g", ctypes.c_double), ("rho_ref", ctypes.c_double),
        ("epsi", ctypes.c_double),
        ("pin_i", ctypes.c_int32), ("pin_j", ctypes.c_int32), ("pin_k", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("mg_omega", ctypes.c_double),
        ("mg_omega_coarse", ctypes.c_double),
    ]


class SpWallBc(ctypes.Structure):
    """BCu/BCv/BCw flags and BCu/BCv/BCw/BCp constants of the wall step
    (faces W,E,S,N,B,T; 0 Dirichlet, 1 Neumann)."""
    _fields_ = [("flag_u", ctypes.c_int32 * 6), ("flag_v", ctypes.c_int32 * 6),
                ("flag_w", ctypes.c_int32 * 6), ("const_u", ctypes.c_double * 6),
                ("const_v", ctypes.c_double * 6), ("const_w", ctypes.c_double * 6),
                ("const_p", ctypes.c_double * 6)]


def make_wall(flags, consts):
    """flags: 18 values (BCu W..T, BCv, BCw); consts: 24 (BCu, BCv, BCw, BCp)."""
    w = SpWallBc()
    for f in range(6):
        w.flag_u[f], w.flag_v[f], w.flag_w[f] = (int(flags[6 * q + f]) for q in range(3))
        w.const_u[f], w.const_v[f], w.const_w[f], w.const_p[f] = (
            float(consts[6 * q + f]) for q in range(4))
    return w


VTR_FIELDS = ("u", "v", "w", "p", "phi", "phi_l", "vof", "vof_l", "mres")


class SpVtrFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("x", "y", "z") + VTR_FIELDS] + [
        ("halo_mask", ctypes.c_uint32)]


VTK_FIELDS = ("p", "u", "v", "w", "phi_l", "phi", "vof_l", "vof")


class SpVtkFields(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("x", "y", "z", "dx", "dy", "dz") + VTK_FIELDS] + [
        ("halo_mask", ctypes.c_uint32)]


class SpSolidBody(ctypes.Structure):
    _fields_ = [("imax", ctypes.c_int32), ("jmax", ctypes.c_int32), ("kmax", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("phi", ctypes.c_void_p), ("cm", ctypes.c_double * 3),
                ("rot", ctypes.c_double * 9)]


LVS_GRID_FIELDS = ("x", "y", "z", "dx", "dy", "dz", "phi", "vof", "nx", "ny", "nz")


class SpLvsGrid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in LVS_GRID_FIELDS]


CLSVOF_IN_PTRS = ("u", "v", "w", "x", "y", "z", "dx", "dy", "dz", "nx", "ny", "nz", "bcu_w",
                  "bcv_s", "bcw_b")


class SpClsvofInputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in CLSVOF_IN_PTRS] + [
        ("lvs_flag", ctypes.c_int32 * 6), ("lvs_const", ctypes.c_double * 6),
        ("vof_flag", ctypes.c_int32 * 6), ("vof_const", ctypes.c_double * 6),
        ("dt", ctypes.c_double)]


CLSVOF_STATE_PTRS = ("vof_l", "phi_l", "nx_l", "ny_l", "nz_l")


class SpClsvofState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in CLSVOF_STATE_PTRS] + [
        ("lvs_var", ctypes.c_void_p * 6), ("vof_var", ctypes.c_void_p * 6)]


class SpInputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "u_star", "v_star", "w_star", "dp_u", "dp_v", "dp_w", "ee_area", "ne_area", "te_area",
        "vof", "vof_old", "vof_l", "phi_l", "cell_type_p", "cell_type_u", "cell_type_v",
        "cell_type_w", "moex_u", "moex_v", "moex_w")] + [
        ("bc_flag", ctypes.c_int32 * 6), ("bc_val", ctypes.c_void_p * 6), ("dt", ctypes.c_double),
        ("beta_e", ctypes.c_void_p), ("beta_n", ctypes.c_void_p), ("beta_t", ctypes.c_void_p),
        ("wall", ctypes.c_void_p), ("static_geometry", ctypes.c_int32)]


class SpOutputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("p", "u", "v", "w", "po_coef")] + [
        ("bc_var", ctypes.c_void_p * 24)]


class SpStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("levels", ctypes.c_int32),
        ("singular", ctypes.c_int32), ("active_cells", ctypes.c_int64),
        ("stale_faces", ctypes.c_int64), ("initial_div", ctypes.c_double),
        ("final_div", ctypes.c_double), ("residual_div", ctypes.c_double),
        ("t_assemble_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double),
        ("t_correct_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double),
        ("wall_div", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SpAdjInputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("gu", "gv", "gw", "gp")]


class SpAdjOutputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "gu_star", "gv_star", "gw_star", "gk_e", "gk_n", "gk_t", "gbeta_e", "gbeta_n", "gbeta_t")]


class SpKernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_int32),
                ("total_ms", ctypes.c_double), ("bytes", ctypes.c_double)]


class SpCellGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in GEOM_FIELDS] + [("halo_fields", ctypes.c_int32)]


class SpCellInterp(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in INTERP_FIELDS]


class SpLiquidField(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in LIQUID_FIELDS] + [("halo_fields", ctypes.c_int32)]


class SpPredInputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "p_ee_area", "p_ne_area", "p_te_area", "u_old", "v_old", "w_old", "u_old_x", "v_old_x",
        "w_old_x", "u_cur", "v_cur", "w_cur", "u", "v", "w", "u_n", "v_n", "w_n",
        "flux_div_old")] + [
        ("bc_flag", ctypes.c_int32 * 18), ("bc_var", ctypes.c_void_p * 18),
        ("dt", ctypes.c_double), ("itt", ctypes.c_int64), ("time_order", ctypes.c_int32),
        ("space_order", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in ("mu_w_ref", "mu_a_ref", "rey", "fr", "gx", "gy", "gz", "g")]


class SpPredOutputs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("u", "v", "w", "dp_u", "dp_v", "dp_w")]


# predictor cell members (configs.PRED_CELL_FIELDS) -> sp_cell_geom / sp_cell_interp
_PRED_GEOM = {"ee": "ee_area", "ne": "ne_area", "te": "te_area", "cent": "cell_cent"}
_PRED_INTERP = ("moex", "sxe", "syn", "szt", "etae", "etan", "etat", "ale", "aln", "alt")


class SpError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


_lib = None


def load():
    """Load the engine library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ENGINE_LIB):
        raise RuntimeError(f"HIP engine library missing: {ENGINE_LIB} (run __graft_entry__.build())")
    L = ctypes.CDLL(ENGINE_LIB)
    P = ctypes.POINTER
    V = ctypes.c_void_p
    L.sp_params_default.argtypes = [P(SpParams)]
    L.sp_params_default.restype = None
    L.sp_create.argtypes = [P(V), P(SpParams), ctypes.c_int]
    L.sp_create.restype = ctypes.c_int
    L.sp_destroy.argtypes = [V]
    L.sp_destroy.restype = ctypes.c_int
    L.sp_set_params.argtypes = [V, P(SpParams)]
    L.sp_set_params.restype = ctypes.c_int
    L.sp_set_grid.argtypes = [V] + [F64P] * 6
    L.sp_set_grid.restype = ctypes.c_int
    L.sp_set_grid_1d.argtypes = [V] + [F64P] * 6
    L.sp_set_grid_1d.restype = ctypes.c_int
    for name in ("sp_project", "sp_project_dev"):
        f = getattr(L, name)
        f.argtypes = [V, P(SpInputs), P(SpOutputs), P(SpStats)]
        f.restype = ctypes.c_int
    for name in ("sp_adjoint", "sp_adjoint_dev"):
        f = getattr(L, name)
        f.argtypes = [V, P(SpAdjInputs), P(SpAdjOutputs), P(SpStats)]
        f.restype = ctypes.c_int
    for name in ("sp_wall_step", "sp_wall_step_dev"):
        f = getattr(L, name)
        f.argtypes = [V, P(SpInputs), P(SpOutputs), P(SpStats)]
        f.restype = ctypes.c_int
    L.sp_write_vtr.argtypes = [V, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64,
                               P(SpVtrFields), ctypes.c_int]
    L.sp_write_vtr.restype = ctypes.c_int
    L.sp_write_vtk.argtypes = [V, ctypes.c_char_p, ctypes.c_int64, P(SpVtkFields), ctypes.c_int]
    L.sp_write_vtk.restype = ctypes.c_int
    L.sp_interpolate_lvs.argtypes = [V, P(SpSolidBody), P(SpLvsGrid), ctypes.c_int]
    L.sp_interpolate_lvs.restype = ctypes.c_int
    L.sp_clsvof_scheme.argtypes = [V, P(SpClsvofInputs), P(SpClsvofState), ctypes.c_int]
    L.sp_clsvof_scheme.restype = ctypes.c_int
    for name in ("sp_grid_preprocess", "sp_grid_preprocess_dev"):
        f = getattr(L, name)
        f.argtypes = [V] + [P(SpCellGeom)] * 4
        f.restype = ctypes.c_int
    L.sp_compute_uvw_liquid_field.argtypes = [V] + [P(SpLiquidField)] * 4 + [ctypes.c_int]
    L.sp_compute_uvw_liquid_field.restype = ctypes.c_int
    L.sp_new_cell_face.argtypes = [V] + [P(SpCellGeom)] * 4 + [P(SpCellInterp)] * 3 + [
        ctypes.c_int]
    L.sp_new_cell_face.restype = ctypes.c_int
    L.sp_define_momentum_exchange_cell.argtypes = [V] + [V] * 7 + [ctypes.c_int]
    L.sp_define_momentum_exchange_cell.restype = ctypes.c_int
    L.sp_number_external_cell.argtypes = [V] + [ctypes.c_int] * 3 + [V, V, I32P, ctypes.c_int]
    L.sp_number_external_cell.restype = ctypes.c_int
    L.sp_number_external_cell_noalloc.argtypes = [V] + [ctypes.c_int] * 3 + [V, V, V]
    L.sp_number_external_cell_noalloc.restype = ctypes.c_int
    L.sp_predict_uvw.argtypes = [V, ctypes.POINTER(SpPredInputs), ctypes.POINTER(SpPredOutputs),
                                 ctypes.c_int]
    L.sp_predict_uvw.restype = ctypes.c_int
    L.sp_stream.argtypes = [V]
    L.sp_stream.restype = V
    L.sp_kernel_times.argtypes = [V, P(SpKernelTime), ctypes.c_int]
    L.sp_kernel_times.restype = ctypes.c_int
    L.sp_set_profiling.argtypes = [V, ctypes.c_int]
    L.sp_set_profiling.restype = None
    L.sp_set_profiling_filter.argtypes = [V, ctypes.c_char_p]
    L.sp_set_profiling_filter.restype = None
    L.sp_strerror.argtypes = [ctypes.c_int]
    L.sp_strerror.restype = ctypes.c_char_p
    L.sp_last_error.argtypes = [V]
    L.sp_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def default_params(shape, **kw):
    L = load()
    p = SpParams()
    L.sp_params_default(ctypes.byref(p))
    p.imax, p.jmax, p.kmax = shape
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _addr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


class Projector:
    """One engine context (one grid, one set of solver parameters)."""

    def __init__(self, shape, device=0, **params):
        self.lib = load()
        self.shape = tuple(shape)
        self.params = default_params(shape, **params)
        h = ctypes.c_void_p()
        st = self.lib.sp_create(ctypes.byref(h), ctypes.byref(self.params), device)
        if st != SP_OK:
            raise SpError(st, "sp_create failed: " + self.lib.sp_strerror(st).decode())
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.sp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what, allow=()):
        if st != SP_OK and st not in allow:
            raise SpError(st, f"{what}: {self.lib.sp_strerror(st).decode()}: "
                              f"{self.lib.sp_last_error(self.h).decode()}")
        return st

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._check(self.lib.sp_set_params(self.h, ctypes.byref(self.params)), "sp_set_params")

    def set_grid_1d(self, dxp, dyp, dzp, dxu, dyv, dzw):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (dxp, dyp, dzp, dxu, dyv, dzw)]
        I, J, K = self.shape
        if [a.size for a in arrs] != [I, J, K, I, J, K]:
            raise ValueError("sp_set_grid_1d takes spacings of lengths Imax, Jmax, Kmax")
        self._check(self.lib.sp_set_grid_1d(self.h, *[a.ctypes.data_as(F64P) for a in arrs]),
                    "sp_set_grid_1d")

    def set_grid(self, dxp, dyp, dzp, dxu, dyv, dzw):
        arrs = [np.asfortranarray(a, dtype=np.float64) for a in (dxp, dyp, dzp, dxu, dyv, dzw)]
        n = self.shape[0] * self.shape[1] * self.shape[2]
        if any(a.size != n for a in arrs):
            raise ValueError("sp_set_grid takes (Imax, Jmax, Kmax) arrays; use set_grid_1d for spacings")
        self._check(self.lib.sp_set_grid(self.h, *[a.ctypes.data_as(F64P) for a in arrs]),
                    "sp_set_grid")

    @staticmethod
    def make_inputs(f, bc_flag, dt, bc_val=None, beta=None, wall=None):
        """f: mapping with the sp_inputs field names (numpy arrays or torch tensors);
        wall: an SpWallBc (kept alive by the caller) or None."""
        inp = SpInputs()
        for n in ("u_star", "v_star", "w_star", "dp_u", "dp_v", "dp_w", "ee_area", "ne_area",
                  "te_area", "vof", "vof_old", "vof_l", "phi_l", "cell_type_p", "cell_type_u",
                  "cell_type_v", "cell_type_w", "moex_u", "moex_v", "moex_w"):
            setattr(inp, n, _addr(f.get(n)))
        for q in range(6):
            inp.bc_flag[q] = int(bc_flag[q])
            inp.bc_val[q] = _addr(bc_val[q]) if bc_val is not None and bc_val[q] is not None else None
        inp.dt = dt
        if beta is not None:
            inp.beta_e, inp.beta_n, inp.beta_t = (_addr(b) for b in beta)
        if wall is not None:
            inp.wall = ctypes.addressof(wall)
        return inp

    def _bc_var_arrays(self):
        I, J, K = self.shape
        shapes = [(J, K), (J, K), (I, K), (I, K), (I, J), (I, J)]
        return [np.zeros(shapes[q % 6], order="F") for q in range(24)]

    def project(self, fields, bc_flag, dt, out=None, bc_val=None, beta=None, po_coef=False,
                wall=None, static_geometry=False):
        """Host-array projection (sp_project).  Returns (outputs dict, stats).
        With ``wall`` (SpWallBc) the wall step runs after the correction and
        out["bc_var"] holds the 24 face values.  ``static_geometry``: the
        apertures, vof, vofO, cell types and MoExCell equal the previous
        call's, so their device copies are reused."""
        I, J, K = self.shape
        hs = (I + 2, J + 2, K + 2)
        inp = self.make_inputs(fields, bc_flag, dt, bc_val, beta, wall)
        inp.static_geometry = 1 if static_geometry else 0
        if out is None:
            out = {n: np.zeros(hs, order="F") for n in ("p", "u", "v", "w")}
        if po_coef:
            out["po_coef"] = np.zeros((I, J, K, 6), order="F")
        o = SpOutputs(*(_addr(out.get(n)) for n in ("p", "u", "v", "w", "po_coef")))
        if wall is not None:
            out["bc_var"] = self._bc_var_arrays()
            for q in range(24):
                o.bc_var[q] = _addr(out["bc_var"][q])
        st = SpStats()
        rc = self.lib.sp_project(self.h, ctypes.byref(inp), ctypes.byref(o), ctypes.byref(st))
        self._check(rc, "sp_project", allow=(SP_ERR_NOT_CONVERGED,))
        return out, st

    def project_dev(self, fields, bc_flag, dt, out, bc_val=None, beta=None, wall=None):
        """Device-array projection (sp_project_dev) on torch tensors; with
        ``wall`` (SpWallBc) the wall step follows, and out["bc_var"] (a list
        of 24 device tensors or None entries), when present, receives the
        face values."""
        inp = self.make_inputs(fields, bc_flag, dt, bc_val, beta, wall)
        o = SpOutputs(*(_addr(out.get(n)) for n in ("p", "u", "v", "w", "po_coef")))
        for q, a in enumerate(out.get("bc_var") or ()):
            o.bc_var[q] = _addr(a)
        st = SpStats()
        rc = self.lib.sp_project_dev(self.h, ctypes.byref(inp), ctypes.byref(o), ctypes.byref(st))
        self._check(rc, "sp_project_dev", allow=(SP_ERR_NOT_CONVERGED,))
        return st

    def wall_step(self, fields, bc_flag, wall, u, v, w, p):
        """The wall step alone (sp_wall_step) on copies of the host arrays
        u, v, w, p.  Returns (dict u, v, w, p, bc_var; stats)."""
        inp = self.make_inputs(fields, bc_flag, 1.0, wall=wall)
        out = {n: np.array(a, dtype=np.float64, order="F", copy=True)
               for n, a in (("u", u), ("v", v), ("w", w), ("p", p))}
        out["bc_var"] = self._bc_var_arrays()
        o = SpOutputs(*(_addr(out.get(n)) for n in ("p", "u", "v", "w", "po_coef")))
        for q in range(24):
            o.bc_var[q] = _addr(out["bc_var"][q])
        st = SpStats()
        rc = self.lib.sp_wall_step(self.h, ctypes.byref(inp), ctypes.byref(o), ctypes.byref(st))
        self._check(rc, "sp_wall_step")
        return out, st

    def grid_preprocess(self, cells, device=False, halo_fields=False):
        """GridPreprocess (sp_grid_preprocess) on the P, U, V, W grids.
        ``cells``: four dicts keyed like sp_cell_geom.  Host (numpy) inputs
        are copied; the result is four dicts with the updated and written
        arrays.  With device=True every member is a torch tensor on the
        context's device (Fortran shape reversed, C-contiguous), updated in
        place; nothing is returned.  ``halo_fields``: phi, nx, ny, nz, vof
        and vof_l are halo arrays, as the reference's drivers allocate them."""
        I, J, K = self.shape
        sh = geom_shapes(I, J, K, halo_fields)
        structs = []
        if device:
            for c in cells:
                for n in GEOM_FIELDS:
                    t = c[n]
                    if tuple(t.shape) != sh[n][::-1] or not t.is_contiguous():
                        raise ValueError(f"{n}: expected a contiguous tensor of shape "
                                         f"{sh[n][::-1]}, got {tuple(t.shape)}")
                structs.append(SpCellGeom(*(_addr(c[n]) for n in GEOM_FIELDS),
                                          1 if halo_fields else 0))
            rc = self.lib.sp_grid_preprocess_dev(self.h, *(ctypes.byref(s) for s in structs))
            self._check(rc, "sp_grid_preprocess_dev")
            return None
        outs = []
        for c in cells:
            o = {}
            for n in GEOM_FIELDS:
                dt = np.int32 if n == "cell_type" else np.float64
                if n in GEOM_WRITTEN and n not in c:
                    o[n] = np.zeros(sh[n], dtype=dt, order="F")
                else:
                    o[n] = np.array(c[n], dtype=dt, order="F", copy=True)
                if o[n].shape != sh[n]:
                    raise ValueError(f"{n}: expected shape {sh[n]}, got {o[n].shape}")
            outs.append(o)
            structs.append(SpCellGeom(*(_addr(o[n]) for n in GEOM_FIELDS), 1 if halo_fields else 0))
        rc = self.lib.sp_grid_preprocess(self.h, *(ctypes.byref(s) for s in structs))
        self._check(rc, "sp_grid_preprocess")
        return outs

    def uvw_liquid_field(self, cells, device=False, halo_fields=False):
        """ComputeUVWLiquidField (sp_compute_uvw_liquid_field).  cells: four
        dicts (P, U, V, W) keyed like sp_liquid_field (cell arrays; with
        ``halo_fields`` phi_l, nx_l, ny_l, nz_l, vof_l are halo arrays).
        Host: returns the U, V, W dicts with new phi_l, nx_l, ny_l, nz_l,
        vof_l; device (torch tensors): writes them in place."""
        I, J, K = self.shape
        hf = 1 if halo_fields else 0
        if device:
            gs = [SpLiquidField(*(_addr(c.get(n)) for n in LIQUID_FIELDS), hf) for c in cells]
            rc = self.lib.sp_compute_uvw_liquid_field(self.h, *(ctypes.byref(a) for a in gs), 1)
            self._check(rc, "sp_compute_uvw_liquid_field")
            return None
        outs, gs = [], []
        for m, c in enumerate(cells):
            o = {}
            for n in LIQUID_FIELDS:
                a = c.get(n)
                if a is None and (m == 0 or n in ("dx", "dy", "dz")):
                    raise ValueError(f"{n} is required")
                shp = (I + 2, J + 2, K + 2) if halo_fields and n not in ("dx", "dy", "dz") \
                    else (I, J, K)
                o[n] = (np.zeros(shp, order="F") if a is None else
                        np.array(a, dtype=np.float64, order="F", copy=True))
                if o[n].shape != shp:
                    raise ValueError(f"{n}: expected shape {shp}, got {o[n].shape}")
            outs.append(o)
            gs.append(SpLiquidField(*(_addr(o[n]) for n in LIQUID_FIELDS), hf))
        rc = self.lib.sp_compute_uvw_liquid_field(self.h, *(ctypes.byref(a) for a in gs), 0)
        self._check(rc, "sp_compute_uvw_liquid_field")
        return outs[1:]

    def new_cell_face(self, cells, interp, device=False, halo_fields=False):
        """NewCellFace (sp_new_cell_face).  cells: four dicts (P, U, V, W)
        keyed like sp_cell_geom; interp: three dicts (U, V, W) keyed like
        sp_cell_interp (moex required; msce and the interpolation arrays
        start values, zeros when missing).  Host: returns updated copies
        (cells, interp); device (torch tensors): updates in place.
        ``halo_fields`` as for grid_preprocess."""
        I, J, K = self.shape
        sh = geom_shapes(I, J, K, halo_fields)
        ish = {n: (I + 2, J + 2, K + 2) for n in INTERP_FIELDS}
        ish["moex"] = (I, J, K)
        ish["msce"] = (I, J, K, 3)
        if device:
            gs = [SpCellGeom(*(_addr(c.get(n)) for n in GEOM_FIELDS), 1 if halo_fields else 0)
                  for c in cells]
            xs = [SpCellInterp(*(_addr(d.get(n)) for n in INTERP_FIELDS)) for d in interp]
            rc = self.lib.sp_new_cell_face(self.h, *(ctypes.byref(a) for a in gs),
                                           *(ctypes.byref(a) for a in xs), 1)
            self._check(rc, "sp_new_cell_face")
            return None
        cs, gs = [], []
        for c in cells:
            o = {}
            for n in GEOM_FIELDS:
                dt = np.int32 if n == "cell_type" else np.float64
                a = c.get(n)
                o[n] = (np.zeros(sh[n], dtype=dt, order="F") if a is None else
                        np.array(a, dtype=dt, order="F", copy=True))
                if o[n].shape != sh[n]:
                    raise ValueError(f"{n}: expected shape {sh[n]}, got {o[n].shape}")
            cs.append(o)
            gs.append(SpCellGeom(*(_addr(o[n]) for n in GEOM_FIELDS), 1 if halo_fields else 0))
        xs, xstructs = [], []
        for d in interp:
            o = {}
            for n in INTERP_FIELDS:
                dt = np.int32 if n in ("moex", "msce") else np.float64
                a = d.get(n)
                o[n] = (np.zeros(ish[n], dtype=dt, order="F") if a is None else
                        np.array(a, dtype=dt, order="F", copy=True))
                if o[n].shape != ish[n]:
                    raise ValueError(f"{n}: expected shape {ish[n]}, got {o[n].shape}")
            xs.append(o)
            xstructs.append(SpCellInterp(*(_addr(o[n]) for n in INTERP_FIELDS)))
        rc = self.lib.sp_new_cell_face(self.h, *(ctypes.byref(a) for a in gs),
                                       *(ctypes.byref(a) for a in xstructs), 0)
        self._check(rc, "sp_new_cell_face")
        return cs, xs

    def define_moex(self, ctp, ctu, ctv, ctw, mu, mv, mw, device=False):
        """DefineMomentumExchangeCell (sp_define_momentum_exchange_cell).  Host:
        returns updated copies of mu, mv, mw; device (torch int32 tensors):
        updates mu, mv, mw in place."""
        if device:
            args = [_addr(a) for a in (ctp, ctu, ctv, ctw, mu, mv, mw)]
            rc = self.lib.sp_define_momentum_exchange_cell(self.h, *args, 1)
            self._check(rc, "sp_define_momentum_exchange_cell")
            return None
        shape = self.shape
        cts = [np.asfortranarray(a, dtype=np.int32) for a in (ctp, ctu, ctv, ctw)]
        ms = [np.array(a, dtype=np.int32, order="F", copy=True) for a in (mu, mv, mw)]
        for a in cts + ms:
            if a.shape != shape:
                raise ValueError(f"expected cell arrays of shape {shape}, got {a.shape}")
        rc = self.lib.sp_define_momentum_exchange_cell(self.h, *(_addr(a) for a in cts + ms), 0)
        self._check(rc, "sp_define_momentum_exchange_cell")
        return ms

    def number_external(self, cell_type, iu, iv, iw, posnu=None, device=False):
        """NumberExternalCell (sp_number_external_cell): (PosNu, ExtCell).
        Device: cell_type and posnu are torch int32 tensors (posnu written)."""
        I, J, K = self.shape
        ext = ctypes.c_int32()
        if device:
            rc = self.lib.sp_number_external_cell(self.h, iu, iv, iw, _addr(cell_type),
                                                  _addr(posnu), ctypes.byref(ext), 1)
            self._check(rc, "sp_number_external_cell")
            return posnu, ext.value
        ct = np.asfortranarray(cell_type, dtype=np.int32)
        if ct.shape != (I, J, K):
            raise ValueError(f"expected Cell_Type of shape {(I, J, K)}, got {ct.shape}")
        pos = np.zeros((I - iu, J - iv, K - iw), dtype=np.int32, order="F")
        rc = self.lib.sp_number_external_cell(self.h, iu, iv, iw, _addr(ct), _addr(pos),
                                              ctypes.byref(ext), 0)
        self._check(rc, "sp_number_external_cell")
        return pos, ext.value

    def predict_uvw(self, p_areas, old, cur, fields, pred, dp, device=False, halo_fields=False):
        """PredictingUVW (sp_predict_uvw) with SpaceOrder 1 or 2 (from
        fields["params"]["space_order"]).  ``p_areas``:
        PCellO%EEArea, NEArea, TEArea; ``old``/``cur``: three dicts (U, V, W)
        keyed by configs.PRED_CELL_FIELDS (UCellO.. / UCell..); ``fields``: the
        dict of configs.predictor_fields (vel, vel_n, flux_div_old, bc_flag,
        bcv, params).  ``pred`` (Pred%u, v, w) and ``dp`` (PU/PV/PW%Dp) are
        written in place, and fields["flux_div_old"] is updated in place:
        numpy arrays in the host's layouts, or torch tensors with
        device=True.  ``halo_fields``: the cells' phi, nx, ny, nz, vof and
        vof_l are halo arrays."""
        keep = []

        def ptr(a, dt=None):
            if a is None:
                return None
            if not device:
                a = np.asfortranarray(a, dtype=dt or a.dtype)
                keep.append(a)
            return _addr(a)

        def geom(d):
            g = SpCellGeom(*(None for _ in GEOM_FIELDS), 1 if halo_fields else 0)
            for n in d:
                if n in _PRED_INTERP:
                    continue
                setattr(g, _PRED_GEOM.get(n, n),
                        ptr(d[n], np.int32 if n == "cell_type" else np.float64))
            keep.append(g)
            return g

        def interp(d):
            x = SpCellInterp(*(None for _ in INTERP_FIELDS))
            for n in _PRED_INTERP:
                if n in d:
                    setattr(x, n, ptr(d[n], np.int32 if n == "moex" else np.float64))
            keep.append(x)
            return x

        inp = SpPredInputs()
        inp.p_ee_area, inp.p_ne_area, inp.p_te_area = (ptr(a, np.float64) for a in p_areas)
        og = [geom(d) for d in old]
        ox = [interp(d) for d in old]
        cg = [geom(d) for d in cur]
        inp.u_old, inp.v_old, inp.w_old = (ctypes.addressof(g) for g in og)
        inp.u_old_x, inp.v_old_x, inp.w_old_x = (ctypes.addressof(x) for x in ox)
        inp.u_cur, inp.v_cur, inp.w_cur = (ctypes.addressof(g) for g in cg)
        inp.u, inp.v, inp.w = (ptr(a, np.float64) for a in fields["vel"])
        if fields.get("vel_n") is not None:
            inp.u_n, inp.v_n, inp.w_n = (ptr(a, np.float64) for a in fields["vel_n"])
        fdo = fields["flux_div_old"]
        if not device and not (isinstance(fdo, np.ndarray) and fdo.flags.f_contiguous and
                               fdo.dtype == np.float64):
            raise ValueError("flux_div_old must be a Fortran-ordered float64 array (updated in place)")
        inp.flux_div_old = _addr(fdo)
        for q in range(18):
            inp.bc_flag[q] = int(fields["bc_flag"][q])
            inp.bc_var[q] = ptr(fields["bcv"][q], np.float64)
        p = fields["params"]
        inp.dt, inp.itt = p["dt"], p["itt"]
        inp.time_order, inp.space_order = p["time_order"], p["space_order"]
        for n in ("mu_w_ref", "mu_a_ref", "rey", "fr", "gx", "gy", "gz", "g"):
            setattr(inp, n, p[n])
        for a in list(pred) + list(dp):
            if not device and not (isinstance(a, np.ndarray) and a.flags.f_contiguous and
                                   a.dtype == np.float64):
                raise ValueError("pred and dp must be Fortran-ordered float64 arrays")
        out = SpPredOutputs(*(_addr(a) for a in list(pred) + list(dp)))
        rc = self.lib.sp_predict_uvw(self.h, ctypes.byref(inp), ctypes.byref(out),
                                     1 if device else 0)
        self._check(rc, "sp_predict_uvw")

    def number_external_noalloc(self, cell_type, iu, iv, iw, posnu, ext_cell):
        """NumberExternalCell without device allocation
        (sp_number_external_cell_noalloc): cell_type, posnu and ext_cell
        (one int32) are torch tensors on the context's device; the work is
        enqueued on the context's stream; this wrapper then synchronizes the
        device (every stream on it, the context's included)."""
        import torch
        rc = self.lib.sp_number_external_cell_noalloc(self.h, iu, iv, iw, _addr(cell_type),
                                                      _addr(posnu), _addr(ext_cell))
        self._check(rc, "sp_number_external_cell_noalloc")
        torch.cuda.synchronize()

    def write_vtr(self, directory, prefix, itt, fields, device=False):
        """PrintResultVTR3D's file (sp_write_vtr): fields maps x, y, z (1-D) and
        u, v, w, p, phi, phi_l, vof, vof_l, mres (halo or cell arrays; numpy
        arrays, or torch device tensors with device=True)."""
        I, J, K = self.shape
        f = SpVtrFields()
        keep = []
        mask = 0
        for n in ("x", "y", "z") + VTR_FIELDS:
            a = fields[n]
            if isinstance(a, np.ndarray):
                a = np.asfortranarray(a, dtype=np.float64)
                keep.append(a)
            setattr(f, n, _addr(a))
            if n in VTR_FIELDS:
                shape = tuple(a.shape) if isinstance(a, np.ndarray) else tuple(a.shape)[::-1]
                if shape == (I + 2, J + 2, K + 2):
                    mask |= 1 << VTR_FIELDS.index(n)
                elif shape != (I, J, K):
                    raise ValueError(f"{n}: shape {shape} is neither a halo nor a cell array")
        f.halo_mask = mask
        d = directory if directory.endswith("/") else directory + "/"
        rc = self.lib.sp_write_vtr(self.h, d.encode(), prefix.encode(), int(itt), ctypes.byref(f),
                                   1 if device else 0)
        self._check(rc, "sp_write_vtr")
        return os.path.join(d, "Paraview", f"{prefix}_{int(itt):08d}.vtr")

    def write_vtk(self, directory, itt, fields, device=False):
        """PrintResultVTK's file (sp_write_vtk): fields maps x, y, z (cell
        centres) and dx, dy, dz (cell widths), 1-D, and p, u, v, w, phi_l,
        phi, vof_l, vof (halo or cell arrays); numpy arrays, or torch device
        tensors with device=True.  directory "" writes into the working
        folder.  Returns the path of the file."""
        I, J, K = self.shape
        f = SpVtkFields()
        keep = []
        mask = 0
        for n in ("x", "y", "z", "dx", "dy", "dz") + VTK_FIELDS:
            a = fields[n]
            if isinstance(a, np.ndarray):
                a = np.asfortranarray(a, dtype=np.float64)
                keep.append(a)
            setattr(f, n, _addr(a))
            if n in VTK_FIELDS:
                shape = tuple(a.shape) if isinstance(a, np.ndarray) else tuple(a.shape)[::-1]
                if shape == (I + 2, J + 2, K + 2):
                    mask |= 1 << VTK_FIELDS.index(n)
                elif shape != (I, J, K):
                    raise ValueError(f"{n}: shape {shape} is neither a halo nor a cell array")
        f.halo_mask = mask
        d = directory if (not directory or directory.endswith("/")) else directory + "/"
        rc = self.lib.sp_write_vtk(self.h, d.encode(), int(itt), ctypes.byref(f),
                                   1 if device else 0)
        self._check(rc, "sp_write_vtk")
        itt = int(itt)
        curd = f"{itt:08d}" if 0 <= itt <= 99999999 else "*" * 8
        return d + f"PCell_Binary_3D{curd}.vtr"

    def interpolate_lvs(self, body, grid, device=False):
        """interpolateLvS of a solid body onto one grid (sp_interpolate_lvs).
        body maps x, y, z (local centres), dx, dy, dz, phi (imax, jmax, kmax),
        cm (3) and rot (3x3 rotMatLtoG); grid maps x, y, z, dx, dy, dz (cell
        arrays) and phi, vof, nx, ny, nz (halo arrays, updated in place).
        numpy arrays, or torch device tensors with device=True (body arrays
        included)."""
        keep = []

        def addr(a, out=False):
            if isinstance(a, np.ndarray):
                if out:
                    if not (a.flags.f_contiguous and a.dtype == np.float64):
                        raise ValueError("output arrays must be Fortran-ordered float64")
                else:
                    a = np.asfortranarray(a, dtype=np.float64)
                    keep.append(a)
            return _addr(a)
        phi = body["phi"]
        bshape = tuple(phi.shape) if isinstance(phi, np.ndarray) else tuple(phi.shape)[::-1]
        rot = np.asarray(body["rot"], dtype=np.float64).reshape(3, 3).ravel(order="F")
        b = SpSolidBody(*bshape, float(body["dx"]), float(body["dy"]), float(body["dz"]),
                        addr(body["x"]), addr(body["y"]), addr(body["z"]), addr(phi),
                        (ctypes.c_double * 3)(*np.asarray(body["cm"], dtype=np.float64)),
                        (ctypes.c_double * 9)(*rot))
        g = SpLvsGrid(*(addr(grid[n], out=n in ("phi", "vof", "nx", "ny", "nz"))
                        for n in LVS_GRID_FIELDS))
        rc = self.lib.sp_interpolate_lvs(self.h, ctypes.byref(b), ctypes.byref(g),
                                         1 if device else 0)
        self._check(rc, "sp_interpolate_lvs")

    def clsvof_scheme(self, inputs, state, device=False):
        """Clsvof_Scheme (sp_clsvof_scheme).  inputs maps u, v, w, nx, ny, nz
        (halo), x, y, z, dx, dy, dz (cell), bcu_w, bcv_s, bcw_b (face
        arrays), lvs_flag, lvs_const, vof_flag, vof_const (6 each) and dt;
        state maps vof_l, phi_l, nx_l, ny_l, nz_l (halo) and lvs_var,
        vof_var (six face arrays each, vof_var entries may be None), all
        updated in place.  numpy arrays (Fortran order), or torch device
        tensors with device=True."""
        keep = []

        def inp(a):
            if isinstance(a, np.ndarray):
                a = np.asfortranarray(a, dtype=np.float64)
                keep.append(a)
            return _addr(a)

        def out(a):
            if isinstance(a, np.ndarray) and not (a.flags.f_contiguous and a.dtype == np.float64):
                raise ValueError("state arrays must be Fortran-ordered float64")
            return _addr(a)
        ci = SpClsvofInputs(*(inp(inputs[n]) for n in CLSVOF_IN_PTRS),
                            (ctypes.c_int32 * 6)(*inputs["lvs_flag"]),
                            (ctypes.c_double * 6)(*inputs["lvs_const"]),
                            (ctypes.c_int32 * 6)(*inputs["vof_flag"]),
                            (ctypes.c_double * 6)(*inputs["vof_const"]), float(inputs["dt"]))
        vv = list(state.get("vof_var") or [None] * 6)
        cs = SpClsvofState(*(out(state[n]) for n in CLSVOF_STATE_PTRS),
                           (ctypes.c_void_p * 6)(*(out(a) for a in state["lvs_var"])),
                           (ctypes.c_void_p * 6)(*(out(a) for a in vv)))
        rc = self.lib.sp_clsvof_scheme(self.h, ctypes.byref(ci), ctypes.byref(cs),
                                       1 if device else 0)
        self._check(rc, "sp_clsvof_scheme")

    def adjoint(self, gu, gv, gw, beta_grad=False, gp=None):
        """dL/du*, dL/dK (and dL/dbeta) for dL/du, dL/dv, dL/dw and, when gp
        is given, dL/dPp (halo arrays)."""
        I, J, K = self.shape
        hs = (I + 2, J + 2, K + 2)
        keep = [np.asfortranarray(g) for g in (gu, gv, gw)]
        keep.append(None if gp is None else np.asfortranarray(gp))
        ai = SpAdjInputs(*(_addr(g) for g in keep))
        res = {n: np.zeros(hs, order="F") for n in ("gu_star", "gv_star", "gw_star")}
        for n in ("gk_e", "gk_n", "gk_t") + (("gbeta_e", "gbeta_n", "gbeta_t") if beta_grad else ()):
            res[n] = np.zeros((I, J, K), order="F")
        ao = SpAdjOutputs(*(_addr(res.get(n)) for n in (
            "gu_star", "gv_star", "gw_star", "gk_e", "gk_n", "gk_t", "gbeta_e", "gbeta_n", "gbeta_t")))
        st = SpStats()
        rc = self.lib.sp_adjoint(self.h, ctypes.byref(ai), ctypes.byref(ao), ctypes.byref(st))
        self._check(rc, "sp_adjoint", allow=(SP_ERR_NOT_CONVERGED,))
        return res, st

    def adjoint_dev(self, g, res):
        """g: dict gu, gv, gw and optionally gp (device tensors); res: dict of
        device output tensors."""
        ai = SpAdjInputs(*(_addr(g.get(n)) for n in ("gu", "gv", "gw", "gp")))
        ao = SpAdjOutputs(*(_addr(res.get(n)) for n in (
            "gu_star", "gv_star", "gw_star", "gk_e", "gk_n", "gk_t", "gbeta_e", "gbeta_n", "gbeta_t")))
        st = SpStats()
        rc = self.lib.sp_adjoint_dev(self.h, ctypes.byref(ai), ctypes.byref(ao), ctypes.byref(st))
        self._check(rc, "sp_adjoint_dev", allow=(SP_ERR_NOT_CONVERGED,))
        return st

    def set_profiling(self, on):
        self.lib.sp_set_profiling(self.h, 1 if on else 0)

    def set_profiling_filter(self, name=None):
        """Time only the launches of kernel `name` (None: every kernel)."""
        self.lib.sp_set_profiling_filter(self.h, name.encode() if name else None)

    def kernel_times(self):
        n = self.lib.sp_kernel_times(self.h, None, 0)
        arr = (SpKernelTime * max(n, 1))()
        self.lib.sp_kernel_times(self.h, arr, n)
        return [dict(name=a.name.decode(), launches=a.launches, total_ms=a.total_ms, bytes=a.bytes)
                for a in arr[:n]]

    def stream(self):
        return self.lib.sp_stream(self.h)
