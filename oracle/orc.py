"""ctypes binding of ``oracle/_build/liborc.so`` (test infrastructure only).

Every function restates a piece of the reference (file:line in orc.h).  The
high-level ``project`` runs the whole path the way ``PoissonEquationSolver``
plus the correction block do: GFM coefficients -> reference rows/RHS -> a
tightly converged solve -> pressure field with ghosts -> corrected velocity.
"""

import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

F64P = ctypes.POINTER(ctypes.c_double)
I32P = ctypes.POINTER(ctypes.c_int32)
I64P = ctypes.POINTER(ctypes.c_int64)


class OrcGrid(ctypes.Structure):
    _fields_ = [("I", ctypes.c_int32), ("J", ctypes.c_int32), ("K", ctypes.c_int32)] + [
        (n, F64P) for n in ("dxp", "dyp", "dzp", "dxu", "dyv", "dzw")]


class OrcOpts(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("pre", ctypes.c_int32), ("post", ctypes.c_int32),
                ("coarse_sweeps", ctypes.c_int32), ("max_iter", ctypes.c_int32),
                ("coarse_scale", ctypes.c_double), ("div_tol", ctypes.c_double),
                ("use_rows", ctypes.c_int32), ("precond", ctypes.c_int32),
                ("omega", ctypes.c_double), ("omega_coarse", ctypes.c_double)]


class OrcStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("levels", ctypes.c_int32), ("initial_div", ctypes.c_double),
                ("final_div", ctypes.c_double), ("seconds", ctypes.c_double)]


GEOM_FIELDS = ("dx", "dy", "dz", "phi", "nx", "ny", "nz", "vof", "vof_l", "cell_type",
               "ee_area", "ne_area", "te_area", "fce", "fcn", "fct", "cell_cent", "wllh",
               "delh")
GEOM_INPUTS = GEOM_FIELDS[:7]


class OrcCell(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in GEOM_FIELDS]


INTERP_FIELDS = ("sxe", "syn", "szt", "etae", "etan", "etat", "ale", "aln", "alt", "dale",
                 "daln", "dalt")


class OrcInterp(ctypes.Structure):
    _fields_ = [("moex", ctypes.c_void_p), ("msce", ctypes.c_void_p)] + [
        (n, ctypes.c_void_p) for n in INTERP_FIELDS]


LIQUID_FIELDS = ("dx", "dy", "dz", "phi_l", "nx_l", "ny_l", "nz_l", "vof_l")


class OrcLiquid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in LIQUID_FIELDS]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        G = ctypes.POINTER(OrcGrid)
        L.orc_gfm_coef.argtypes = [G] + [F64P] * 6 + [ctypes.c_double] * 4 + [F64P]
        L.orc_gfm_coef.restype = None
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_gfm_coef_fast.argtypes = L.orc_gfm_coef.argtypes
        L.orc_gfm_coef_fast.restype = None
        L.orc_matrix_rows.argtypes = [G, F64P, F64P, F64P, F64P, I32P, I32P, F64P, I64P]
        L.orc_matrix_rows.restype = None
        L.orc_rhs.argtypes = [G, F64P, F64P, F64P, F64P, I32P, I32P, ctypes.POINTER(F64P),
                              F64P, F64P, F64P, F64P, F64P, ctypes.c_double, F64P]
        L.orc_rhs.restype = None
        L.orc_rows_apply.argtypes = [ctypes.c_int64, F64P, I64P, F64P, F64P]
        L.orc_rows_apply.restype = None
        L.orc_sym_operator.argtypes = [G, F64P, F64P, F64P, F64P, I32P, I32P, F64P, F64P, F64P, F64P]
        L.orc_sym_operator.restype = None
        L.orc_sym_diag.argtypes = [ctypes.c_int] * 3 + [F64P] * 5
        L.orc_sym_diag.restype = ctypes.c_int64
        L.orc_sym_apply.argtypes = [ctypes.c_int] * 3 + [F64P] * 6
        L.orc_sym_apply.restype = None
        L.orc_pcg.argtypes = [G, F64P, F64P, F64P, F64P, F64P, I64P, F64P,
                              ctypes.POINTER(OrcOpts), F64P, ctypes.POINTER(OrcStats)]
        L.orc_pcg.restype = ctypes.c_int
        L.orc_gauge.argtypes = [ctypes.c_int] * 3 + [F64P, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int, F64P]
        L.orc_gauge.restype = None
        L.orc_pressure_field.argtypes = [G, I32P, I32P, F64P, F64P]
        L.orc_pressure_field.restype = None
        L.orc_correct.argtypes = [G, F64P, F64P] + [I32P] * 6 + [F64P] * 6
        L.orc_correct.restype = None
        L.orc_mres.argtypes = [G, F64P, F64P, F64P, F64P, F64P, F64P, F64P]
        L.orc_mres.restype = ctypes.c_double
        L.orc_adjoint.argtypes = ([G, F64P, F64P, F64P, F64P, I32P, I32P] + [I32P] * 6 +
                                  [F64P] * 5 + [ctypes.c_int] * 3 +
                                  [ctypes.POINTER(OrcOpts)] + [F64P] * 7)
        L.orc_adjoint.restype = ctypes.c_int
        L.orc_wall_step.argtypes = [G, F64P, F64P, F64P] + [I32P] * 5 + [F64P] * 5 + [
            ctypes.POINTER(F64P)]
        L.orc_wall_step.restype = None
        L.orc_face_geo.argtypes = [ctypes.c_double] * 7 + [F64P, F64P]
        L.orc_face_geo.restype = None
        C = ctypes.POINTER(OrcCell)
        L.orc_grid_preprocess.argtypes = [ctypes.c_int] * 3 + [C] * 4
        L.orc_grid_preprocess.restype = None
        L.orc_new_cell_face.argtypes = [ctypes.c_int] * 3 + [C] * 4 + [ctypes.POINTER(OrcInterp)] * 3
        L.orc_new_cell_face.restype = None
        L.orc_volume_fraction.argtypes = [ctypes.c_double] * 7
        L.orc_volume_fraction.restype = ctypes.c_double
        Lq = ctypes.POINTER(OrcLiquid)
        L.orc_uvw_liquid_field.argtypes = [ctypes.c_int] * 3 + [Lq] * 4
        L.orc_uvw_liquid_field.restype = None
        L.orc_define_moex.argtypes = [ctypes.c_int] * 3 + [I32P] * 7
        L.orc_define_moex.restype = None
        L.orc_number_external.argtypes = [ctypes.c_int] * 6 + [I32P, I32P]
        L.orc_number_external.restype = ctypes.c_int32
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        assert a.flags.f_contiguous or a.ndim == 1
        return a.ctypes.data_as(F64P)
    if a.dtype == np.int32:
        return a.ctypes.data_as(I32P)
    if a.dtype == np.int64:
        return a.ctypes.data_as(I64P)
    raise TypeError(a.dtype)


def grid_of(case):
    g = OrcGrid(case.I, case.J, case.K, _p(case.dxp), _p(case.dyp), _p(case.dzp),
                _p(case.dxu), _p(case.dyv), _p(case.dzw))
    g._keep = case
    return g


def opts(levels=0, pre=2, post=2, coarse_sweeps=4, max_iter=200, coarse_scale=1.0,
         div_tol=1e-10, use_rows=0, precond=2, omega=1.0, omega_coarse=1.0):
    return OrcOpts(levels, pre, post, coarse_sweeps, max_iter, coarse_scale, div_tol, use_rows,
                   precond, omega, omega_coarse)


def gfm_coef(case, rho_l=None, rho_g=None, rho_ref=None, epsi=1.0e-3, phi_l=None, vof_l=None,
             vof=None):
    rho_l = case.rho_l if rho_l is None else rho_l
    rho_g = case.rho_g if rho_g is None else rho_g
    rho_ref = rho_l if rho_ref is None else rho_ref
    pc = np.zeros((case.I, case.J, case.K, 6), order="F")
    g = grid_of(case)
    lib().orc_gfm_coef(ctypes.byref(g), _p(case.phi_l if phi_l is None else phi_l),
                       _p(case.vof_l if vof_l is None else vof_l),
                       _p(case.vof if vof is None else vof), _p(case.dp_u), _p(case.dp_v),
                       _p(case.dp_w), rho_l, rho_g, rho_ref, epsi, _p(pc))
    return pc


def gfm_coef_fast(case, epsi=1.0e-3):
    pc = np.zeros((case.I, case.J, case.K, 6), order="F")
    g = grid_of(case)
    lib().orc_gfm_coef_fast(ctypes.byref(g), _p(case.phi_l), _p(case.vof_l), _p(case.vof),
                            _p(case.dp_u), _p(case.dp_v), _p(case.dp_w), case.rho_l, case.rho_g,
                            case.rho_l, epsi, _p(pc))
    return pc


def num_threads():
    return int(lib().orc_num_threads())


def baseline_project(case, levels=0, pre=2, post=2, coarse_sweeps=2, coarse_scale=0.5,
                     div_tol=1e-10, max_iter=200, omega=1.25, omega_coarse=1.5, pin=(1, 1, 1)):
    """The CPU baseline: the same path and multigrid-preconditioned CG as the
    engine, FP64, OpenMP over the host cores.  Returns (seconds, info); info
    holds the pressure field Pp (gauge ``pin``) of the solve."""
    import time
    t0 = time.perf_counter()
    pc = gfm_coef_fast(case)
    b = rhs(case, pc)
    op = sym_operator(case, pc)
    d, _ = sym_diag(case, *op)
    t1 = time.perf_counter()
    o = opts(levels=levels, pre=pre, post=post, coarse_sweeps=coarse_sweeps,
             coarse_scale=coarse_scale, div_tol=div_tol, max_iter=max_iter, omega=omega,
             omega_coarse=omega_coarse)
    x, st, sing = pcg(case, op, b, o)
    x = gauge(case, d, sing, x, pin)
    t2 = time.perf_counter()
    Pp = pressure_field(case, x)
    u, v, w = correct(case, pc, Pp)
    _, dinf = mres(case, u, v, w)
    t3 = time.perf_counter()
    return t3 - t0, dict(assemble_s=t1 - t0, solve_s=t2 - t1, correct_s=t3 - t2,
                         iterations=st.iterations, converged=st.converged, div_inf=dinf,
                         Pp=Pp)


def matrix_rows(case, pc):
    n = case.I * case.J * case.K
    vals = np.zeros(7 * n)
    cols = np.zeros(7 * n, dtype=np.int64)
    g = grid_of(case)
    lib().orc_matrix_rows(ctypes.byref(g), _p(pc), _p(case.ee_area), _p(case.ne_area),
                          _p(case.te_area), _p(case.cell_type_p), _p(case.bc_flag), _p(vals),
                          _p(cols))
    return vals.reshape(n, 7), cols.reshape(n, 7)


def rows_to_scipy(vals, cols):
    import scipy.sparse as sp
    n = vals.shape[0]
    r = np.repeat(np.arange(n), 7)
    c = cols.ravel()
    m = c >= 0
    return sp.csr_matrix((vals.ravel()[m], (r[m], c[m])), shape=(n, n))


def rhs(case, pc, bc_val=None):
    n = case.I * case.J * case.K
    b = np.zeros(n)
    g = grid_of(case)
    bv = (F64P * 6)()
    keep = []
    for f in range(6):
        if bc_val is not None and bc_val[f] is not None:
            a = np.asfortranarray(bc_val[f], dtype=np.float64)
            keep.append(a)
            bv[f] = _p(a)
    lib().orc_rhs(ctypes.byref(g), _p(pc), _p(case.ee_area), _p(case.ne_area), _p(case.te_area),
                  _p(case.cell_type_p), _p(case.bc_flag), bv, _p(case.u_star), _p(case.v_star),
                  _p(case.w_star), _p(case.vof), _p(case.vof_old), case.dt, _p(b))
    return b


def sym_operator(case, pc):
    n = case.I * case.J * case.K
    cx, cy, cz, e = (np.zeros(n) for _ in range(4))
    g = grid_of(case)
    lib().orc_sym_operator(ctypes.byref(g), _p(pc), _p(case.ee_area), _p(case.ne_area),
                           _p(case.te_area), _p(case.cell_type_p), _p(case.bc_flag), _p(cx),
                           _p(cy), _p(cz), _p(e))
    return cx, cy, cz, e


def sym_diag(case, cx, cy, cz, e):
    d = np.zeros_like(cx)
    act = lib().orc_sym_diag(case.I, case.J, case.K, _p(cx), _p(cy), _p(cz), _p(e), _p(d))
    return d, act


def sym_apply(case, cx, cy, cz, d, x):
    y = np.zeros_like(x)
    lib().orc_sym_apply(case.I, case.J, case.K, _p(cx), _p(cy), _p(cz), _p(d), _p(x), _p(y))
    return y


def pcg(case, op, b, o=None, rows=None):
    cx, cy, cz, e = op
    o = o or opts()
    x = np.zeros_like(b)
    st = OrcStats()
    g = grid_of(case)
    vals = cols = None
    if rows is not None:
        vals = np.ascontiguousarray(rows[0].ravel())
        cols = np.ascontiguousarray(rows[1].ravel())
    singular = lib().orc_pcg(ctypes.byref(g), _p(cx), _p(cy), _p(cz), _p(e), _p(vals), _p(cols),
                             _p(np.ascontiguousarray(b)), ctypes.byref(o), _p(x), ctypes.byref(st))
    return x, st, bool(singular)


def gauge(case, d, singular, x, pin=(1, 1, 1)):
    x = x.copy()
    lib().orc_gauge(case.I, case.J, case.K, _p(d), int(singular), pin[0], pin[1], pin[2], _p(x))
    return x


def pressure_field(case, x):
    Pp = np.zeros(case.halo_shape, order="F")
    g = grid_of(case)
    lib().orc_pressure_field(ctypes.byref(g), _p(case.cell_type_p), _p(case.bc_flag),
                             _p(np.ascontiguousarray(x)), _p(Pp))
    return Pp


def correct(case, pc, Pp, u=None, v=None, w=None):
    u = np.zeros(case.halo_shape, order="F") if u is None else u.copy(order="F")
    v = np.zeros(case.halo_shape, order="F") if v is None else v.copy(order="F")
    w = np.zeros(case.halo_shape, order="F") if w is None else w.copy(order="F")
    g = grid_of(case)
    lib().orc_correct(ctypes.byref(g), _p(pc), _p(Pp), _p(case.cell_type_u), _p(case.cell_type_v),
                      _p(case.cell_type_w), _p(case.moex_u), _p(case.moex_v), _p(case.moex_w),
                      _p(case.u_star), _p(case.v_star), _p(case.w_star), _p(u), _p(v), _p(w))
    return u, v, w


def mres(case, u, v, w):
    out = np.zeros((case.I, case.J, case.K), order="F")
    g = grid_of(case)
    m = lib().orc_mres(ctypes.byref(g), _p(case.ee_area), _p(case.ne_area), _p(case.te_area),
                       _p(u), _p(v), _p(w), _p(out))
    return out, m


@dataclasses.dataclass
class OracleResult:
    pocoef: np.ndarray
    b: np.ndarray
    x: np.ndarray
    Pp: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    div_inf: float
    stats: OrcStats
    singular: bool
    op: tuple
    d: np.ndarray


def project(case, pin=(1, 1, 1), tol=None, rows=True, o=None, u_init=None, bc_val=None,
            gfm_fast=False, max_iter=40):
    """Reference path on the CPU.  ``rows=True`` solves the reference's own
    assembled rows (ProjectionP.f90:330-478); ``rows=False`` the engine's
    symmetric face operator.  ``tol`` defaults to a machine-precision stop
    (repeated solves until the divergence stops falling); a given ``tol``
    runs one solve to max |r|/V <= tol or ``max_iter`` iterations, with the
    over-relaxed multigrid of the CPU baseline.  ``gfm_fast``: the parallel
    GFM restatement (bit-identical to the serial one, pinned on every
    configuration).  ``bc_val``: face values of the Dirichlet pressure faces
    (W,E,S,N,B,T)."""
    pc = gfm_coef_fast(case) if gfm_fast else gfm_coef(case)
    b = rhs(case, pc, bc_val)
    op = sym_operator(case, pc)
    d, _ = sym_diag(case, *op)
    if o is None:
        o = (opts(max_iter=1000, div_tol=0.0) if tol is None else
             opts(max_iter=max_iter, div_tol=tol, coarse_scale=0.5, omega=1.25,
                  omega_coarse=1.5))
    R = matrix_rows(case, pc) if rows else None
    if tol is None:
        # converge until the residual stops improving: run at a tiny target
        o.div_tol = 1e-300
        o.max_iter = 1000
        x, st, sing = _pcg_to_floor(case, op, b, o, R)
    else:
        x, st, sing = pcg(case, op, b, o, R)
    x = gauge(case, d, sing, x, pin)
    Pp = pressure_field(case, x)
    u, v, w = correct(case, pc, Pp, u_init, u_init, u_init)
    _, dinf = mres(case, u, v, w)
    return OracleResult(pc, b, x, Pp, u, v, w, dinf, st, sing, op, d)


def _pcg_to_floor(case, op, b, o, R):
    """Run PCG until the divergence stops decreasing (machine-precision floor)."""
    best = None
    o.max_iter = 4
    o.div_tol = 0.0
    x, st, sing = pcg(case, op, b, o, R)
    # iterate by increasing max_iter until no further improvement
    prev = st.final_div
    it = 4
    while it < 400:
        it = int(it * 1.5) + 1
        o.max_iter = it
        x2, st2, sing = pcg(case, op, b, o, R)
        if st2.final_div >= 0.5 * prev:
            if st2.final_div < prev:
                x, st = x2, st2
            break
        x, st, prev = x2, st2, st2.final_div
    return x, st, sing


def adjoint(case, pc, Pp, gu, gv, gw, o=None, gp=None, pin=None):
    """Gradients of L for dL/du, dL/dv, dL/dw and, optionally, dL/dPp = gp
    (halo array) for the pressure in the gauge `pin` (a cell (i, j, k),
    1-based, or (0, 0, 0) for the zero mean); pin is required with gp."""
    o = o or opts(max_iter=1000, div_tol=1e-300)
    if gp is not None:
        assert pin is not None, "the pressure cotangent needs the forward's gauge"
        gp = np.asfortranarray(gp, dtype=np.float64)
    pi, pj, pk = pin if pin is not None else (0, 0, 0)
    hs = case.halo_shape
    gus, gvs, gws = (np.zeros(hs, order="F") for _ in range(3))
    gkx, gky, gkz = (np.zeros((case.I, case.J, case.K), order="F") for _ in range(3))
    lam = np.zeros(case.I * case.J * case.K)
    g = grid_of(case)
    rc = lib().orc_adjoint(ctypes.byref(g), _p(pc), _p(case.ee_area), _p(case.ne_area),
                           _p(case.te_area), _p(case.cell_type_p), _p(case.bc_flag),
                           _p(case.cell_type_u), _p(case.cell_type_v), _p(case.cell_type_w),
                           _p(case.moex_u), _p(case.moex_v), _p(case.moex_w), _p(Pp),
                           _p(np.asfortranarray(gu)), _p(np.asfortranarray(gv)),
                           _p(np.asfortranarray(gw)), _p(gp), pi, pj, pk, ctypes.byref(o),
                           _p(gus), _p(gvs),
                           _p(gws), _p(gkx), _p(gky), _p(gkz), _p(lam))
    return dict(gus=gus, gvs=gvs, gws=gws, gkx=gkx, gky=gky, gkz=gkz, lam=lam, rc=rc)


def wall_plane_shapes(case):
    I, J, K = case.I, case.J, case.K
    return [(J, K), (J, K), (I, K), (I, K), (I, J), (I, J)]


def wall_step(case, flags, consts, u, v, w, p):
    """BoundaryConditionVarNew + mass fix + VariablesInternalCellCondition on
    copies of u, v, w, p.  flags/consts: 24 values (BCu W..T, BCv, BCw, BCp).
    Returns (u, v, w, p, var) with var the 24 face-value arrays."""
    u, v, w, p = (np.array(a, dtype=np.float64, order="F", copy=True) for a in (u, v, w, p))
    fl = np.ascontiguousarray(flags, dtype=np.int32)
    cs = np.ascontiguousarray(consts, dtype=np.float64)
    assert fl.shape == (24,) and cs.shape == (24,)
    shapes = wall_plane_shapes(case)
    var = [np.zeros(shapes[q % 6], order="F") for q in range(24)]
    vp = (F64P * 24)(*[_p(a) for a in var])
    g = grid_of(case)
    lib().orc_wall_step(ctypes.byref(g), _p(case.ee_area), _p(case.ne_area), _p(case.te_area),
                        _p(case.cell_type_p), _p(case.cell_type_u), _p(case.cell_type_v),
                        _p(case.cell_type_w), _p(fl), _p(cs), _p(u), _p(v), _p(w), _p(p), vp)
    return u, v, w, p, var


def face_geo(d1, d2, n1, n2, n3, d3, phi):
    """FaceGeoCal: (area fraction, centroid[3])."""
    a = ctypes.c_double()
    c = (ctypes.c_double * 3)()
    lib().orc_face_geo(d1, d2, n1, n2, n3, d3, phi, ctypes.byref(a), c)
    return a.value, np.array(c[:])


def geom_shapes(I, J, K):
    cell, halo = (I, J, K), (I + 2, J + 2, K + 2)
    sh = {n: cell for n in GEOM_FIELDS}
    for n in ("ee_area", "ne_area", "te_area"):
        sh[n] = halo
    for n in ("fce", "fcn", "fct"):
        sh[n] = halo + (3,)
    sh["cell_cent"] = cell + (3,)
    return sh


def grid_preprocess(cells):
    """GridPreprocess on copies of four dicts of host arrays (P, U, V, W,
    keyed like sp_cell_geom).  Returns four dicts with every field."""
    I, J, K = cells[0]["dx"].shape
    sh = geom_shapes(I, J, K)
    outs, structs = [], []
    for c in cells:
        o = {}
        for n in GEOM_FIELDS:
            dt = np.int32 if n == "cell_type" else np.float64
            if n in c:
                o[n] = np.array(c[n], dtype=dt, order="F", copy=True)
            else:
                o[n] = np.zeros(sh[n], dtype=dt, order="F")
            assert o[n].shape == sh[n], (n, o[n].shape)
        outs.append(o)
        structs.append(OrcCell(*(o[n].ctypes.data for n in GEOM_FIELDS)))
    lib().orc_grid_preprocess(I, J, K, *(ctypes.byref(s) for s in structs))
    return outs


def define_moex(ctp, ctu, ctv, ctw, mu, mv, mw):
    """DefineMomentumExchangeCell on copies of MoExCell of U, V, W."""
    I, J, K = ctp.shape
    cts = [np.asfortranarray(a, dtype=np.int32) for a in (ctp, ctu, ctv, ctw)]
    ms = [np.array(a, dtype=np.int32, order="F", copy=True) for a in (mu, mv, mw)]
    lib().orc_define_moex(I, J, K, *(_p(a) for a in cts), *(_p(a) for a in ms))
    return ms


def number_external(ct, iu, iv, iw):
    """NumberExternalCell: (PosNu of shape (I-iu, J-iv, K-iw), ExtCell)."""
    I, J, K = ct.shape
    ct = np.asfortranarray(ct, dtype=np.int32)
    pos = np.zeros((I - iu, J - iv, K - iw), dtype=np.int32, order="F")
    ext = lib().orc_number_external(I, J, K, iu, iv, iw, _p(ct), _p(pos))
    return pos, int(ext)


def new_cell_face(cells, interp):
    """NewCellFace on copies: cells = four dicts (P, U, V, W) keyed like
    sp_cell_geom (GridPreprocess output), interp = three dicts (U, V, W)
    with moex (and optionally msce and the interpolation arrays' start
    values).  Returns (cells, interp) with the updated copies."""
    I, J, K = cells[0]["dx"].shape
    sh = geom_shapes(I, J, K)
    cs, structs = [], []
    for c in cells:
        o = {n: np.array(c[n], dtype=np.int32 if n == "cell_type" else np.float64, order="F",
                         copy=True) for n in GEOM_FIELDS}
        cs.append(o)
        structs.append(OrcCell(*(o[n].ctypes.data for n in GEOM_FIELDS)))
    halo = (I + 2, J + 2, K + 2)
    xs, xstructs = [], []
    for d in interp:
        o = {"moex": np.array(d["moex"], dtype=np.int32, order="F", copy=True)}
        o["msce"] = np.array(d.get("msce", np.zeros((I, J, K, 3), np.int32)), dtype=np.int32,
                             order="F", copy=True)
        for n in INTERP_FIELDS:
            o[n] = np.array(d.get(n, np.zeros(halo)), dtype=np.float64, order="F", copy=True)
        xs.append(o)
        xstructs.append(OrcInterp(*(o[n].ctypes.data for n in ("moex", "msce") + INTERP_FIELDS)))
    lib().orc_new_cell_face(I, J, K, *(ctypes.byref(s) for s in structs),
                            *(ctypes.byref(s) for s in xstructs))
    return cs, xs


def volume_fraction(dx, dy, dz, nx, ny, nz, s):
    """Volume_Fraction_Calc: the volume on the corner side of the plane."""
    return lib().orc_volume_fraction(dx, dy, dz, nx, ny, nz, s)


def uvw_liquid_field(cells):
    """ComputeUVWLiquidField on copies: cells = four dicts (P, U, V, W) keyed
    by LIQUID_FIELDS (cell arrays); P's are read, U/V/W's phi_l, nx_l,
    ny_l, nz_l and vof_l returned in copies."""
    I, J, K = cells[0]["dx"].shape
    outs, structs = [], []
    for c in cells:
        o = {}
        for n in LIQUID_FIELDS:
            a = c.get(n)
            o[n] = (np.zeros((I, J, K), order="F") if a is None else
                    np.array(a, dtype=np.float64, order="F", copy=True))
        outs.append(o)
        structs.append(OrcLiquid(*(o[n].ctypes.data for n in LIQUID_FIELDS)))
    lib().orc_uvw_liquid_field(I, J, K, *(ctypes.byref(s) for s in structs))
    return outs


# ---------------------------------------------------------------------------
# Momentum predictor (PredictUVW.f90), SpaceOrder 1
# ---------------------------------------------------------------------------

PRED_CELL_FIELDS = ("phi", "nx", "ny", "nz", "vof", "vof_l", "cell_type", "moex", "ee", "ne",
                    "te", "fce", "fcn", "fct", "cent", "sxe", "syn", "szt", "etae", "etan",
                    "etat", "ale", "aln", "alt", "wllh", "delh")


class OrcPredCell(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in PRED_CELL_FIELDS]


class OrcPredParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("itt", ctypes.c_int64), ("time_order", ctypes.c_int32),
                ("space_order", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in ("rho_a_ref", "rho_w_ref", "mu_a_ref", "mu_w_ref", "rey",
                                       "fr", "gx", "gy", "gz", "g")]


def _pred_lib():
    L = lib()
    if not getattr(L, "_pred_ready", False):
        C = ctypes.POINTER(OrcPredCell)
        G = ctypes.POINTER(OrcGrid)
        L.orc_conv_flux_first.argtypes = [G, F64P, F64P, F64P, C, C, C, F64P, F64P, F64P,
                                          ctypes.POINTER(F64P), F64P, F64P, F64P]
        L.orc_conv_flux_first.restype = None
        L.orc_conv_flux_second.argtypes = L.orc_conv_flux_first.argtypes
        L.orc_conv_flux_second.restype = None
        L.orc_predict.argtypes = [G, F64P, F64P, F64P, ctypes.POINTER(C), ctypes.POINTER(C),
                                  ctypes.POINTER(F64P), ctypes.POINTER(F64P), F64P, I32P,
                                  ctypes.POINTER(F64P), ctypes.POINTER(OrcPredParams),
                                  ctypes.POINTER(F64P), ctypes.POINTER(F64P)]
        L.orc_predict.restype = ctypes.c_int
        L._pred_ready = True
    return L


def _pred_cell(d, keep):
    arrs = {}
    for n in PRED_CELL_FIELDS:
        a = d[n]
        dt = np.int32 if n in ("cell_type", "moex") else np.float64
        a = np.asfortranarray(a, dtype=dt)
        keep.append(a)
        arrs[n] = a.ctypes.data
    return OrcPredCell(*(arrs[n] for n in PRED_CELL_FIELDS))


def _f64_array(arrs, keep):
    out = (F64P * len(arrs))()
    for q, a in enumerate(arrs):
        if a is None:
            continue
        a = np.asfortranarray(a, dtype=np.float64)
        keep.append(a)
        out[q] = _p(a)
    return out


def conv_flux_first(grid, p_areas, cells, vel, bcv, second=False):
    """ModifiedConvectiveFluxFirstOrder x3 (ModifiedConvectiveFlux x3, the
    SpaceOrder 2 fluxes, with ``second``): (CFEW, CFNS, CFTB) with shapes
    (I+2,J,K,3), (I,J+2,K,3), (I,J,K+2,3) (index 0 of the extended axis is
    the reference's 0)."""
    L = _pred_lib()
    I, J, K = grid.I, grid.J, grid.K
    keep = []
    cs = [_pred_cell(c, keep) for c in cells]
    pa = [np.asfortranarray(a, dtype=np.float64) for a in p_areas]
    vv = [np.asfortranarray(a, dtype=np.float64) for a in vel]
    cfew = np.zeros((I + 2, J, K, 3), order="F")
    cfns = np.zeros((I, J + 2, K, 3), order="F")
    cftb = np.zeros((I, J, K + 2, 3), order="F")
    g = grid_of(grid)
    f = L.orc_conv_flux_second if second else L.orc_conv_flux_first
    f(ctypes.byref(g), *(_p(a) for a in pa), *(ctypes.byref(c) for c in cs),
      *(_p(a) for a in vv), _f64_array(bcv, keep), _p(cfew), _p(cfns), _p(cftb))
    return cfew, cfns, cftb


def conv_flux_second(grid, p_areas, cells, vel, bcv):
    """ModifiedConvectiveFlux x3 (SpaceOrder 2), shapes as conv_flux_first."""
    return conv_flux_first(grid, p_areas, cells, vel, bcv, second=True)


def second_order_carries():
    """(carried, uncarried) of the last SpaceOrder 2 flux evaluation: cut
    faces whose neighbour tests both failed and that read the delh of an
    earlier face, or had none to read."""
    L = _pred_lib()
    a, b = ctypes.c_int64(), ctypes.c_int64()
    L.orc_conv_flux_second_carries(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def predict(grid, p_areas, old, cur, vel, vel_n, flux_div_old, bc_flag, bcv, params, pred=None):
    """PredictingUVW on copies: returns (pred (u, v, w halo arrays), dp (PU/PV/PW%Dp),
    flux_div_old (I,J,K,3)).  ``old``/``cur``: three dicts keyed by
    PRED_CELL_FIELDS (UCellO.. and UCell..); ``vel``/``vel_n``: TVar and TVar_n
    velocities; ``bc_flag``: 18 flags (BCu, BCv, BCw); ``bcv``: 18 face arrays;
    ``params``: dict of OrcPredParams fields; ``pred``: the Pred arrays before
    the call (entries the predictor does not write are kept), zeros if None."""
    L = _pred_lib()
    I, J, K = grid.I, grid.J, grid.K
    hs = (I + 2, J + 2, K + 2)
    keep = []
    co = (OrcPredCell * 3)(*(_pred_cell(c, keep) for c in old))
    cc = (OrcPredCell * 3)(*(_pred_cell(c, keep) for c in cur))
    po = (ctypes.POINTER(OrcPredCell) * 3)(*(ctypes.pointer(co[q]) for q in range(3)))
    pc = (ctypes.POINTER(OrcPredCell) * 3)(*(ctypes.pointer(cc[q]) for q in range(3)))
    pa = [np.asfortranarray(a, dtype=np.float64) for a in p_areas]
    fdo = np.array(flux_div_old, dtype=np.float64, order="F", copy=True)
    assert fdo.shape == (I, J, K, 3)
    outs = [np.zeros(hs, order="F") if pred is None else
            np.array(pred[q], dtype=np.float64, order="F", copy=True) for q in range(3)]
    dps = [np.zeros(hs, order="F") for _ in range(3)]
    fl = np.ascontiguousarray(bc_flag, dtype=np.int32)
    assert fl.shape == (18,)
    prm = OrcPredParams(**params)
    g = grid_of(grid)
    rc = L.orc_predict(ctypes.byref(g), *(_p(a) for a in pa), po, pc, _f64_array(vel, keep),
                       _f64_array(vel_n, keep), _p(fdo), _p(fl), _f64_array(bcv, keep),
                       ctypes.byref(prm), _f64_array(outs, keep), _f64_array(dps, keep))
    if rc != 0:
        raise ValueError("orc_predict: SpaceOrder 1 and 2 are restated")
    return outs, dps, fdo


# ---------------------------------------------------------------------------
# Solid level set and interface transport (orc_clsvof.c)
# ---------------------------------------------------------------------------

class OrcSolidBody(ctypes.Structure):
    _fields_ = [("imax", ctypes.c_int32), ("jmax", ctypes.c_int32), ("kmax", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("phi", ctypes.c_void_p), ("cm", ctypes.c_double * 3),
                ("rot", ctypes.c_double * 9)]


def _lvs_lib():
    L = lib()
    if not getattr(L, "_lvs_ready", False):
        L.orc_interpolate_lvs.argtypes = ([ctypes.POINTER(OrcSolidBody)] + [ctypes.c_int] * 3
                                          + [F64P] * 11)
        L.orc_interpolate_lvs.restype = None
        L.orc_normal_vector_irre.argtypes = ([ctypes.c_int] * 3 + [F64P] * 4 + [ctypes.c_int] * 3
                                             + [ctypes.POINTER(ctypes.c_double)] * 3)
        L.orc_normal_vector_irre.restype = None
        L._lvs_ready = True
    return L


def solid_body(body, keep):
    """OrcSolidBody from a dict: x, y, z (local centres), dx, dy, dz, phi
    (imax, jmax, kmax), cm (3), rot (3x3 rotMatLtoG)."""
    arrs = {}
    for n in ("x", "y", "z", "phi"):
        a = np.asfortranarray(body[n], dtype=np.float64)
        keep.append(a)
        arrs[n] = a.ctypes.data
    I, J, K = body["phi"].shape
    rot = np.asarray(body["rot"], dtype=np.float64).reshape(3, 3).ravel(order="F")
    return OrcSolidBody(I, J, K, float(body["dx"]), float(body["dy"]), float(body["dz"]),
                        arrs["x"], arrs["y"], arrs["z"], arrs["phi"],
                        (ctypes.c_double * 3)(*np.asarray(body["cm"], dtype=np.float64)),
                        (ctypes.c_double * 9)(*rot))


def interpolate_lvs(body, grid, nx=None, ny=None, nz=None):
    """interpolateLvS of `body` onto one grid: grid maps x, y, z, dx, dy, dz
    to cell arrays (I, J, K); nx, ny, nz are the normals before the call
    (halo arrays, default zero).  Returns phi, vof, nx, ny, nz (halo)."""
    L = _lvs_lib()
    keep = []
    I, J, K = grid["dx"].shape
    g = [np.asfortranarray(grid[n], dtype=np.float64) for n in ("x", "y", "z", "dx", "dy", "dz")]
    hs = (I + 2, J + 2, K + 2)
    phi = np.zeros(hs, order="F")
    vof = np.zeros(hs, order="F")
    n = [np.zeros(hs, order="F") if a is None else np.array(a, dtype=np.float64, order="F",
                                                             copy=True) for a in (nx, ny, nz)]
    B = solid_body(body, keep)
    L.orc_interpolate_lvs(ctypes.byref(B), I, J, K, *(_p(a) for a in g), _p(phi), _p(vof),
                          *(_p(a) for a in n))
    return phi, vof, n[0], n[1], n[2]


def normal_vector_irre(grid, phi, i, j, k):
    """Normal_Vector_Irre at interior cell (i, j, k) (1-based): grid x, y, z
    cell arrays, phi a halo array."""
    L = _lvs_lib()
    I, J, K = grid["x"].shape
    g = [np.asfortranarray(grid[n], dtype=np.float64) for n in ("x", "y", "z")]
    ph = np.asfortranarray(phi, dtype=np.float64)
    out = [ctypes.c_double() for _ in range(3)]
    L.orc_normal_vector_irre(I, J, K, *(_p(a) for a in g), _p(ph), i, j, k,
                             *(ctypes.byref(o) for o in out))
    return tuple(o.value for o in out)


CLSVOF_IN_PTRS = ("u", "v", "w", "x", "y", "z", "dx", "dy", "dz", "nx", "ny", "nz", "bcu_w",
                  "bcv_s", "bcw_b")


class OrcClsvofIn(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in CLSVOF_IN_PTRS] + [
        ("lvs_flag", ctypes.c_int32 * 6), ("lvs_const", ctypes.c_double * 6),
        ("vof_flag", ctypes.c_int32 * 6), ("vof_const", ctypes.c_double * 6),
        ("dt", ctypes.c_double)]


class OrcClsvofState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("vof", "phi", "nx", "ny", "nz")] + [
        ("lvs_var", ctypes.c_void_p * 6), ("vof_var", ctypes.c_void_p * 6)]


def _clsvof_lib():
    L = _lvs_lib()
    if not getattr(L, "_clsvof_ready", False):
        L.orc_clsvof_scheme.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(OrcClsvofIn),
                                                             ctypes.POINTER(OrcClsvofState)]
        L.orc_clsvof_scheme.restype = ctypes.c_int
        L.orc_find_distance.argtypes = [ctypes.c_double] * 7
        L.orc_find_distance.restype = ctypes.c_double
        L.orc_redist_candidate.argtypes = [ctypes.c_double] * 7 + [ctypes.c_int] * 3 + [
            ctypes.c_double]
        L.orc_redist_candidate.restype = ctypes.c_double
        L._clsvof_ready = True
    return L


def find_distance(dx, dy, dz, nx, ny, nz, f):
    return _clsvof_lib().orc_find_distance(dx, dy, dz, nx, ny, nz, f)


def redist_candidate(nx, ny, nz, dis, dx, dy, dz, ii, jj, kk, phit):
    return _clsvof_lib().orc_redist_candidate(nx, ny, nz, dis, dx, dy, dz, ii, jj, kk, phit)


def clsvof_face_shapes(shape):
    """Shapes of the W, E, S, N, B, T face arrays: (J,K), (J,K), (I,K), (I,K),
    (I,J), (I,J)."""
    I, J, K = shape
    return [(J, K), (J, K), (I, K), (I, K), (I, J), (I, J)]


def clsvof_scheme(inp, state):
    """Clsvof_Scheme on copies.  inp maps u, v, w, nx, ny, nz (halo), x, y,
    z, dx, dy, dz (cell), bcu_w, bcv_s, bcw_b (face arrays), lvs_flag,
    lvs_const, vof_flag, vof_const (6 each) and dt; state maps vof_l, phi_l,
    nx_l, ny_l, nz_l (halo) and lvs_var, vof_var (six face arrays each).
    Returns the updated state."""
    L = _clsvof_lib()
    I, J, K = inp["dx"].shape
    keep = []
    ptr = {}
    for n in CLSVOF_IN_PTRS:
        a = np.asfortranarray(inp[n], dtype=np.float64)
        keep.append(a)
        ptr[n] = a.ctypes.data
    ci = OrcClsvofIn(*(ptr[n] for n in CLSVOF_IN_PTRS),
                     (ctypes.c_int32 * 6)(*inp["lvs_flag"]), (ctypes.c_double * 6)(*inp["lvs_const"]),
                     (ctypes.c_int32 * 6)(*inp["vof_flag"]), (ctypes.c_double * 6)(*inp["vof_const"]),
                     float(inp["dt"]))
    out = {n: np.array(state[n], dtype=np.float64, order="F", copy=True)
           for n in ("vof_l", "phi_l", "nx_l", "ny_l", "nz_l")}
    shapes = clsvof_face_shapes((I, J, K))
    out["lvs_var"] = [np.array(a, dtype=np.float64, order="F", copy=True).reshape(s, order="F")
                      for a, s in zip(state["lvs_var"], shapes)]
    out["vof_var"] = [np.array(a, dtype=np.float64, order="F", copy=True).reshape(s, order="F")
                      for a, s in zip(state["vof_var"], shapes)]
    cs = OrcClsvofState(*(out[n].ctypes.data for n in ("vof_l", "phi_l", "nx_l", "ny_l", "nz_l")),
                        (ctypes.c_void_p * 6)(*(a.ctypes.data for a in out["lvs_var"])),
                        (ctypes.c_void_p * 6)(*(a.ctypes.data for a in out["vof_var"])))
    rc = L.orc_clsvof_scheme(I, J, K, ctypes.byref(ci), ctypes.byref(cs))
    if rc != 0:
        raise ValueError(f"orc_clsvof_scheme: {rc}")
    return out
