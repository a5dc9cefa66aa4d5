"""Python mirror of SIMCOFlow's projection interface on the HIP engine.

* :func:`poisson_equation_solver` follows ``PoissonEquationSolver``
  (ProjectionP.f90:33-121) followed by the correction block of
  ``UpdatePUVW`` (ComputePUVW.f90:179-232): same arguments (grids, cells,
  old solid vof, variables, predictor, d-coefficients, pressure BC,
  PoCoef, projection, dt), same outputs written in place (PoCoef, Proj%Pp,
  TVar%u/v/w/p).
* :func:`boundary_condition_var_new` follows the lines after the correction
  in ``UpdatePUVW``: ``BoundaryConditionVarNew`` (ComputePUVW.f90:418-694),
  the mass fix of the first Dirichlet pressure face (:256-384) and
  ``VariablesInternalCellCondition`` (Solver.f90:467-501), on the engine.
* :func:`grid_preprocess` follows ``GridPreprocess`` (CutCell.f90:62-325):
  the cut-cell geometry of the four staggered grids, written into the
  Cell containers in place.
* :class:`DeviceProjection` runs the projection on torch tensors resident on
  the GPU and keeps the forward pressure for the adjoint.
* :class:`ProjectionFunction` is the ``torch.autograd.Function`` whose
  backward is the engine's hand-written adjoint: gradients of a loss on the
  corrected velocities and the pressure with respect to u*, v*, w* and to
  the face coefficients beta = 1/rho' (E, N, T faces).

Array layouts are the host's Fortran arrays.  On the torch side a Fortran
(I+2, J+2, K+2) halo array is a C-contiguous tensor of shape (K+2, J+2, I+2)
(same bytes); cell arrays are (K, J, I).
"""

import dataclasses
from typing import Optional

import numpy as np

from . import engine as _eng

# ---------------------------------------------------------------------------
# Fortran-shaped containers (only the members the projection reads or writes)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Grid:  # Mesh.f90:14-18 (tensor product: 1-D spacings)
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray


@dataclasses.dataclass
class Cell:  # CutCell.f90:13-39
    Cell_Type: np.ndarray
    vof: Optional[np.ndarray] = None
    vofL: Optional[np.ndarray] = None
    phiL: Optional[np.ndarray] = None
    EEArea: Optional[np.ndarray] = None
    NEArea: Optional[np.ndarray] = None
    TEArea: Optional[np.ndarray] = None
    MoExCell: Optional[np.ndarray] = None
    phi: Optional[np.ndarray] = None         # solid level set and normal
    nx: Optional[np.ndarray] = None
    ny: Optional[np.ndarray] = None
    nz: Optional[np.ndarray] = None
    FCE: Optional[np.ndarray] = None         # (0:I+1, 0:J+1, 0:K+1, 3)
    FCN: Optional[np.ndarray] = None
    FCT: Optional[np.ndarray] = None
    Cell_Cent: Optional[np.ndarray] = None   # (I, J, K, 3)
    WlLh: Optional[np.ndarray] = None
    delh: Optional[np.ndarray] = None


@dataclasses.dataclass
class Variables:  # StateVariables.f90:21-26
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    p: np.ndarray


@dataclasses.dataclass
class Predictor:  # PredictUVW.f90:36-38
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray


@dataclasses.dataclass
class PoissonCoefficient:  # PredictUVW.f90:39-41
    Dp: np.ndarray


@dataclasses.dataclass
class BCBase:  # BoundaryInterface.f90:7-21
    flag: np.ndarray  # W,E,S,N,B,T: 0 Dirichlet, 1 Neumann
    Const: Optional[np.ndarray] = None  # %Const(1..6) of TypicalBC
    VarW: Optional[np.ndarray] = None
    VarE: Optional[np.ndarray] = None
    VarS: Optional[np.ndarray] = None
    VarN: Optional[np.ndarray] = None
    VarB: Optional[np.ndarray] = None
    VarT: Optional[np.ndarray] = None


@dataclasses.dataclass
class Projection:  # ProjectionP.f90:21-23
    Pp: np.ndarray


class ProjectionError(RuntimeError):
    pass


_projectors = {}


def _projector(shape, params):
    key = (shape, tuple(sorted(params.items())))
    P = _projectors.get(key)
    if P is None:
        P = _eng.Projector(shape, **params)
        _projectors[key] = P
    return P


def poisson_equation_solver(PGrid, UGrid, VGrid, WGrid, PCell, UCell, VCell, WCell, vofO, TVar,
                            TPred, PU, PV, PW, BCp, PoCoef, Proj, dt, *, rho_l=998.3, rho_g=1.205,
                            precision=_eng.SP_PREC_FP64, div_tol=1e-10, pin=(0, 0, 0), **solver):
    """Projection with the reference's argument list; outputs in place.

    Writes PoCoef (Imax,Jmax,Kmax,6), Proj.Pp (with ghosts) and the corrected
    TVar.u/v/w on (1:I,1:J,1:K) plus TVar.p = Pp there, like
    ComputePUVW.f90:181-221.  Returns the engine statistics."""
    I, J, K = PCell.Cell_Type.shape
    params = dict(precision=precision, div_tol=div_tol, rho_l=rho_l, rho_g=rho_g, rho_ref=rho_l,
                  pin_i=pin[0], pin_j=pin[1], pin_k=pin[2], **solver)
    P = _projector((I, J, K), params)
    P.set_grid(PGrid.dx, PGrid.dy, PGrid.dz, UGrid.dx, VGrid.dy, WGrid.dz)
    f = dict(u_star=TPred.u, v_star=TPred.v, w_star=TPred.w, dp_u=PU.Dp, dp_v=PV.Dp,
             dp_w=PW.Dp, ee_area=PCell.EEArea, ne_area=PCell.NEArea, te_area=PCell.TEArea,
             vof=PCell.vof, vof_old=vofO, vof_l=PCell.vofL, phi_l=PCell.phiL,
             cell_type_p=PCell.Cell_Type, cell_type_u=UCell.Cell_Type,
             cell_type_v=VCell.Cell_Type, cell_type_w=WCell.Cell_Type, moex_u=UCell.MoExCell,
             moex_v=VCell.MoExCell, moex_w=WCell.MoExCell)
    f = {k: np.asfortranarray(v) for k, v in f.items()}
    bcv = [BCp.VarW, BCp.VarE, BCp.VarS, BCp.VarN, BCp.VarB, BCp.VarT]
    bcv = [None if v is None else np.asfortranarray(v, dtype=np.float64) for v in bcv]
    out = {"p": Proj.Pp, "u": TVar.u, "v": TVar.v, "w": TVar.w}
    for k, a in out.items():
        if not (isinstance(a, np.ndarray) and a.flags.f_contiguous and a.dtype == np.float64):
            raise ProjectionError(f"{k} must be a Fortran-ordered float64 array")
    res, st = P.project(f, np.asarray(BCp.flag, dtype=np.int32), dt, out=out, bc_val=bcv,
                        po_coef=True)
    PoCoef[...] = res["po_coef"]
    TVar.p[1:-1, 1:-1, 1:-1] = Proj.Pp[1:-1, 1:-1, 1:-1]
    return st


def boundary_condition_var_new(PGrid, PCell, UCell, VCell, WCell, TVar, BCp, BCu, BCv, BCw):
    """BoundaryConditionVarNew + the Dirichlet-pressure mass fix +
    VariablesInternalCellCondition on the engine (sp_wall_step), with the
    reference's arguments.  TVar.u/v/w/p are updated in place; BCu/BCv/BCw/BCp
    VarW..VarT receive the new face values.  Returns the engine statistics
    (wall_div = max |mres|/V of the result, Solver.f90:236-250)."""
    I, J, K = PCell.Cell_Type.shape
    for a in (PGrid.dx, PGrid.dy, PGrid.dz):
        if np.shape(a) != (I, J, K):
            raise ProjectionError("PGrid%dx, dy, dz must be (Imax, Jmax, Kmax) arrays")
    P = _projector((I, J, K), dict(pin_i=0, pin_j=0, pin_k=0))
    # the wall step reads only the P-cell spacings; the staggered ones of
    # this context are never used
    P.set_grid(PGrid.dx, PGrid.dy, PGrid.dz, PGrid.dx, PGrid.dy, PGrid.dz)
    f = dict(ee_area=PCell.EEArea, ne_area=PCell.NEArea, te_area=PCell.TEArea,
             cell_type_p=PCell.Cell_Type, cell_type_u=UCell.Cell_Type,
             cell_type_v=VCell.Cell_Type, cell_type_w=WCell.Cell_Type)
    f = {k: np.asfortranarray(v) for k, v in f.items()}
    flags = np.concatenate([np.asarray(b.flag, dtype=np.int32) for b in (BCu, BCv, BCw)])
    consts = np.concatenate([np.asarray(b.Const, dtype=np.float64) for b in (BCu, BCv, BCw, BCp)])
    wall = _eng.make_wall(flags, consts)
    res, st = P.wall_step(f, np.asarray(BCp.flag, dtype=np.int32), wall, TVar.u, TVar.v, TVar.w,
                          TVar.p)
    for n in ("u", "v", "w", "p"):
        getattr(TVar, n)[...] = res[n]
    for q, b in enumerate((BCu, BCv, BCw, BCp)):
        for f_, name in enumerate(("VarW", "VarE", "VarS", "VarN", "VarB", "VarT")):
            setattr(b, name, res["bc_var"][6 * q + f_])
    return st


# sp_cell_geom member -> Cell member
_GEOM_CELL = dict(phi="phi", nx="nx", ny="ny", nz="nz", vof="vof", vof_l="vofL",
                  cell_type="Cell_Type", ee_area="EEArea", ne_area="NEArea", te_area="TEArea",
                  fce="FCE", fcn="FCN", fct="FCT", cell_cent="Cell_Cent", wllh="WlLh",
                  delh="delh")


def grid_preprocess(PGrid, UGrid, VGrid, WGrid, PCell, UCell, VCell, WCell):
    """GridPreprocess (CutCell.f90:62-325) on the engine (sp_grid_preprocess),
    with the reference's arguments.  Grids carry 3-D dx, dy, dz; each Cell
    needs phi, nx, ny, nz, vof, vofL and delh (and FCE/FCN/FCT where their
    untouched entries matter); Cell_Type, vof, vofL, EEArea, NEArea,
    TEArea, FCE, FCN, FCT, Cell_Cent, WlLh and delh are written back (the
    ones that are None are allocated).  phi, nx, ny, nz, vof and vofL may
    be (0:Imax+1, 0:Jmax+1, 0:Kmax+1) arrays, as the drivers allocate them
    (FlowSphere2Phases.f90:316-322), or (Imax, Jmax, Kmax) ones, as
    AllocateVar does (Main.f90:264-270): PCell%phi's shape says which."""
    I, J, K = np.shape(PGrid.dx)
    halo = np.shape(PCell.phi) == (I + 2, J + 2, K + 2)
    sh = _eng.geom_shapes(I, J, K, halo)
    cells = []
    for G, C in ((PGrid, PCell), (UGrid, UCell), (VGrid, VCell), (WGrid, WCell)):
        d = {}
        for n in ("dx", "dy", "dz"):
            a = getattr(G, n)
            if np.shape(a) != (I, J, K):
                raise ProjectionError(f"TGrid%{n} must be an (Imax, Jmax, Kmax) array")
            d[n] = a
        for n, m in _GEOM_CELL.items():
            a = getattr(C, m)
            if a is None:
                if n in ("phi", "nx", "ny", "nz", "vof", "vof_l", "delh"):
                    raise ProjectionError(f"TCell%{m} is required")
                a = np.zeros(sh[n], dtype=np.int32 if n == "cell_type" else np.float64,
                             order="F")
            if np.shape(a) != sh[n]:
                raise ProjectionError(f"TCell%{m} must have shape {sh[n]}")
            d[n] = a
        cells.append(d)
    P = _projector((I, J, K), dict(pin_i=0, pin_j=0, pin_k=0))
    outs = P.grid_preprocess(cells, halo_fields=halo)
    for C, o in zip((PCell, UCell, VCell, WCell), outs):
        for n, m in _GEOM_CELL.items():
            if n in ("phi", "nx", "ny", "nz"):
                continue
            cur = getattr(C, m)
            if cur is None or not isinstance(cur, np.ndarray) or cur.dtype != o[n].dtype:
                setattr(C, m, o[n])
            else:
                cur[...] = o[n]


# ---------------------------------------------------------------------------
# Device path and autograd
# ---------------------------------------------------------------------------


def to_torch_halo(a, device="cuda"):
    import torch
    return torch.from_numpy(np.asfortranarray(a).T.copy()).to(device)


def to_torch_cell(a, device="cuda"):
    import torch
    return torch.from_numpy(np.asfortranarray(a).T.copy()).to(device)


class DeviceProjection:
    """Projection of fixed geometry on device tensors.

    ``geometry`` holds the sp_inputs fields other than u*, v*, w* as device
    tensors (d-coefficients, apertures, vof, level set, cell flags);
    ``bc_flag`` the six pressure BC flags (Neumann only here: the beta
    parametrisation leaves Dirichlet wall terms fixed)."""

    def __init__(self, shape, grid, geometry, bc_flag, dt, device_index=0, **params):
        import torch
        self.shape = tuple(shape)
        I, J, K = self.shape
        self.torch = torch
        self.geometry = geometry
        self.bc_flag = [int(v) for v in bc_flag]
        self.dt = float(dt)
        self.P = _eng.Projector(self.shape, device=device_index, **params)
        self.P.set_grid_1d(*grid)
        self.grid = [np.asarray(g, dtype=np.float64) for g in grid]
        self.forward_id = 0
        self.hs = (K + 2, J + 2, I + 2)
        self.cs = (K, J, I)

    def _fields(self, us, vs, ws):
        f = dict(self.geometry)
        f.update(u_star=us, v_star=vs, w_star=ws)
        return f

    def forward(self, us, vs, ws, beta=None, po_coef=False):
        torch = self.torch
        out = {n: torch.zeros(self.hs, dtype=torch.float64, device=us.device)
               for n in ("p", "u", "v", "w")}
        if po_coef:
            out["po_coef"] = torch.zeros((6,) + self.cs, dtype=torch.float64, device=us.device)
        tens = [us, vs, ws] + (list(beta) if beta is not None else [])
        for t in tens:
            if not t.is_contiguous() or t.dtype != torch.float64:
                raise ProjectionError("inputs must be contiguous float64 device tensors")
        torch.cuda.synchronize()
        st = self.P.project_dev(self._fields(us, vs, ws), self.bc_flag, self.dt, out,
                                beta=beta)
        self.forward_id += 1
        self.last_stats = st
        return out, st

    def gfm_beta(self, us, vs, ws):
        """beta = PoCoef(:,4..6) * delta / Dp from the GFM coefficients."""
        out, _ = self.forward(us, vs, ws, po_coef=True)
        torch = self.torch
        I, J, K = self.shape
        dxu, dyv, dzw = (torch.from_numpy(self.grid[q]).to(us.device) for q in (3, 4, 5))
        pc = out["po_coef"]
        dp = [self.geometry[n][1:-1, 1:-1, 1:-1] for n in ("dp_u", "dp_v", "dp_w")]
        be = pc[3] * dxu.view(1, 1, I) / dp[0]
        bn = pc[4] * dyv.view(1, J, 1) / dp[1]
        bt = pc[5] * dzw.view(K, 1, 1) / dp[2]
        return [be.contiguous(), bn.contiguous(), bt.contiguous()]

    def adjoint(self, gu, gv, gw, gp=None):
        """Gradients for dL/du, dL/dv, dL/dw and, if given, dL/dPp (device
        halo tensors)."""
        torch = self.torch
        res = {n: torch.zeros(self.hs, dtype=torch.float64, device=gu.device)
               for n in ("gu_star", "gv_star", "gw_star")}
        for n in ("gk_e", "gk_n", "gk_t", "gbeta_e", "gbeta_n", "gbeta_t"):
            res[n] = torch.zeros(self.cs, dtype=torch.float64, device=gu.device)
        g = {"gu": gu.contiguous(), "gv": gv.contiguous(), "gw": gw.contiguous()}
        if gp is not None:
            g["gp"] = gp.contiguous()
        torch.cuda.synchronize()
        st = self.P.adjoint_dev(g, res)
        return res, st


def _ProjectionFunction():
    import torch

    class ProjectionFunction(torch.autograd.Function):
        """(u*, v*, w*, beta_e, beta_n, beta_t) -> (u, v, w, p).

        The backward reuses the engine's multigrid PCG on the adjoint source;
        only the forward pressure is kept (if another forward ran on the same
        engine in between, the forward is recomputed first)."""

        @staticmethod
        def forward(ctx, dp_engine, us, vs, ws, be, bn, bt):
            out, _ = dp_engine.forward(us, vs, ws, beta=[be, bn, bt])
            ctx.dp_engine = dp_engine
            ctx.fid = dp_engine.forward_id
            ctx.save_for_backward(us, vs, ws, be, bn, bt)
            return out["u"], out["v"], out["w"], out["p"]

        @staticmethod
        def backward(ctx, gu, gv, gw, gp):
            E = ctx.dp_engine
            us, vs, ws, be, bn, bt = ctx.saved_tensors
            if E.forward_id != ctx.fid:
                E.forward(us, vs, ws, beta=[be, bn, bt])
            z = torch.zeros_like(us)
            res, _ = E.adjoint(gu if gu is not None else z, gv if gv is not None else z,
                               gw if gw is not None else z, gp=gp)
            return (None, res["gu_star"], res["gv_star"], res["gw_star"], res["gbeta_e"],
                    res["gbeta_n"], res["gbeta_t"])

    return ProjectionFunction


_pf = None


def projection_function():
    global _pf
    if _pf is None:
        _pf = _ProjectionFunction()
    return _pf


def project(dp_engine, us, vs, ws, beta):
    """Differentiable projection: returns (u, v, w, p) tensors."""
    return projection_function().apply(dp_engine, us, vs, ws, *beta)
