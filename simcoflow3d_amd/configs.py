"""Synthetic projection inputs for the five benchmark configurations.

``generate(cfg)`` returns a :class:`ProjectionCase` holding every array a
SIMCOFlow host passes to the projection, in the host's own layouts:

* halo fields ``(I+2, J+2, K+2)`` Fortran-ordered float64 (index 0 is the
  ghost layer ``0`` of the Fortran bounds ``0:Imax+1``);
* cell flags ``(I, J, K)`` Fortran-ordered int32;
* 1-D grid spacings.

Configurations (all no-slip walls, Neumann pressure, rho_l=1000, rho_g=1):

0. 64^3 unit cube, interface z=0.5+0.05 sin(2 pi x) cos(2 pi y), sphere
   r=0.15 at (0.5,0.5,0.3), u* ~ U(-1,1), dt=1e-3.
1. 128^3, z stretched by 1.02 per cell toward z=0.6, 8-mode seeded Fourier
   interface of amplitude 0.02 about z=0.6, cylinder r=0.1 along x at
   (y,z)=(0.5,0.35).
2. 256^3, interface z=0.5+0.1 tanh(20(x-0.5))(1+0.3 sin 4 pi y), no solid.
3. 128^3 copy of configuration 0 (used for the adjoint).
4. 512x256x256 on [0,2]x[0,1]^2, wave z=0.5+0.03 cos(2 pi x + phase), three
   seeded submerged boxes, u* ~ U(-0.5,0.5).

Any configuration can be generated at another size for tests.
"""

import ctypes
import dataclasses
import os

import numpy as np

from ._paths import GEN_LIB

_F64P = ctypes.POINTER(ctypes.c_double)
_I32P = ctypes.POINTER(ctypes.c_int32)


class _SgFields(ctypes.Structure):
    _fields_ = [(n, _F64P) for n in (
        "dxp", "dyp", "dzp", "dxu", "dyv", "dzw", "xc", "yc", "zc",
        "u_star", "v_star", "w_star", "dp_u", "dp_v", "dp_w",
        "ee_area", "ne_area", "te_area", "vof", "vof_old", "vof_l", "phi_l")] + [
        (n, _I32P) for n in (
            "cell_type_p", "cell_type_u", "cell_type_v", "cell_type_w",
            "moex_u", "moex_v", "moex_w")]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(GEN_LIB):
            raise RuntimeError(f"generator library missing: {GEN_LIB} (run __graft_entry__.build())")
        lib = ctypes.CDLL(GEN_LIB)
        lib.sg_generate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(_SgFields)]
        lib.sg_generate.restype = ctypes.c_int
        lib.sg_config_dims.argtypes = [ctypes.c_int, _I32P]
        lib.sg_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int64, _F64P]
        lib.sg_uniform.restype = None
        _lib = lib
    return _lib


HALO_FIELDS = ("u_star", "v_star", "w_star", "dp_u", "dp_v", "dp_w", "ee_area", "ne_area",
               "te_area", "vof", "vof_old", "vof_l", "phi_l")
FLAG_FIELDS = ("cell_type_p", "cell_type_u", "cell_type_v", "cell_type_w",
               "moex_u", "moex_v", "moex_w")


@dataclasses.dataclass
class ProjectionCase:
    cfg: int
    I: int
    J: int
    K: int
    seed: int
    dt: float
    rho_l: float
    rho_g: float
    dxp: np.ndarray
    dyp: np.ndarray
    dzp: np.ndarray
    dxu: np.ndarray
    dyv: np.ndarray
    dzw: np.ndarray
    xc: np.ndarray
    yc: np.ndarray
    zc: np.ndarray
    fields: dict
    bc_flag: np.ndarray

    @property
    def shape(self):
        return (self.I, self.J, self.K)

    @property
    def halo_shape(self):
        return (self.I + 2, self.J + 2, self.K + 2)

    def __getattr__(self, name):
        f = self.__dict__.get("fields")
        if f is not None and name in f:
            return f[name]
        raise AttributeError(name)


def config_dims(cfg):
    d = (ctypes.c_int32 * 3)()
    if _load().sg_config_dims(cfg, d) != 0:
        raise ValueError(f"unknown configuration {cfg}")
    return tuple(d)


def uniform(seed, stream, lo, hi, shape):
    """Seeded counter-based U(lo,hi) draws (Fortran order)."""
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.float64)
    _load().sg_uniform(seed, stream, lo, hi, n, out.ctypes.data_as(_F64P))
    return out.reshape(shape, order="F")


def generate(cfg, size=None, seed=0, dt=1.0e-3):
    """Build configuration ``cfg`` (0..4), optionally at another grid size."""
    lib = _load()
    I, J, K = size if size is not None else config_dims(cfg)
    hs = (I + 2, J + 2, K + 2)
    arrs = {}
    for n in ("dxp", "dxu", "xc"):
        arrs[n] = np.zeros(I)
    for n in ("dyp", "dyv", "yc"):
        arrs[n] = np.zeros(J)
    for n in ("dzp", "dzw", "zc"):
        arrs[n] = np.zeros(K)
    for n in HALO_FIELDS:
        arrs[n] = np.zeros(hs, order="F")
    for n in FLAG_FIELDS:
        arrs[n] = np.zeros((I, J, K), dtype=np.int32, order="F")
    st = _SgFields()
    for n, _ in _SgFields._fields_:
        a = arrs[n]
        setattr(st, n, a.ctypes.data_as(_I32P if a.dtype == np.int32 else _F64P))
    rc = lib.sg_generate(cfg, I, J, K, seed, dt, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"sg_generate failed ({rc})")
    fields = {n: arrs[n] for n in HALO_FIELDS + FLAG_FIELDS}
    return ProjectionCase(cfg=cfg, I=I, J=J, K=K, seed=seed, dt=dt, rho_l=1000.0, rho_g=1.0,
                          dxp=arrs["dxp"], dyp=arrs["dyp"], dzp=arrs["dzp"], dxu=arrs["dxu"],
                          dyv=arrs["dyv"], dzw=arrs["dzw"], xc=arrs["xc"], yc=arrs["yc"],
                          zc=arrs["zc"], fields=fields,
                          bc_flag=np.ones(6, dtype=np.int32))


# ---------------------------------------------------------------------------
# Cut-cell geometry inputs (GridPreprocess, CutCell.f90:62-325)
# ---------------------------------------------------------------------------

GEOM_SHAPES = ("sphere", "plane", "cylinder", "bodies", "axis_plane")


def _centres(d):
    return np.cumsum(d) - 0.5 * d


def staggered_spacings(dxp, dyp, dzp):
    """Cell centres and 3-D spacing arrays of the P, U, V, W grids as the
    reference builds them (Mesh.f90:160-247): a U cell is centred on the
    east face of its P cell, its width is the distance to the next P centre
    (the P width in the last cell); V and W likewise."""
    xc, yc, zc = _centres(dxp), _centres(dyp), _centres(dzp)
    I, J, K = len(dxp), len(dyp), len(dzp)

    def stag(c, d):
        w = np.empty_like(d)
        w[:-1] = c[1:] - c[:-1]
        w[-1] = d[-1]
        return c + 0.5 * d, w

    xu, dxu = stag(xc, dxp)
    yv, dyv = stag(yc, dyp)
    zw, dzw = stag(zc, dzp)
    grids = []
    for (x, dx), (y, dy), (z, dz) in (((xc, dxp), (yc, dyp), (zc, dzp)),
                                      ((xu, dxu), (yc, dyp), (zc, dzp)),
                                      ((xc, dxp), (yv, dyv), (zc, dzp)),
                                      ((xc, dxp), (yc, dyp), (zw, dzw))):
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        DX, DY, DZ = np.meshgrid(dx, dy, dz, indexing="ij")
        grids.append(dict(x=np.asfortranarray(X), y=np.asfortranarray(Y), z=np.asfortranarray(Z),
                          dx=np.asfortranarray(DX), dy=np.asfortranarray(DY),
                          dz=np.asfortranarray(DZ)))
    assert all(g["dx"].shape == (I, J, K) for g in grids)
    return grids


def solid_level_set(shape, X, Y, Z, rng):
    """Signed distance to a solid (positive in the fluid) and its unit
    normal (the gradient, pointing into the fluid)."""
    if shape == "sphere":
        c = np.array([0.5, 0.5, 0.45]) + rng.uniform(-0.05, 0.05, 3)
        r = 0.3
        dX, dY, dZ = X - c[0], Y - c[1], Z - c[2]
        rho = np.sqrt(dX * dX + dY * dY + dZ * dZ)
        rho = np.where(rho > 0.0, rho, 1.0)
        return rho - r, dX / rho, dY / rho, dZ / rho
    if shape in ("plane", "axis_plane"):
        if shape == "plane":
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
        else:
            n = np.zeros(3)
            n[rng.integers(0, 3)] = 1.0 if rng.uniform() < 0.5 else -1.0
        d = float(n @ np.array([0.5, 0.5, 0.5])) + rng.uniform(-0.1, 0.1)
        phi = n[0] * X + n[1] * Y + n[2] * Z - d
        ones = np.ones_like(X)
        return phi, n[0] * ones, n[1] * ones, n[2] * ones
    if shape == "cylinder":
        yc, zc = 0.5 + rng.uniform(-0.05, 0.05, 2)
        dY, dZ = Y - yc, Z - zc
        rho = np.sqrt(dY * dY + dZ * dZ)
        rho = np.where(rho > 0.0, rho, 1.0)
        return rho - 0.2, np.zeros_like(X), dY / rho, dZ / rho
    if shape == "bodies":
        best = None
        for _ in range(3):
            c = rng.uniform(0.2, 0.8, 3)
            r = rng.uniform(0.08, 0.2)
            dX, dY, dZ = X - c[0], Y - c[1], Z - c[2]
            rho = np.sqrt(dX * dX + dY * dY + dZ * dZ)
            rho = np.where(rho > 0.0, rho, 1.0)
            cand = (rho - r, dX / rho, dY / rho, dZ / rho)
            if best is None:
                best = cand
            else:
                take = cand[0] < best[0]
                best = tuple(np.where(take, a, b) for a, b in zip(cand, best))
        return best
    raise ValueError(f"unknown solid shape {shape}")


def geometry_case(size, shape="sphere", seed=0, stretch=False, edge_values=False):
    """Inputs of GridPreprocess for the P, U, V, W grids on [0,1]^3: a list
    of four dicts keyed like sp_cell_geom (Fortran-ordered arrays).  The
    solid vof is a smooth function of the level set; the arrays the
    reference updates in place (vof, vof_l, cell_type, fce/fcn/fct, delh)
    hold seeded values.  ``stretch`` grows dz by 2% per cell;
    ``edge_values`` sets some vof to the DefineCell thresholds, 0, 1 and NaN."""
    I, J, K = size
    rng = np.random.default_rng(seed)
    dxp = np.full(I, 1.0 / I)
    dyp = np.full(J, 1.0 / J)
    dzp = np.full(K, 1.0 / K)
    if stretch:
        dzp = 1.02 ** np.arange(K)
        dzp /= dzp.sum()
    grids = staggered_spacings(dxp, dyp, dzp)
    hs = (I + 2, J + 2, K + 2)
    cells = []
    shape_seed = int(rng.integers(0, 2**31))
    for g in grids:
        phi, nx, ny, nz = solid_level_set(shape, g["x"], g["y"], g["z"],
                                          np.random.default_rng(shape_seed))
        h = np.abs(nx) * g["dx"] + np.abs(ny) * g["dy"] + np.abs(nz) * g["dz"]
        vof = np.clip(0.5 + phi / h, 0.0, 1.0)
        if edge_values:
            sel = rng.uniform(size=vof.shape) < 0.02
            vals = np.array([1.0e-3, 1.0 - 1.0e-3, 0.0, 1.0, 0.5, np.nan])
            vof = np.where(sel, vals[rng.integers(0, len(vals), vof.shape)], vof)
        c = dict(dx=g["dx"], dy=g["dy"], dz=g["dz"])
        for n, a in (("phi", phi), ("nx", nx), ("ny", ny), ("nz", nz), ("vof", vof)):
            c[n] = np.asfortranarray(a, dtype=np.float64)
        c["vof_l"] = np.asfortranarray(vof * rng.uniform(0.0, 1.0, vof.shape))
        c["cell_type"] = np.asfortranarray(rng.integers(0, 3, (I, J, K)).astype(np.int32))
        for n in ("fce", "fcn", "fct"):
            c[n] = np.asfortranarray(rng.uniform(-1.0, 1.0, hs + (3,)))
        c["delh"] = np.asfortranarray(rng.uniform(0.0, 1.0, (I, J, K)))
        cells.append(c)
    return cells


# ---------------------------------------------------------------------------
# Momentum predictor inputs (PredictingUVW, PredictUVW.f90:49-803)
# ---------------------------------------------------------------------------

# members of one U/V/W cell set the predictor reads, and where they come
# from in the geometry update's outputs (sp_cell_geom / sp_cell_interp names)
PRED_CELL_FIELDS = ("phi", "nx", "ny", "nz", "vof", "vof_l", "cell_type", "moex", "ee", "ne",
                    "te", "fce", "fcn", "fct", "cent", "sxe", "syn", "szt", "etae", "etan",
                    "etat", "ale", "aln", "alt", "wllh", "delh")
_PRED_SOURCE = {"ee": "ee_area", "ne": "ne_area", "te": "te_area", "cent": "cell_cent"}

# reference constants (StateVariables.f90:9-10) in the drivers' scaling:
# densities and viscosities relative to the liquid's
PRED_PARAMS = dict(dt=1.0e-3, itt=1, time_order=1, space_order=1,
                   rho_a_ref=1.205 / 998.3, rho_w_ref=1.0, mu_a_ref=1.506e-5 / 1.002e-3,
                   mu_w_ref=1.0, rey=150.0, fr=0.5, gx=0.0, gy=0.0, gz=-9.80665, g=9.80665)


def pred_cells(geo, interp):
    """The predictor's view of the U, V, W cells: three dicts keyed by
    PRED_CELL_FIELDS, from the geometry update's four cell dicts (P, U, V,
    W) and the three interpolation dicts (U, V, W)."""
    out = []
    for c, x in zip(geo[1:], interp):
        d = {}
        for n in PRED_CELL_FIELDS:
            src = _PRED_SOURCE.get(n, n)
            d[n] = c[src] if src in c else x[src]
        out.append(d)
    return out


def predictor_fields(size, seed=0, time_order=1, itt=1, walls=None):
    """The non-geometry inputs of one predictor step on a grid of ``size``:
    TVar and TVar_n velocities (U(-1,1), halo arrays), FluxDivOld (I,J,K,3),
    the 18 velocity BC flags (BCu, BCv, BCw; W,E,S,N,B,T; ``walls`` or
    seeded) and face values, and the parameters."""
    I, J, K = size
    hs = (I + 2, J + 2, K + 2)
    vel = [uniform(seed, 80 + q, -1.0, 1.0, hs) for q in range(3)]
    vel_n = [uniform(seed, 83 + q, -1.0, 1.0, hs) for q in range(3)]
    fdo = uniform(seed, 86, -1.0e-3, 1.0e-3, (I, J, K, 3))
    rng = np.random.default_rng(seed + 17)
    flags = (np.asarray(walls, dtype=np.int32) if walls is not None else
             rng.integers(0, 2, 18).astype(np.int32))
    dims = [(J, K), (J, K), (I, K), (I, K), (I, J), (I, J)]
    bcv = [np.asfortranarray(rng.uniform(-0.5, 0.5, dims[q % 6])) for q in range(18)]
    params = dict(PRED_PARAMS, itt=itt, time_order=time_order)
    return dict(vel=vel, vel_n=vel_n, flux_div_old=fdo, bc_flag=flags, bcv=bcv, params=params)
