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
