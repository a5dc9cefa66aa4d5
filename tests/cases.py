"""Hand-built inputs for the oracle pinning tests (independent of the
synthetic generator): uniform grids, controlled phases, apertures and flags."""

import dataclasses

import numpy as np


@dataclasses.dataclass
class Case:
    I: int
    J: int
    K: int
    dt: float
    rho_l: float
    rho_g: float
    dxp: np.ndarray
    dyp: np.ndarray
    dzp: np.ndarray
    dxu: np.ndarray
    dyv: np.ndarray
    dzw: np.ndarray
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


def uniform(I, J, K, L=(1.0, 1.0, 1.0), dt=1e-3, rho_l=1000.0, rho_g=1.0, phase="wet"):
    """All-fluid box, apertures 1 everywhere, Neumann walls, u* = 0."""
    hs = (I + 2, J + 2, K + 2)
    dx, dy, dz = L[0] / I, L[1] / J, L[2] / K
    f = {}
    for n in ("u_star", "v_star", "w_star", "vof_l", "phi_l"):
        f[n] = np.zeros(hs, order="F")
    for n in ("dp_u", "dp_v", "dp_w"):
        f[n] = np.full(hs, dt, order="F")
    for n in ("ee_area", "ne_area", "te_area", "vof", "vof_old"):
        f[n] = np.ones(hs, order="F")
    if phase == "wet":
        f["vof_l"][:] = 1.0
        f["phi_l"][:] = -1.0
    else:
        f["vof_l"][:] = 0.0
        f["phi_l"][:] = 1.0
    for n in ("cell_type_p", "cell_type_u", "cell_type_v", "cell_type_w", "moex_u", "moex_v",
              "moex_w"):
        f[n] = np.zeros((I, J, K), dtype=np.int32, order="F")
    return Case(I, J, K, dt, rho_l, rho_g, np.full(I, dx), np.full(J, dy), np.full(K, dz),
                np.full(I, dx), np.full(J, dy), np.full(K, dz), f, np.ones(6, dtype=np.int32))


def zcentres(c):
    return (np.arange(c.K + 2) - 0.5) * c.dzp[0]


def layered(I, J, K, zi, **kw):
    """Flat interface at z = zi: liquid below (phi_l = z - zi < 0)."""
    c = uniform(I, J, K, **kw)
    z = zcentres(c)
    phi = (z - zi)[None, None, :] * np.ones(c.halo_shape)
    c.fields["phi_l"][:] = phi
    c.fields["vof_l"][:] = (phi < 0).astype(float)
    return c
