"""Shared helpers for the parity tests: engine inputs from a generated case
and the oracle's reference path on the same inputs."""

import numpy as np

from simcoflow3d_amd import configs

FIELD_NAMES = ("u_star", "v_star", "w_star", "dp_u", "dp_v", "dp_w", "ee_area", "ne_area",
               "te_area", "vof", "vof_old", "vof_l", "phi_l", "cell_type_p", "cell_type_u",
               "cell_type_v", "cell_type_w", "moex_u", "moex_v", "moex_w")

SIZES = {0: (24, 24, 24), 1: (24, 24, 32), 2: (32, 32, 32), 3: (20, 20, 20), 4: (48, 24, 24)}


def case(cfg, size=None, seed=0):
    return configs.generate(cfg, size=size or SIZES[cfg], seed=seed)


def fields(c):
    return {n: c.fields[n] for n in FIELD_NAMES}


def projector(c, **kw):
    from simcoflow3d_amd.engine import Projector
    params = dict(rho_l=c.rho_l, rho_g=c.rho_g, rho_ref=c.rho_l, pin_i=1, pin_j=1, pin_k=1)
    params.update(kw)
    P = Projector(c.shape, **params)
    P.set_grid_1d(c.dxp, c.dyp, c.dzp, c.dxu, c.dyv, c.dzw)
    return P


def rel_inf(a, b):
    d = np.max(np.abs(a - b))
    s = max(np.max(np.abs(b)), 1e-300)
    return d / s
