"""Predictor inputs built with the oracle's geometry update (GridPreprocess,
DefineMomentumExchangeCell, NewCellFace): the U/V/W cells a predictor step
reads, in the reference's own conventions for apertures, face and cell
centres, interpolation factors and distances."""

import dataclasses

import numpy as np

from oracle import orc
from simcoflow3d_amd import configs


@dataclasses.dataclass
class PredGrid:
    I: int
    J: int
    K: int
    dxp: np.ndarray
    dyp: np.ndarray
    dzp: np.ndarray
    dxu: np.ndarray
    dyv: np.ndarray
    dzw: np.ndarray

    @property
    def shape(self):
        return (self.I, self.J, self.K)


def grid_of(cells):
    """1-D spacings of the P, U, V, W grids from geometry_case's cell dicts."""
    P, U, V, W = cells
    return PredGrid(*P["dx"].shape, P["dx"][:, 0, 0].copy(), P["dy"][0, :, 0].copy(),
                    P["dz"][0, 0, :].copy(), U["dx"][:, 0, 0].copy(), V["dy"][0, :, 0].copy(),
                    W["dz"][0, 0, :].copy())


def geometry(size, shape="bodies", seed=0, stretch=False, fluid=False):
    """(grid, geo, interp): the four cell dicts after GridPreprocess and
    NewCellFace and the three U/V/W interpolation dicts.  ``fluid``: no
    solid anywhere (the level set far outside the domain)."""
    cells = configs.geometry_case(size, shape, seed=seed, stretch=stretch)
    if fluid:
        for c in cells:
            c["phi"] = np.full_like(c["phi"], 10.0)
            c["nx"] = np.ones_like(c["nx"])
            c["ny"] = np.zeros_like(c["ny"])
            c["nz"] = np.zeros_like(c["nz"])
            c["vof"] = np.ones_like(c["vof"])
    geo = orc.grid_preprocess(cells)
    zeros = [np.zeros(size, np.int32, order="F") for _ in range(3)]
    moex = orc.define_moex(*(g["cell_type"] for g in geo), *zeros)
    geo, interp = orc.new_cell_face(geo, [dict(moex=m) for m in moex])
    return grid_of(cells), geo, interp


def p_areas(geo):
    return [geo[0]["ee_area"], geo[0]["ne_area"], geo[0]["te_area"]]


def oracle_step(grid, geo_old, int_old, geo_cur, int_cur, f, pred=None):
    old = configs.pred_cells(geo_old, int_old)
    cur = configs.pred_cells(geo_cur, int_cur)
    return orc.predict(grid, p_areas(geo_old), old, cur, f["vel"], f["vel_n"], f["flux_div_old"],
                       f["bc_flag"], f["bcv"], f["params"], pred=pred)
