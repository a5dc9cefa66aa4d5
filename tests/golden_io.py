"""Loader of the golden fixtures written by tests/golden/make_golden.py."""

import os

import numpy as np

from tests import cases

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("config0_8x8x8", "config1_8x8x12", "config4_16x8x8")


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    meta = z["meta"]
    I, J, K = (int(v) for v in meta[1:4])
    fields = {k[3:]: np.asfortranarray(z[k]) for k in z.files if k.startswith("in_")}
    g = {k[5:]: z[k] for k in z.files if k.startswith("grid_")}
    c = cases.Case(I, J, K, float(meta[4]), float(meta[5]), float(meta[6]), g["dxp"], g["dyp"],
                   g["dzp"], g["dxu"], g["dyv"], g["dzw"], fields, z["bc_flag"].astype(np.int32))
    exp = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return int(meta[0]), c, exp
