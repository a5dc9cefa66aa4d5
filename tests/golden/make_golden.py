"""Writes the golden fixtures: small seeded inputs of three configurations and
the oracle's outputs on them (PoCoef, pressure with ghosts, corrected
velocity).  The oracle is pinned by closed forms and independent solvers
(tests/test_oracle_*.py); these fixtures freeze its outputs so the engine
can be checked against them without running the oracle.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import orc  # noqa: E402
from simcoflow3d_amd import configs  # noqa: E402

FIXTURES = {
    "config0_8x8x8": (0, (8, 8, 8)),
    "config1_8x8x12": (1, (8, 8, 12)),
    "config4_16x8x8": (4, (16, 8, 8)),
}


def make(name, cfg, size):
    c = configs.generate(cfg, size=size)
    ref = orc.project(c, pin=(1, 1, 1))
    out = {f"in_{k}": v for k, v in c.fields.items()}
    for k in ("dxp", "dyp", "dzp", "dxu", "dyv", "dzw"):
        out[f"grid_{k}"] = getattr(c, k)
    out["bc_flag"] = c.bc_flag
    out["meta"] = np.array([cfg, *size, c.dt, c.rho_l, c.rho_g])
    out["out_pocoef"] = ref.pocoef
    out["out_p"] = ref.Pp
    out["out_u"], out["out_v"], out["out_w"] = ref.u, ref.v, ref.w
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


if __name__ == "__main__":
    for name, (cfg, size) in FIXTURES.items():
        make(name, cfg, size)
        print("wrote", name)
