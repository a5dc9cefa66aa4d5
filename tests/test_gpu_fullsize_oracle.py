"""Every benchmark configuration at its full size against the oracle, in
both precisions: 64^3 (C0), 128^3 stretched (C1), 256^3 (C2), 128^3 (C3)
and 512x256x256 (C4).

The oracle assembles the reference's own rows (ProjectionP.f90:301-576,
GFM coefficients :123-265) and solves them with its FP64 multigrid PCG to
max |r|/V <= 1e-12 (all five reach it within 18 iterations), then gathers
the pressure, sets its ghosts (:578-643) and corrects the velocity
(ComputePUVW.f90:181-232).  The GFM coefficients come from the serial
restatement up to 128^3 and from the parallel one (bit-identical to it on
every configuration) above.  Bars as for the small cases: PoCoef
bit-identical, pressure within 1e-8 relative, corrected velocity within
1e-8 of max |u|, max |div u| <= 1e-10.
"""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs
from tests import helpers as H

pytestmark = pytest.mark.gpu

CONFIGS = [0, 1, 2, 3, 4]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_config_at_full_size_matches_oracle(gpu_available, cfg):
    dims = configs.config_dims(cfg)
    print(f"[fullsize-oracle] cfg={cfg} size={dims}: generating", flush=True)
    c = configs.generate(cfg)
    ref = orc.project(c, pin=(1, 1, 1), tol=1e-12, gfm_fast=c.I * c.J * c.K > 128 ** 3)
    print(f"[fullsize-oracle] cfg={cfg} oracle: {ref.stats.iterations} iterations, "
          f"max|r|/V={ref.stats.final_div:.2e}", flush=True)
    assert ref.stats.converged == 1
    for prec in (0, 1):
        print(f"[fullsize-oracle] cfg={cfg} prec={prec}", flush=True)
        out, st = H.run_engine(c, prec, device=True)
        assert st.converged == 1 and st.final_div <= 1e-10, (prec, st.as_dict())
        H.assert_matches_oracle(c, out, st, ref)
        print(f"[fullsize-oracle] cfg={cfg} prec={prec}: p rel err "
              f"{H.rel_inf(out['p'], ref.Pp):.2e}, {st.iterations} iterations", flush=True)
        del out
