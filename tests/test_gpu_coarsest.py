"""Multigrid hierarchies whose coarsest level does not fit the LDS-resident
tail kernel (seven arrays per level in 144 KB): a hierarchy cut short by
mg_levels, and a grid too thin to coarsen.  The coarsest level is then
smoothed in global memory (the tail's coarsest sweeps with the plain
red-black kernels); a coarsest level above the tail's 512-cell limit that
fits its LDS stays in the tail.  Bar: the parity bars of the oracle comparison
(PoCoef bit-identical, pressure and corrected velocity within 1e-8) and
convergence to max |div u| <= 1e-10, both precisions."""

import pytest

from oracle import orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", (0, 1))
@pytest.mark.parametrize("size,levels", [
    ((48, 40, 32), 2),  # coarsest 24x20x16: 7680 cells, 215 KB in FP32
    ((64, 3, 40), 0),   # J = 3 does not coarsen: one level of 7680 cells
    ((40, 40, 3), 0),   # one level of 4800 cells: the tail in FP32 (134 KB), not in FP64
])
def test_coarsest_level_outside_the_tail(gpu_available, size, levels, prec):
    print(f"[coarsest] size={size} mg_levels={levels} precision={prec}", flush=True)
    c = H.case(0, size=size)
    out, st = H.run_engine(c, prec, mg_levels=levels)
    assert st.converged == 1, st.as_dict()
    assert st.final_div <= 1e-10
    ref = orc.project(c, pin=(1, 1, 1))
    H.assert_matches_oracle(c, out, st, ref)
