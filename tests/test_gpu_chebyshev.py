"""The Chebyshev-Jacobi smoother (SP_FLAG_CHEBYSHEV) inside the same V-cycle
and flexible CG: the same solution as the oracle on every configuration
(both precisions, pre/post degrees 1 to 4), and an iteration count within a
few of the Gauss-Seidel V-cycle's."""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd.engine import SP_FLAG_CHEBYSHEV
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,size", [(0, None), (1, None), (2, None), (3, (20, 24, 16)),
                                      (4, None), (2, (33, 17, 40))])
@pytest.mark.parametrize("prec", (0, 1))
def test_chebyshev_matches_oracle(gpu_available, cfg, size, prec):
    c = H.case(cfg, size=size)
    print(f"[cheb] cfg={cfg} size={c.shape} prec={prec}", flush=True)
    ref = orc.project(c, pin=(1, 1, 1))
    P = H.projector(c, precision=prec, flags=SP_FLAG_CHEBYSHEV)
    out, st = P.project(H.fields(c), c.bc_flag, c.dt, po_coef=True)
    P.close()
    assert st.converged == 1 and st.final_div <= 1e-10, st.as_dict()
    assert np.array_equal(out["po_coef"], ref.pocoef)
    assert H.rel_inf(out["p"], ref.Pp) <= 1e-8
    umax = max(np.max(np.abs(c.u_star)), np.max(np.abs(c.v_star)), np.max(np.abs(c.w_star)))
    for n, r in (("u", ref.u), ("v", ref.v), ("w", ref.w)):
        d = out[n][1:-1, 1:-1, 1:-1] - r[1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(d)) <= 1e-8 * umax, n
    # the Gauss-Seidel V-cycle on the same case, for the iteration count
    P = H.projector(c, precision=prec)
    _, sg = P.project(H.fields(c), c.bc_flag, c.dt)
    P.close()
    print(f"  iterations: Chebyshev {st.iterations}, Gauss-Seidel {sg.iterations}", flush=True)
    assert st.iterations <= 2 * sg.iterations + 4


@pytest.mark.parametrize("deg", (1, 3, 4))
def test_chebyshev_degrees(gpu_available, deg):
    """Odd degrees end the post-smoother in the second buffer and copy back;
    every degree reaches the oracle's solution."""
    c = H.case(2, size=(40, 36, 44))
    ref = orc.project(c, pin=(1, 1, 1))
    P = H.projector(c, precision=1, flags=SP_FLAG_CHEBYSHEV, mg_pre=deg, mg_post=deg)
    out, st = P.project(H.fields(c), c.bc_flag, c.dt)
    P.close()
    print(f"[cheb] degree {deg}: {st.iterations} iterations", flush=True)
    assert st.converged == 1 and st.final_div <= 1e-10
    assert H.rel_inf(out["p"], ref.Pp) <= 1e-8
