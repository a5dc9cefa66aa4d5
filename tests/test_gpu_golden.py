"""The engine against the committed golden fixtures (no oracle at run time).

PoCoef bit-identical; pressure within 1e-8 relative; corrected velocity
within 1e-8 of max|u*|; the reference-interface mirror
(poisson_equation_solver) writes the same outputs in place."""

import numpy as np
import pytest

from tests import golden_io
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_io.NAMES)
@pytest.mark.parametrize("prec", (0, 1))
def test_engine_matches_golden(gpu_available, name, prec):
    cfg, c, exp = golden_io.load(name)
    print(f"[golden] {name} prec={prec}", flush=True)
    P = H.projector(c, precision=prec)
    out, st = P.project(H.fields(c), c.bc_flag, c.dt, po_coef=True)
    P.close()
    assert st.converged == 1 and st.final_div <= 1e-10, st.as_dict()
    assert np.array_equal(out["po_coef"], exp["pocoef"])
    assert H.rel_inf(out["p"], exp["p"]) <= 1e-8
    umax = max(np.max(np.abs(c.u_star)), np.max(np.abs(c.v_star)), np.max(np.abs(c.w_star)))
    for n in ("u", "v", "w"):
        d = out[n][1:-1, 1:-1, 1:-1] - exp[n][1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(d)) <= 1e-8 * umax, n


@pytest.mark.parametrize("name", golden_io.NAMES)
def test_reference_interface_mirror(gpu_available, name):
    from simcoflow3d_amd import projection as PR
    cfg, c, exp = golden_io.load(name)
    print(f"[mirror] {name}", flush=True)
    f = c.fields
    I, J, K = c.shape
    ones = np.ones((I, J, K), order="F")
    PGrid = PR.Grid(c.dxp[:, None, None] * ones, c.dyp[None, :, None] * ones,
                    c.dzp[None, None, :] * ones)
    UGrid = PR.Grid(c.dxu[:, None, None] * ones, None, None)
    VGrid = PR.Grid(None, c.dyv[None, :, None] * ones, None)
    WGrid = PR.Grid(None, None, c.dzw[None, None, :] * ones)
    PCell = PR.Cell(f["cell_type_p"], f["vof"], f["vof_l"], f["phi_l"], f["ee_area"],
                    f["ne_area"], f["te_area"])
    UCell = PR.Cell(f["cell_type_u"], MoExCell=f["moex_u"])
    VCell = PR.Cell(f["cell_type_v"], MoExCell=f["moex_v"])
    WCell = PR.Cell(f["cell_type_w"], MoExCell=f["moex_w"])
    hs = c.halo_shape
    TVar = PR.Variables(*(np.zeros(hs, order="F") for _ in range(4)))
    TPred = PR.Predictor(f["u_star"], f["v_star"], f["w_star"])
    PoCoef = np.zeros((I, J, K, 6), order="F")
    Proj = PR.Projection(np.zeros(hs, order="F"))
    st = PR.poisson_equation_solver(
        PGrid, UGrid, VGrid, WGrid, PCell, UCell, VCell, WCell, f["vof_old"], TVar, TPred,
        PR.PoissonCoefficient(f["dp_u"]), PR.PoissonCoefficient(f["dp_v"]),
        PR.PoissonCoefficient(f["dp_w"]), PR.BCBase(c.bc_flag), PoCoef, Proj, c.dt,
        rho_l=c.rho_l, rho_g=c.rho_g, pin=(1, 1, 1))
    assert st.converged == 1
    assert np.array_equal(PoCoef, exp["pocoef"])
    assert H.rel_inf(Proj.Pp, exp["p"]) <= 1e-8
    assert np.array_equal(TVar.p[1:-1, 1:-1, 1:-1], Proj.Pp[1:-1, 1:-1, 1:-1])
    umax = max(np.max(np.abs(c.u_star)), np.max(np.abs(c.v_star)), np.max(np.abs(c.w_star)))
    for n in ("u", "v", "w"):
        d = getattr(TVar, n)[1:-1, 1:-1, 1:-1] - exp[n][1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(d)) <= 1e-8 * umax, n
