"""Pins the oracle's GFM face coefficients (Compute1DGFMCoefficient,
ProjectionP.f90:123-265) with closed forms, not with the formula itself:

* single phase: coef = Dp/delta (liquid) and rho_l/rho_g * Dp/delta (gas);
* flat interface between two cell centres: the coefficient is the harmonic
  (series) average Dp/delta / (theta rho_l' + (1-theta) rho_g') with theta
  the liquid share of the centre-to-centre segment computed from geometry;
* the two cells of a face give the same value (to rounding);
* the first/last plane copies the neighbouring interior plane (:256-264);
* a face next to a solid cell reuses the beta left by the previous cell of
  the i-j-k loop (:170-248);
* the parallel restatement used by the CPU baseline is bit-identical.
"""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs
from tests import cases


def coef(c):
    return orc.gfm_coef(c, rho_l=c.rho_l, rho_g=c.rho_g, rho_ref=c.rho_l)


def test_single_phase_liquid_and_gas():
    for phase, factor in (("wet", 1.0), ("dry", 1000.0)):
        c = cases.uniform(5, 4, 6, L=(1.0, 2.0, 3.0), phase=phase)
        pc = coef(c)
        for f, d in enumerate((c.dxu, c.dyv, c.dzw) * 2):
            ref = c.dt / d[0] * factor
            np.testing.assert_allclose(pc[..., f], ref, rtol=4e-16, atol=0)


def test_flat_interface_series_average():
    I, J, K = 3, 3, 8
    zi = 0.43  # between the centres of k=3 (z=0.3125) and k=4 (z=0.4375)
    c = cases.layered(I, J, K, zi)
    pc = coef(c)
    z = cases.zcentres(c)
    h = c.dzp[0]
    kP = int(np.floor(zi / h + 0.5))  # last liquid cell (1-based)
    theta = (zi - z[kP]) / (z[kP + 1] - z[kP])  # liquid share of the segment
    rl, rg = 1.0, c.rho_g / c.rho_l
    expect = c.dt / h / (theta * rl + (1 - theta) * rg)
    # top face of the liquid cell and bottom face of the gas cell above
    np.testing.assert_allclose(pc[:, :, kP - 1, 5], expect, rtol=1e-13)
    np.testing.assert_allclose(pc[:, :, kP, 2], expect, rtol=1e-13)
    # faces away from the interface: pure phases
    np.testing.assert_allclose(pc[:, :, 0, 5], c.dt / h, rtol=4e-16)
    np.testing.assert_allclose(pc[:, :, K - 2, 5], c.dt / h * c.rho_l / c.rho_g, rtol=4e-16)


def test_face_symmetry_and_boundary_copy():
    c = configs.generate(0, size=(12, 12, 12))
    pc = coef(c)
    ct = c.cell_type_p
    # interior x faces between two fluid cells: E of i equals W of i+1
    E, W = pc[:-1, :, :, 3], pc[1:, :, :, 0]
    fluid = (ct[:-1] == 0) & (ct[1:] == 0)
    inner = np.zeros_like(fluid)
    inner[1:-1] = True  # skip the copied first/last planes
    m = fluid & inner
    assert m.sum() > 500
    np.testing.assert_allclose(E[m], W[m], rtol=3e-16)
    # boundary copies
    np.testing.assert_array_equal(pc[0, :, :, 0], pc[1, :, :, 0])
    np.testing.assert_array_equal(pc[-1, :, :, 3], pc[-2, :, :, 3])
    np.testing.assert_array_equal(pc[:, 0, :, 1], pc[:, 1, :, 1])
    np.testing.assert_array_equal(pc[:, -1, :, 4], pc[:, -2, :, 4])
    np.testing.assert_array_equal(pc[:, :, 0, 2], pc[:, :, 1, 2])
    np.testing.assert_array_equal(pc[:, :, -1, 5], pc[:, :, -2, 5])


def test_stale_beta_reuses_previous_loop_cell():
    """Column (1,1,:) = [wet, dry, wet, solid]: the top face of the third cell
    points at a solid cell, so the reference reuses BetaW from the last cell
    of its loop that assigned one -- the wet-over-dry face of the first cell."""
    I, J, K = 1 + 1, 1 + 1, 4
    c = cases.uniform(I, J, K)
    f = c.fields
    # column (i=1, j=1) in halo indices [1,1,k]
    phi = np.array([-0.05, 0.02, -0.07, 0.3])
    for k in range(K):
        f["phi_l"][1, 1, k + 1] = phi[k]
        f["vof_l"][1, 1, k + 1] = 1.0 if phi[k] < 0 else 0.0
    f["vof"][1, 1, 4] = 0.0          # solid (vof <= epsi): unclassified
    # make every other cell gas so no later/earlier column assigns BetaW
    for i in range(I + 2):
        for j in range(J + 2):
            if (i, j) != (1, 1):
                f["phi_l"][i, j, :] = 1.0
                f["vof_l"][i, j, :] = 0.0
    pc = coef(c)
    T = pc[0, 0, :, 5]
    lam = abs(phi[0]) / (abs(phi[0]) + abs(phi[1]))
    betaP, betaM = 1.0, c.rho_l / c.rho_g
    expect0 = c.dt / c.dzw[0] * betaP * betaM / (lam * betaM + (1 - lam) * betaP)
    np.testing.assert_allclose(T[0], expect0, rtol=1e-14)
    assert T[2] == T[0]  # stale: the same BetaW
    assert T[3] == T[2]  # the last plane copies the plane below (:262-264)


@pytest.mark.parametrize("cfg,size", [(0, (16, 16, 16)), (1, (12, 14, 18)), (3, (10, 10, 10)),
                                      (4, (32, 16, 16)), (2, (16, 16, 16)), (1, None),
                                      (4, (128, 64, 64))])
def test_parallel_restatement_is_bit_identical(cfg, size):
    c = configs.generate(cfg, size=size)
    a = orc.gfm_coef(c)
    b = orc.gfm_coef_fast(c)
    assert np.array_equal(a, b)
