"""Pins the oracle's SpaceOrder 2 convective fluxes (ModifiedConvectiveFlux,
PredictUVW.f90:2253-2644) by closed forms, not by its formulas:

* the first step (itt = 1) takes the first-order fluxes whatever SpaceOrder
  says (:172-185), and later steps differ from them;
* uniform flow through an all-fluid box is kept, on uniform and stretched
  grids;
* 1-D self-advection on a uniform all-fluid grid, in each direction: the
  textbook central flux a^2 A with a the average of the two neighbours;
* a v field carried by a uniform u: U0 times the average of the two v
  values, with the BC face values at the walls.

The cut-face branches (P aperture <= 1e-5) follow the reference's
statements, including the carry of its local delh between iterations when
neither neighbour test holds; no closed form covers them here (parity
unpinned at the oracle), and the engine is checked against this
restatement bit for bit."""

import numpy as np
import pytest

from simcoflow3d_amd import configs
from tests import pred_cases as PC


def _step(size, vel, bcv=None, stretch=False, itt=2, space_order=2):
    grid, geo, interp = PC.geometry(size, fluid=True, stretch=stretch)
    f = configs.predictor_fields(size, seed=5, walls=[1] * 18, itt=itt)
    f["vel"] = [np.asfortranarray(a) for a in vel]
    f["bcv"] = bcv if bcv is not None else [np.zeros_like(a) for a in f["bcv"]]
    f["params"].update(gx=0.0, gy=0.0, gz=0.0, space_order=space_order)
    return PC.oracle_step(grid, geo, interp, geo, interp, f), f


def test_first_step_takes_first_order_fluxes():
    size = (10, 8, 9)
    grid, geo, interp = PC.geometry(size, shape="sphere", seed=6)
    res = {}
    for so in (1, 2):
        for itt in (1, 2):
            f = configs.predictor_fields(size, seed=8, itt=itt)
            f["params"]["space_order"] = so
            res[so, itt] = PC.oracle_step(grid, geo, interp, geo, interp, f)
    for a, b in zip(res[1, 1][0] + res[1, 1][1], res[2, 1][0] + res[2, 1][1]):
        assert np.array_equal(a, b)
    assert not np.array_equal(res[1, 2][2], res[2, 2][2])


@pytest.mark.parametrize("stretch", (False, True))
def test_uniform_flow_is_kept(stretch):
    size = (9, 7, 8)
    U0 = 0.7
    hs = (11, 9, 10)
    vel = [np.full(hs, U0), np.zeros(hs), np.zeros(hs)]
    f0 = configs.predictor_fields(size, seed=5, walls=[1] * 18)
    bcv = [np.full(a.shape, U0 if q in (0, 1) else 0.0, order="F")
           for q, a in enumerate(f0["bcv"])]
    ((u, v, w), dps, fdo), f = _step(size, vel, bcv, stretch=stretch)
    assert np.allclose(u[1:-1, 1:-1, 1:-1], U0, rtol=1e-13, atol=0)
    assert np.max(np.abs(v[1:-1, 1:-1, 1:-1])) <= 1e-14
    assert np.max(np.abs(w[1:-1, 1:-1, 1:-1])) <= 1e-14
    I, J, K = size
    for part in (fdo[:I - 1, :, :, 0], fdo[:, :J - 1, :, 1], fdo[:, :, :K - 1, 2]):
        assert np.max(np.abs(part)) <= 1e-13


@pytest.mark.parametrize("axis", (0, 1, 2))
def test_one_dimensional_central_self_advection(axis):
    """q = q(x_axis) of either sign, the other two velocities zero: the faces
    across the axis carry (q_{m-1} + q_m)^2 / 4 A, every face the same
    (interior faces, the last interior one and the wall ones alike)."""
    size = [8, 7, 6]
    size[axis] = 12
    size = tuple(size)
    n = size[axis]
    prof = np.random.default_rng(11 + axis).uniform(-1.0, 1.0, n + 2)  # index 0..n+1
    hs = tuple(s + 2 for s in size)
    shape = [1, 1, 1]
    shape[axis] = n + 2
    vel = [np.zeros(hs) for _ in range(3)]
    vel[axis] = np.broadcast_to(prof.reshape(shape), hs).copy()
    ((u, v, w), dps, fdo), f = _step(size, vel)
    h = [1.0 / s for s in size]
    A = np.prod([h[d] for d in range(3) if d != axis])
    F = (0.5 * (prof[:-1] + prof[1:])) ** 2 * A     # F[m]: face between index m and m+1
    div = F[1:n] - F[0:n - 1]                        # cells 1..n-1 of the component
    lim = [slice(None)] * 3
    lim[axis] = slice(0, n - 1)
    got = fdo[tuple(lim) + (axis,)]
    shape2 = [1, 1, 1]
    shape2[axis] = n - 1
    assert np.allclose(got, div.reshape(shape2), rtol=1e-12, atol=1e-15)


def test_one_dimensional_advection_of_v_by_uniform_u():
    """u = U0, v = v(x): the x faces of the V cells carry U0 (v_{i-1} + v_i)/2
    inside and U0 VarW / U0 VarE at the walls."""
    I, J, K = 11, 6, 4
    U0 = -0.8
    vx = np.random.default_rng(8).uniform(-1.0, 1.0, I + 2)
    hs = (I + 2, J + 2, K + 2)
    vel = [np.full(hs, U0), np.broadcast_to(vx[:, None, None], hs).copy(), np.zeros(hs)]
    f0 = configs.predictor_fields((I, J, K), seed=5, walls=[1] * 18)
    bcv = [np.zeros_like(a) for a in f0["bcv"]]
    bcv[6][:] = 0.31   # BCv%VarW
    bcv[7][:] = -0.27  # BCv%VarE
    ((u, v, w), dps, fdo), f = _step((I, J, K), vel, bcv)
    A = (1.0 / J) * (1.0 / K)
    Fv = np.empty(I + 1)              # faces x_{i-1/2}, i = 1..I+1
    Fv[0] = U0 * 0.31 * A
    Fv[1:I] = U0 * (0.5 * (vx[1:I] + vx[2:I + 1])) * A  # between V(m) and V(m+1)
    Fv[I] = U0 * -0.27 * A
    div_v = Fv[1:] - Fv[:-1]
    assert np.allclose(fdo[:, :J - 1, :, 1], div_v[:, None, None], rtol=1e-12, atol=1e-15)
