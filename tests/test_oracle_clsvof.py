"""Pins the oracle's Clsvof_Scheme (Clsvof.f90:355-798 with its sweeps,
Interface_Reconstruct, Find_Distance, BoundaryConditionLvsVof and
Redistance_Levelset) with what the geometry fixes:

* Find_Distance inverts Volume_Fraction_Calc (the plane it returns cuts the
  given volume fraction);
* a liquid at rest keeps its volume fractions exactly, its level set
  becomes the signed distance to the reconstructed plane within four cells
  of the interface cells and +-5 sqrt(3) dx beyond; the boundary values and
  ghosts follow BoundaryConditionLvsVof's forms;
* a plane translated by a uniform velocity along an axis, or a plane at 45
  degrees in a coordinate plane translated within it, is carried exactly:
  volume fractions are the cut volumes of the moved plane and the level set
  its signed distance (away from the inflow boundaries, whose first-order
  level-set fluxes are not exact);
* a uniform fraction stays uniform in a divergence-free straining flow
  (this fixes the first sweep's division and the later sweeps' correction
  terms, which a uniform velocity leaves idle);
* a sphere carried there and back keeps its volume to round-off and comes
  back to its start within a few per cent.
"""

import math

import numpy as np
import pytest

from oracle import orc
from tests import clsvof_cases as C


def test_find_distance_inverts_volume_fraction():
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(3000):
        d = rng.uniform(0.2, 2.0, 3)
        n = rng.normal(size=3)
        if rng.random() < 0.2:
            n[rng.integers(3)] = 0.0
        n /= np.linalg.norm(n)
        f = rng.uniform(0.0, 1.0)
        s = orc.find_distance(*d, *n, f)
        v = orc.volume_fraction(*d, *n, s) / d.prod()
        worst = max(worst, abs(v - f))
    assert worst < 1e-11


def _run(shape, n, t, U, dt, **kw):
    g = C.grid(shape, seed=kw.pop("seed", 0), stretch=kw.pop("stretch", None))
    vof, phi, nn = C.plane_liquid(g, n, t)
    st = C.state(g, vof, phi, nn)
    inp = C.inputs(g, C.uniform_velocity(shape, U), dt, **kw)
    return g, vof, phi, nn, st, inp, orc.clsvof_scheme(inp, st)


@pytest.mark.parametrize("n,t,exact", [((0, 0, 1), 0.43, True), ((1, 0, 0), 0.61, True),
                                       ((0.6, 0.0, 0.8), 0.7, False),
                                       ((0.0, -0.8, 0.6), -0.1, False)])
def test_rest_state(n, t, exact):
    """At rest.  For a tilted plane the later sub-steps take their normals from
    the redistanced level set, whose band values past the first neighbours
    are not exact distances, so there only the volume fractions, the signs
    and the normals' directions are fixed."""
    shape = (16, 14, 18)
    lvs_flag = (1, 1, 0, 1, 0, 1)
    lvs_const = (0.0, 0.0, 0.25, 0.5, -0.3, 0.7)
    vof_flag = (0, 1, 1, 0, 1, 1)
    vof_const = (0.4, 0.1, 0.0, 0.9, 0.2, 0.0)
    solid_n = (0.36, 0.48, 0.8)
    g, vof, phi, nn, st, inp, out = _run(shape, n, t, (0, 0, 0), 0.01, lvs_flag=lvs_flag,
                                         lvs_const=lvs_const, vof_flag=vof_flag,
                                         vof_const=vof_const, solid_n=solid_n)
    v = out["vof_l"][1:-1, 1:-1, 1:-1]
    assert np.array_equal(v, vof)
    p = out["phi_l"][1:-1, 1:-1, 1:-1]
    cut = (vof > 0) & (vof < 1)
    assert cut.sum() > 50
    # the band: within four cells of an interface cell
    band = np.zeros(shape, bool)
    for a, b, c in np.argwhere(cut):
        band[max(a - 4, 0):a + 5, max(b - 4, 0):b + 5, max(c - 4, 0):c + 5] = True
    # exact distance where the nearest point of the plane lies in the cut
    # cells' reach: at least the interface cells and their face neighbours
    near = np.zeros(shape, bool)
    for a, b, c in np.argwhere(cut):
        near[max(a - 1, 0):a + 2, max(b - 1, 0):b + 2, max(c - 1, 0):c + 2] = True
    if exact:
        np.testing.assert_allclose(p[near], phi[near], rtol=0, atol=1e-14)
    else:
        np.testing.assert_allclose(p[near], phi[near], rtol=0, atol=0.15 * g["dx"].max())
    far = ~band
    assert far.sum() > 0
    np.testing.assert_allclose(np.abs(p[far]), 5 * g["dx"][far] * math.sqrt(3.0), rtol=1e-15)
    assert np.all(np.sign(p[far]) == np.sign(0.5 - vof[far]))
    # band values: signed by the phase, below the far value
    assert np.all(np.sign(p[band]) == np.sign(0.5 - vof[band]))
    assert np.all(np.abs(p[band]) <= 5 * g["dx"][band] * math.sqrt(3.0))
    # interface normals
    for a, key in enumerate(("nx_l", "ny_l", "nz_l")):
        np.testing.assert_allclose(out[key][1:-1, 1:-1, 1:-1][cut], nn[a],
                                   atol=1e-14 if exact else 0.05)
    # ghosts: vofL copied, phiL from the solid normal of the boundary cell
    # (tilted planes: the last sub-steps still move phi by ~1e-7)
    tol = 1e-15 if exact else 1e-5
    hv, hp = out["vof_l"], out["phi_l"]
    assert np.array_equal(hv[0, 1:-1, 1:-1], hv[1, 1:-1, 1:-1])
    assert np.array_equal(hv[1:-1, 1:-1, -1], hv[1:-1, 1:-1, -2])
    np.testing.assert_allclose(hp[0, 1:-1, 1:-1], hp[1, 1:-1, 1:-1] - solid_n[0] * g["dx"][0],
                               rtol=0, atol=tol)
    np.testing.assert_allclose(hp[1:-1, -1, 1:-1], hp[1:-1, -2, 1:-1] - solid_n[1] * g["dy"][:, -1],
                               rtol=0, atol=tol)
    np.testing.assert_allclose(hp[1:-1, 1:-1, -1], hp[1:-1, 1:-1, -2] + solid_n[2] * g["dz"][:, :, -1],
                               rtol=0, atol=tol)
    # face values of the last boundary call (the level set at rest does not
    # change between that call and the end): West/East from the liquid normal
    lv = out["lvs_var"]
    tol = 1e-14 if exact else 1e-5
    nxl = out["nx_l"][1:-1, 1:-1, 1:-1]  # the same in every sub-step at rest
    np.testing.assert_allclose(lv[0], p[0] - nxl[0] * g["dx"][0] / 2, atol=tol)
    np.testing.assert_allclose(lv[1], p[-1] + nxl[-1] * g["dx"][-1] / 2, atol=tol)
    assert np.all(lv[2] == 0.25) and np.all(lv[4] == -0.3)
    np.testing.assert_allclose(lv[3], p[:, -1] + 0.5 * g["dy"][:, -1] / 2, atol=tol)
    np.testing.assert_allclose(lv[5], p[:, :, -1] + 0.7 * g["dz"][:, :, -1] / 2, atol=tol)
    vv = out["vof_var"]
    assert np.all(vv[0] == 0.4) and np.all(vv[3] == 0.9)
    np.testing.assert_allclose(vv[1], v[-1] + 0.1 * g["dx"][-1] / 2, atol=1e-15)
    np.testing.assert_allclose(vv[4], v[:, :, 0] - 0.2 * g["dz"][:, :, 0] / 2, atol=1e-15)


@pytest.mark.parametrize("n,U", [((1, 0, 0), (0.3, 0, 0)), ((0, 1, 0), (0, -0.25, 0)),
                                 ((0, 0, 1), (0, 0, 0.3)), ((-1, 0, 0), (-0.2, 0, 0))])
def test_axis_translation_is_exact(n, U):
    shape = (16, 16, 16)
    t = 0.43 * sum(n)
    g, vof, phi, nn, st, inp, out = _run(shape, n, t, U, 0.3)
    vof2, phi2, _ = C.plane_liquid(g, nn, t + nn @ np.asarray(U) * 0.3)
    v = out["vof_l"][1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(v, vof2, rtol=0, atol=5e-14)
    assert not np.array_equal(vof2, vof)
    if n[2] == 0:
        # (the z sweep leaves residues of order 1e-14 in the downwind cell,
        # which the reconstruction then treats as interface cells)
        band = np.abs(phi2) < 3 * g["dx"].max()
        np.testing.assert_allclose(out["phi_l"][1:-1, 1:-1, 1:-1][band], phi2[band], atol=1e-14)


@pytest.mark.parametrize("n,U,phi_exact", [((1, 1, 0), (0.3, 0.1, 0), True),
                                           ((1, -1, 0), (-0.1, 0.25, 0), True),
                                           ((1, 0, 1), (-0.2, 0, 0.1), True),
                                           ((-1, 0, 1), (0.2, 0, 0.1), True),
                                           ((0, 1, 1), (0, 0.1, 0.3), False),
                                           ((0, 1, -1), (0, -0.2, -0.15), False)])
def test_diagonal_plane_translation_is_exact_inside(n, U, phi_exact):
    """A plane at 45 degrees in a coordinate plane, moved by a uniform
    velocity: every sweep's flux is the exact cut volume, so the volume
    fractions are those of the moved plane (to the Newton tolerance of
    Find_Distance's cubic branch) away from the boundaries.  (Offset from the
    cell centres: a 45-degree plane through a centre with f just below 0.5
    sends Find_Distance's Newton iteration onto a triple root, where it
    returns inf/NaN in the reference's arithmetic too.)  The level set is the
    exact distance too except where the z sweep leaves 1e-14 residues in
    empty cells, which the reconstruction treats as interface cells."""
    shape = (32, 32, 32)
    nn = np.asarray(n, float) / np.linalg.norm(n)
    t = nn @ np.array([0.5, 0.5, 0.5]) + 0.011
    g, vof, phi, nn, st, inp, out = _run(shape, n, t, U, 0.15)
    vof2, phi2, _ = C.plane_liquid(g, nn, t + nn @ np.asarray(U) * 0.15)
    inner = np.zeros(shape, bool)
    inner[8:-8, 8:-8, 8:-8] = True
    cut = (vof2 > 0) & (vof2 < 1) & inner
    assert cut.sum() > 100
    assert np.abs(vof2 - vof)[inner].max() > 0.1
    v = out["vof_l"][1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(v[inner], vof2[inner], rtol=0, atol=1e-10)
    if phi_exact:
        band = (np.abs(phi2) < 2.5 * g["dx"].max()) & inner
        np.testing.assert_allclose(out["phi_l"][1:-1, 1:-1, 1:-1][band], phi2[band], atol=1e-12)


@pytest.mark.parametrize("f,a", [(0.3, 0.5), (0.7, -0.5), (0.45, 0.8)])
def test_uniform_fraction_in_straining_flow(f, a):
    """Every cell cut by a horizontal plane at the same height (the level set
    the redistancing leaves is then constant, whose zero gradient the
    reconstruction turns into the normal (0, 0, 1) again), in the
    divergence-free straining flow (a x, -a y, 0): the fluxes through the x
    and y faces are f times the face volume flux, so the first sweep's
    division by (1 - dt du/dx) and the later sweeps' dt/dx T1 du terms must
    cancel the flux divergence exactly -- the fraction stays f everywhere (a
    sign slip in either moves it by ~1e-3)."""
    normal = 2
    shape = (16, 16, 16)
    g = C.grid(shape)
    I, J, K = shape
    vof = np.full(shape, f)
    phi = np.asfortranarray((g["x"], g["y"], g["z"])[normal] - 0.5)
    n = np.zeros(3)
    n[normal] = 1.0
    st = C.state(g, vof, phi, n)
    st["vof_l"][:] = f
    p, q = [a for a in range(3) if a != normal]
    vel = [np.zeros((I + 2, J + 2, K + 2)) for _ in range(3)]
    e = g["edges"]
    for axis, sgn in ((p, a), (q, -a)):
        shape_b = [1, 1, 1]
        shape_b[axis] = shape[axis] + 1
        idx = [slice(None)] * 3
        idx[axis] = slice(0, shape[axis] + 1)
        vel[axis][tuple(idx)] = sgn * e[axis].reshape(shape_b)
    vel = [np.asfortranarray(a) for a in vel]
    out = orc.clsvof_scheme(C.inputs(g, vel, 0.2, solid_n=tuple(n)), st)
    np.testing.assert_allclose(out["vof_l"][1:-1, 1:-1, 1:-1], f, rtol=0, atol=1e-15)


def test_sphere_translated_there_and_back():
    """A sphere carried by a uniform velocity for five steps and back: the
    liquid volume is kept to round-off (the boundaries see no liquid) and
    the shape comes back to within a few per cent (L1 over the volume)."""
    shape = (24, 24, 24)
    g = C.grid(shape)
    h = g["dx"][0, 0, 0]
    centre, radius = np.array([0.45, 0.5, 0.5]), 0.2
    r = np.sqrt((g["x"] - centre[0]) ** 2 + (g["y"] - centre[1]) ** 2 + (g["z"] - centre[2]) ** 2)
    # cell volume fractions of the sphere by 4^3 sub-samples
    sub = (np.arange(4) + 0.5) / 4 - 0.5
    vof = np.zeros(shape)
    for a in sub:
        for b in sub:
            for c in sub:
                rr = np.sqrt((g["x"] + a * h - centre[0]) ** 2 + (g["y"] + b * h - centre[1]) ** 2
                             + (g["z"] + c * h - centre[2]) ** 2)
                vof += (rr < radius) / 64.0
    st = C.state(g, vof, r - radius, (0.0, 0.0, 1.0))
    st["phi_l"][1:-1, 1:-1, 1:-1] = r - radius
    vol0 = vof.sum()
    for sgn in (1.0, -1.0):
        for _ in range(5):
            vel = C.uniform_velocity(shape, (0.2 * sgn, 0.1 * sgn, 0.05 * sgn))
            st = orc.clsvof_scheme(C.inputs(g, vel, 0.1), st)
            v = st["vof_l"][1:-1, 1:-1, 1:-1]
            assert v.min() >= 0.0 and v.max() <= 1.0
            assert abs(v.sum() / vol0 - 1.0) < 1e-12
    err = np.abs(v - vof).sum() / vol0
    assert err < 0.03, err
