"""Pins the oracle's interpolateLvS (SolidBody.f90:646-742) and
Normal_Vector_Irre (Clsvof.f90:2361-2456) with what the mathematics fixes:

* trilinear interpolation reproduces a body level set that is linear in the
  body's coordinates exactly, under any rotation and translation; on a
  smooth level set it equals scipy's RegularGridInterpolator at the cell
  centres taken into the body frame;
* for a plane level set the band normals are the rotated plane normal and
  the band vof is the plane-cut volume of each cell (inclusion-exclusion);
* cells outside the body mesh get phi = 1e4; vof stays 1 outside the band,
  inside the body too; normals outside the band keep their values;
* Normal_Vector_Irre takes the one-sided difference on the side where the
  gradient magnitude stays 1 and the central one elsewhere.
"""

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import orc
from tests.test_oracle_liquid import halfspace_volume


def rotation(angles):
    """rotMatLtoG of calcRotationMatrix (SolidBody.f90:589-613), written as
    the product Rz(psi) Ry(theta) Rx(phi) of elementary rotations."""
    a, b, c = angles
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


def body_mesh(n, h, phi_fn):
    """A uniform local mesh of n^3 cells of size h centred on the origin."""
    x = (np.arange(n) - (n - 1) / 2.0) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return dict(x=x, y=x.copy(), z=x.copy(), dx=h, dy=h, dz=h,
                phi=np.asfortranarray(phi_fn(X, Y, Z)))


def global_grid(shape, lo, hi):
    I, J, K = shape
    e = [np.linspace(lo[a], hi[a], s + 1) for a, s in enumerate(shape)]
    c = [0.5 * (q[1:] + q[:-1]) for q in e]
    d = [np.diff(q) for q in e]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    DX, DY, DZ = np.meshgrid(*d, indexing="ij")
    return dict(x=np.asfortranarray(X), y=np.asfortranarray(Y), z=np.asfortranarray(Z),
                dx=np.asfortranarray(DX), dy=np.asfortranarray(DY), dz=np.asfortranarray(DZ))


def to_local(g, rot, cm):
    p = np.stack([g["x"] - cm[0], g["y"] - cm[1], g["z"] - cm[2]], axis=-1)
    return p @ rot  # R^T (x - cm) for each point


def covered(body, pl):
    lo = np.array([body["x"][0], body["y"][0], body["z"][0]])
    hi = np.array([body["x"][-1], body["y"][-1], body["z"][-1]])
    return np.all((pl > lo) & (pl <= hi), axis=-1)


@pytest.mark.parametrize("angles", [(0.0, 0.0, 0.0), (0.3, -0.5, 1.1), (np.pi / 6, np.pi / 4, 0.2)])
def test_linear_level_set_is_reproduced(angles):
    nl = np.array([0.36, -0.48, 0.8])  # unit normal in the body frame
    body = body_mesh(16, 0.05, lambda X, Y, Z: nl[0] * X + nl[1] * Y + nl[2] * Z - 0.07)
    rot = rotation(angles)
    cm = np.array([0.52, 0.47, 0.55])
    body.update(cm=cm, rot=rot)
    g = global_grid((20, 18, 22), (0.0, 0.0, 0.0), (1.0, 0.9, 1.1))
    sentinel = np.full((22, 20, 24), 7.0, order="F")
    phi, vof, nx, ny, nz = orc.interpolate_lvs(body, g, sentinel, sentinel, sentinel)
    pl = to_local(g, rot, cm)
    inside = covered(body, pl)
    assert inside.sum() > 500 and (~inside).sum() > 500
    exact = pl @ nl - 0.07
    core = phi[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(core[inside], exact[inside], rtol=0, atol=1e-13)
    assert np.all(core[~inside] == 1.0e4)
    # ghosts are reset too
    assert np.all(phi[0] == 1.0e4) and np.all(vof[:, :, -1] == 1.0)
    # band: normals and vof
    ng = rot @ nl
    hmax = np.maximum(np.maximum(g["dx"], g["dy"]), g["dz"])
    band = np.abs(core) <= 4.0 * hmax
    I, J, K = g["dx"].shape
    interior = np.zeros(band.shape, bool)
    interior[1:-1, 1:-1, 1:-1] = True
    # the interior band cells whose whole stencil lies inside the body mesh
    ok = band & inside & interior
    for a, n in enumerate((nx, ny, nz)):
        cn = n[1:-1, 1:-1, 1:-1]
        assert np.all(cn[~band] == 7.0)
        full = ok.copy()
        for s in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            for sign in (-2, -1, 1, 2):
                full &= np.roll(inside, tuple(-sign * q for q in s), axis=(0, 1, 2))
        np.testing.assert_allclose(cn[full], ng[a], rtol=0, atol=1e-11)
    assert full.sum() > 100
    cv = vof[1:-1, 1:-1, 1:-1]
    assert np.all(cv[~band] == 1.0)
    idx = np.argwhere(full)
    for i, j, k in idx[:: max(1, len(idx) // 200)]:
        c = np.array([g["x"][i, j, k], g["y"][i, j, k], g["z"][i, j, k]])
        d = np.array([g["dx"][i, j, k], g["dy"][i, j, k], g["dz"][i, j, k]])
        # fluid {n.(R^T (x - cm)) >= 0.07} = {-ng.x <= -(0.07 + ng.cm)}
        vol = halfspace_volume(c - d / 2, c + d / 2, -ng, -(0.07 + ng @ cm))
        assert cv[i, j, k] == pytest.approx(vol / d.prod(), abs=1e-10)


def test_smooth_level_set_matches_library_interpolation():
    def sphere(X, Y, Z):
        return np.sqrt(X ** 2 + 1.3 * Y ** 2 + 0.8 * Z ** 2) - 0.3
    body = body_mesh(30, 0.04, sphere)
    rot = rotation((0.2, 0.4, -0.7))
    cm = np.array([0.5, 0.5, 0.5])
    body.update(cm=cm, rot=rot)
    g = global_grid((16, 16, 16), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    phi, vof, nx, ny, nz = orc.interpolate_lvs(body, g)
    pl = to_local(g, rot, cm)
    inside = covered(body, pl)
    f = RegularGridInterpolator((body["x"], body["y"], body["z"]), body["phi"])
    core = phi[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(core[inside], f(pl[inside]), rtol=0, atol=1e-13)
    # deep inside the body the vof is left at 1 (the reference sets it only
    # in the band)
    hmax = 1.0 / 16
    deep = inside & (core < -4.0 * hmax)
    if deep.any():
        assert np.all(vof[1:-1, 1:-1, 1:-1][deep] == 1.0)
    # normals in the band are unit vectors
    band = inside & (np.abs(core) <= 4.0 * hmax)
    norm = np.sqrt(nx ** 2 + ny ** 2 + nz ** 2)[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(norm[band], 1.0, rtol=1e-14)


def test_normal_vector_one_sided_at_a_kink():
    I, J, K = 12, 6, 6
    g = global_grid((I, J, K), (0.0, 0.0, 0.0), (1.2, 0.6, 0.6))
    x = g["x"][:, 0, 0]
    m = 6  # 1-based cell whose right neighbours see the steeper slope
    xm = x[m - 1]
    phi = np.zeros((I + 2, J + 2, K + 2), order="F")
    xx = np.concatenate([[x[0] - 0.1], x, [x[-1] + 0.1]])
    prof = np.where(xx <= xm + 1e-12, xx, xx + 2.0 * (xx - xm))
    phi[:] = prof[:, None, None]
    # cell m: left slope 1 (q = 0), right two-cell slope 3 (q = 2): backward
    n = orc.normal_vector_irre(g, phi, m, 3, 3)
    assert n[0] == (phi[m, 3, 3] - phi[m - 1, 3, 3]) / (x[m - 1] - x[m - 2])
    assert n[0] == pytest.approx(1.0, rel=1e-14) and n[1] == 0.0 and n[2] == 0.0
    # cell m+1: both sides depart from 1: central difference (slope 3)
    n = orc.normal_vector_irre(g, phi, m + 1, 3, 3)
    assert n[0] == (phi[m + 2, 3, 3] - phi[m, 3, 3]) / (x[m + 1] - x[m - 1])
    assert n[0] == pytest.approx(3.0, rel=1e-13)
    # cell m-1: left slope 1, right two-cell slope 2: backward
    n = orc.normal_vector_irre(g, phi, m - 1, 3, 3)
    assert n[0] == (phi[m - 1, 3, 3] - phi[m - 2, 3, 3]) / (x[m - 2] - x[m - 3])
    # mirrored kink, steep on the left: forward at cell m
    prof = np.where(xx >= xm - 1e-12, xx, xx + 2.0 * (xx - xm))
    phi[:] = prof[:, None, None]
    n = orc.normal_vector_irre(g, phi, m, 3, 3)
    assert n[0] == (phi[m + 1, 3, 3] - phi[m, 3, 3]) / (x[m] - x[m - 1])
    assert n[0] == pytest.approx(1.0, rel=1e-14)
