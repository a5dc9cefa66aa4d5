"""GPU parity of interpolateLvS (sp_interpolate_lvs; SolidBody.f90:646-742)
against the oracle on the same inputs, through host arrays and device
tensors.  Bar: phi, nx, ny, nz bit-identical; vof bit-identical except where
Volume_Fraction_Calc's x**3.d0 (a pow call) rounds differently on the GPU,
within 1e-12 (vof is a fraction in [0, 1])."""

import numpy as np
import pytest

from oracle import orc
from tests.test_oracle_lvs import body_mesh, rotation

pytestmark = pytest.mark.gpu

VOF_TOL = 1e-12


def stretched_grid(shape, seed, stagger=None):
    """Cell centres and widths of a tensor grid with random spacings; with
    stagger = axis the centres are the faces of that axis (a U/V/W grid)."""
    rng = np.random.default_rng(seed)
    d = [0.6 + 0.8 * rng.random(n) for n in shape]
    d = [q / q.sum() * 1.2 for q in d]
    e = [np.concatenate([[0.0], np.cumsum(q)]) - 0.1 for q in d]
    c = [0.5 * (q[1:] + q[:-1]) for q in e]
    if stagger is not None:
        c[stagger] = e[stagger][1:]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    DX, DY, DZ = np.meshgrid(*d, indexing="ij")
    f = np.asfortranarray
    return dict(x=f(X), y=f(Y), z=f(Z), dx=f(DX), dy=f(DY), dz=f(DZ))


def ellipsoid(X, Y, Z):
    return np.sqrt((X / 0.9) ** 2 + (Y / 1.1) ** 2 + (Z / 0.7) ** 2) * 0.35 - 0.28


CASES = [((24, 20, 28), (0.3, -0.2, 0.9), None, 40),
         ((32, 32, 32), (np.pi / 6, np.pi / 4, 0.0), 0, 48),
         ((20, 36, 24), (0.0, 0.0, 0.0), 2, 32),
         ((64, 48, 40), (1.1, 0.4, -2.0), 1, 64)]


@pytest.mark.parametrize("shape,angles,stagger,nb", CASES)
def test_interpolate_lvs_matches_oracle(gpu_available, shape, angles, stagger, nb):
    import torch
    from simcoflow3d_amd.engine import Projector
    print(f"[lvs] {shape} {angles} stagger={stagger}", flush=True)
    g = stretched_grid(shape, sum(shape), stagger)
    body = body_mesh(nb, 1.0 / nb, ellipsoid)
    body.update(cm=np.array([0.52, 0.47, 0.55]), rot=rotation(angles))
    rng = np.random.default_rng(7)
    hs = tuple(s + 2 for s in shape)
    n0 = [np.asfortranarray(rng.standard_normal(hs)) for _ in range(3)]
    ephi, evof, enx, eny, enz = orc.interpolate_lvs(body, g, *n0)
    band = np.abs(ephi[1:-1, 1:-1, 1:-1]) <= 0.2
    assert band.sum() > 200 and (ephi == 1.0e4).sum() > 0

    P = Projector(shape)
    out = dict(g)
    for n, a in zip(("phi", "vof", "nx", "ny", "nz"),
                    (np.zeros(hs), np.zeros(hs), *[a.copy(order="F") for a in n0])):
        out[n] = np.asfortranarray(a)
    P.interpolate_lvs(body, out)
    # device tensors (numpy (I, J, K) Fortran order <-> torch (K, J, I))
    def dev(a):
        return torch.from_numpy(np.asfortranarray(a).T.copy()).cuda()
    dbody = dict(body)
    for n in ("x", "y", "z", "phi"):
        dbody[n] = dev(body[n])
    dgrid = {n: dev(g[n]) for n in g}
    dgrid.update(phi=dev(np.zeros(hs)), vof=dev(np.zeros(hs)), nx=dev(n0[0]), ny=dev(n0[1]),
                 nz=dev(n0[2]))
    P.interpolate_lvs(dbody, dgrid, device=True)
    torch.cuda.synchronize()
    P.close()
    for got in (out, {n: dgrid[n].cpu().numpy().T for n in ("phi", "vof", "nx", "ny", "nz")}):
        assert np.array_equal(got["phi"], ephi)
        assert np.array_equal(got["nx"], enx)
        assert np.array_equal(got["ny"], eny)
        assert np.array_equal(got["nz"], enz)
        d = np.abs(got["vof"] - evof)
        assert d.max() <= VOF_TOL, float(d.max())
        print(f"  vof bit-identical in {100 * np.mean(d == 0):.3f}% of cells", flush=True)


def test_interpolate_lvs_rejects_bad_body(gpu_available):
    from simcoflow3d_amd.engine import Projector, SpError
    shape = (8, 8, 8)
    g = stretched_grid(shape, 1)
    hs = tuple(s + 2 for s in shape)
    out = dict(g)
    for n in ("phi", "vof", "nx", "ny", "nz"):
        out[n] = np.zeros(hs, order="F")
    body = body_mesh(8, 0.1, ellipsoid)
    body.update(cm=np.zeros(3), rot=np.eye(3), dx=0.0)
    P = Projector(shape)
    with pytest.raises(SpError, match="spacings"):
        P.interpolate_lvs(body, out)
    P.close()
