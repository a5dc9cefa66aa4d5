"""GPU parity of ComputeUVWLiquidField (sp_compute_uvw_liquid_field;
Clsvof.f90:174-353) against the oracle on the same inputs.  Bar: U/V/W
phiL, nxL, nyL, nzL bit-identical; vofL bit-identical except where the
reference's x**3.d0 (a pow call, Clsvof.f90:1516-1536) rounds differently
on the GPU than on the host, within 1e-12 absolute (vofL is a fraction in
[0, 1])."""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs
from simcoflow3d_amd.engine import LIQUID_FIELDS, SP_ERR_INVALID_ARG, Projector, SpError

pytestmark = pytest.mark.gpu

VOF_TOL = 1e-12


def liquid_case(size, shape, seed, stretch):
    """P-grid liquid level set and normal of a seeded shape, with vofL."""
    I, J, K = size
    dxp = np.full(I, 1.0 / I)
    dyp = np.full(J, 1.0 / J)
    dzp = np.full(K, 1.0 / K)
    if stretch:
        dzp = 1.02 ** np.arange(K)
        dzp /= dzp.sum()
    grids = configs.staggered_spacings(dxp, dyp, dzp)
    g = grids[0]
    phi, nx, ny, nz = configs.solid_level_set(shape, g["x"], g["y"], g["z"],
                                              np.random.default_rng(seed))
    rng = np.random.default_rng(seed + 1)
    P = dict(dx=g["dx"], dy=g["dy"], dz=g["dz"], phi_l=np.asfortranarray(phi),
             nx_l=np.asfortranarray(nx), ny_l=np.asfortranarray(ny), nz_l=np.asfortranarray(nz),
             vof_l=np.asfortranarray(rng.uniform(size=size)))
    return [P] + [dict(dx=q["dx"], dy=q["dy"], dz=q["dz"]) for q in grids[1:]]


def check(got, exp):
    for m, (o, e) in enumerate(zip(got, exp)):
        for n in ("phi_l", "nx_l", "ny_l", "nz_l"):
            assert np.array_equal(o[n], e[n], equal_nan=True), (m, n)
        d = np.abs(o["vof_l"] - e["vof_l"])
        assert np.all(d <= VOF_TOL), (m, float(d.max()))
        exact = np.mean(o["vof_l"] == e["vof_l"])
        print(f"  grid {'UVW'[m]}: vofL bit-identical in {100 * exact:.3f}% of cells, "
              f"max diff {d.max():.3e}", flush=True)


@pytest.mark.parametrize("shape", configs.GEOM_SHAPES)
@pytest.mark.parametrize("size,stretch", [((13, 11, 9), False), ((9, 7, 12), True),
                                          ((2, 3, 2), False), ((70, 5, 6), True)])
def test_uvw_liquid_field_matches_oracle(gpu_available, shape, size, stretch):
    print(f"[liquid] shape={shape} size={size} stretch={stretch}", flush=True)
    cells = liquid_case(size, shape, sum(size), stretch)
    exp = orc.uvw_liquid_field(cells)[1:]
    P = Projector(size)
    got = P.uvw_liquid_field(cells)
    P.close()
    check(got, exp)


def test_uvw_liquid_field_device_128_cubed(gpu_available):
    import torch
    size = (128, 128, 128)
    print("[liquid] 128^3 host and device", flush=True)
    cells = liquid_case(size, "bodies", 3, True)
    exp = orc.uvw_liquid_field(cells)[1:]
    P = Projector(size)
    check(P.uvw_liquid_field(cells), exp)
    tdev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda:0")
    dev = [{n: (tdev(c[n]) if n in c else torch.zeros(size[::-1], dtype=torch.float64,
                                                       device="cuda:0"))
            for n in LIQUID_FIELDS} for c in cells]
    P.uvw_liquid_field(dev, device=True)
    P.close()
    check([{n: d[n].cpu().numpy().T for n in LIQUID_FIELDS} for d in dev[1:]], exp)


def test_null_member_is_rejected(gpu_available):
    import ctypes
    from simcoflow3d_amd.engine import SpLiquidField
    P = Projector((4, 4, 4))
    z = SpLiquidField()
    rc = P.lib.sp_compute_uvw_liquid_field(P.h, *(ctypes.byref(z) for _ in range(4)), 0)
    assert rc == SP_ERR_INVALID_ARG
    with pytest.raises(SpError):
        P._check(rc, "sp_compute_uvw_liquid_field")
    P.close()
