"""The projection at the workload sizes (128^3 to 512x256x256), where the CPU
oracle is too slow to run beside it, through properties that do not depend on
the size:

* the corrected field is divergence-free: max |div u| <= 1e-10;
* the FP32-preconditioned solve and the all-FP64 solve give the same
  pressure (relative 1e-8) and velocity;
* projecting the projected field again changes nothing: the second pressure
  is a constant (zero with the pinned gauge) and u is returned unchanged;
* an adjoint pass on the largest case converges and satisfies the
  dot-product identity <J v, w> = <v, J^T w> of the linear map u* -> u.
"""

import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

CASES = [(1, (128, 128, 128)), (2, (256, 256, 256)), (3, (128, 128, 128)), (4, (512, 256, 256))]


def face_masks(hs, torch):
    """1 on the faces a predictor may carry: interior faces, not the
    wall-normal ones (u(I), v(J), w(K) are walls; u(0) etc. lie outside)."""
    K2, J2, I2 = hs
    ms = []
    for d in range(3):
        m = torch.zeros(hs, dtype=torch.float64, device="cuda")
        m[1:-1, 1:-1, 1:-1] = 1.0
        if d == 0:
            m[:, :, I2 - 2] = 0.0
        elif d == 1:
            m[:, J2 - 2, :] = 0.0
        else:
            m[K2 - 2, :, :] = 0.0
        ms.append(m)
    return ms


def _dev(c):
    import torch
    from simcoflow3d_amd.projection import to_torch_halo
    return {n: to_torch_halo(a) for n, a in H.fields(c).items()}, torch


def _project(P, c, dev, torch):
    hs = (c.K + 2, c.J + 2, c.I + 2)
    out = {n: torch.zeros(hs, dtype=torch.float64, device="cuda") for n in ("p", "u", "v", "w")}
    st = P.project_dev(dev, c.bc_flag, c.dt, out)
    torch.cuda.synchronize()
    return out, st


@pytest.mark.parametrize("cfg,size", CASES)
def test_fullsize_properties(gpu_available, cfg, size):
    c = H.case(cfg, size=size)
    print(f"[fullsize] cfg={cfg} size={size}", flush=True)
    dev, torch = _dev(c)
    res = {}
    for prec in (1, 0):
        P = H.projector(c, precision=prec)
        out, st = _project(P, c, dev, torch)
        assert st.converged == 1 and st.final_div <= 1e-10, (prec, st.as_dict())
        res[prec] = (out, st, P)
    (om, sm, Pm), (o64, s64, P64) = res[1], res[0]
    pm, p64 = om["p"], o64["p"]
    rel = ((pm - p64).abs().max() / p64.abs().max()).item()
    assert rel <= 1e-8, rel
    umax = max(dev[n].abs().max().item() for n in ("u_star", "v_star", "w_star"))
    for n in ("u", "v", "w"):
        assert (om[n] - o64[n]).abs().max().item() <= 1e-8 * umax, n
    # idempotence: the corrected field as the new predictor
    dev2 = dict(dev)
    for n, m in (("u_star", "u"), ("v_star", "v"), ("w_star", "w")):
        f = dev[n].clone()
        f[1:-1, 1:-1, 1:-1] = om[m][1:-1, 1:-1, 1:-1]
        dev2[n] = f
    o2, s2 = _project(Pm, c, dev2, torch)
    assert s2.final_div <= 1e-10
    assert o2["p"][1:-1, 1:-1, 1:-1].abs().max().item() <= 1e-8 * pm.abs().max().item()
    for n in ("u", "v", "w"):
        d = (o2[n][1:-1, 1:-1, 1:-1] - om[n][1:-1, 1:-1, 1:-1]).abs().max().item()
        assert d <= 1e-8 * umax, n
    Pm.close()
    P64.close()


@pytest.mark.parametrize("cfg,size,dist", [(4, (512, 256, 256), "normal"),
                                           (4, (512, 256, 256), "uniform"),
                                           (2, (256, 256, 256), "uniform")])
def test_fullsize_adjoint_dot_product(gpu_available, cfg, size, dist):
    """One mixed forward + adjoint pass at the workload sizes; <J v, w> =
    <v, J^T w> for the linear map u* -> u with fixed geometry (beta
    parametrisation).  Cotangents w either N(0,1) or U(0,1) (a source with a
    large mean flux) on the interior faces; the adjoint solve must converge
    for both."""
    import torch
    from simcoflow3d_amd import projection as PR
    c = H.case(cfg, size=size)
    print(f"[fullsize-adjoint] cfg={cfg} size={size} dist={dist}", flush=True)
    f = H.fields(c)
    geo = {n: PR.to_torch_halo(a) for n, a in f.items() if n not in ("u_star", "v_star", "w_star")}
    E = PR.DeviceProjection(c.shape, (c.dxp, c.dyp, c.dzp, c.dxu, c.dyv, c.dzw), geo, c.bc_flag,
                            c.dt, precision=1, rho_l=c.rho_l, rho_g=c.rho_g, rho_ref=c.rho_l)
    us = [PR.to_torch_halo(f[n]) for n in ("u_star", "v_star", "w_star")]
    beta = E.gfm_beta(*us)
    gen = torch.Generator(device="cuda").manual_seed(7)
    hs = us[0].shape
    masks = face_masks(hs, torch)
    v = [m * torch.randn(hs, generator=gen, device="cuda", dtype=torch.float64) for m in masks]
    draw = torch.randn if dist == "normal" else torch.rand
    w = [m * draw(hs, generator=gen, device="cuda", dtype=torch.float64) for m in masks]
    out, st = E.forward(*v, beta=beta)
    assert st.converged == 1
    res, ast = E.adjoint(*w)
    assert ast.converged == 1, ast.as_dict()
    lhs = sum((out[n] * wq).sum() for n, wq in zip(("u", "v", "w"), w)).item()
    rhs = sum((vq * res[n]).sum() for vq, n in zip(v, ("gu_star", "gv_star", "gw_star"))).item()
    # the adjoint solve stops at 1e-10 of its source; on the stiff rho 1000:1
    # operator that leaves a relative error up to ~1e-8 in J^T w
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("size", [(128, 128, 128), (256, 256, 256)])
def test_dirichlet_outlet_reaches_tolerance(gpu_available, size):
    """The reference drivers' outlet (pressure Dirichlet on the east face,
    Main.f90:92) with an inflow u*(0) = 1 fixes the pressure level at a few
    hundred: rounding of that level in one FP64 value per cell limits the
    divergence of the corrected field near 3.5e-10 at 256^3.  The engine
    refines the pressure as x + xd and reaches max |div u| <= 1e-10; both
    precisions give the same pressure."""
    import numpy as np
    c = H.case(2, size=size)
    I, J, K = c.shape
    c.bc_flag[:] = (1, 0, 1, 1, 1, 1)
    c.fields["u_star"][0, 1:-1, 1:-1] = 1.0
    print(f"[outlet] size={size}", flush=True)
    dev, torch = _dev(c)
    bcv = [None, torch.zeros((K, J), dtype=torch.float64, device="cuda"), None, None, None, None]
    hs = (K + 2, J + 2, I + 2)
    ps = []
    for prec in (0, 1):
        P = H.projector(c, precision=prec)
        out = {n: torch.zeros(hs, dtype=torch.float64, device="cuda") for n in ("p", "u", "v", "w")}
        st = P.project_dev(dev, c.bc_flag, c.dt, out, bc_val=bcv)
        torch.cuda.synchronize()
        P.close()
        assert st.converged == 1 and st.final_div <= 1e-10, st.as_dict()
        assert st.singular == 0
        ps.append(out["p"].cpu().numpy())
    assert np.max(np.abs(ps[0] - ps[1])) <= 1e-8 * np.max(np.abs(ps[0]))
