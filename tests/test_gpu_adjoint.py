"""The engine's adjoint of the projection on the GPU.

* against the oracle's adjoint (same PoCoef and forward pressure), to 1e-7
  relative in max norm;
* through torch.autograd (simcoflow3d_amd.projection.project) against
  central finite differences of L = sum W (u^2 + v^2 + w^2) in random
  directions of u*, v*, w* and of the face coefficients beta, relative
  error <= 1e-6.  The map u* -> u is linear, so the u* differences are exact
  for any step.  L is not quadratic in beta: central differences at a
  relative step h carry an O(h^2) error (measured ~1e-6 at h = 1e-4 where the
  directional derivative is small) and below h ~ 1e-5 the forward's
  convergence noise takes over, so the beta check uses the Richardson
  combination (4 D(h) - D(2h)) / 3 at h = 1e-4, which cancels the h^2 term;
* the differentiation configuration with its own check (128^3, w ~ U(0,1),
  eps = 1e-6, 64 directions each of u* and beta);
* a loss on the pressure too (the pressure cotangent), against the oracle in
  both gauges and with Dirichlet faces, and through autograd."""

import numpy as np
import pytest

from oracle import orc
from tests import helpers as H
from tests.test_gpu_fullsize import face_masks

pytestmark = pytest.mark.gpu


def interior(a):
    z = np.zeros_like(a)
    z[1:-1, 1:-1, 1:-1] = a[1:-1, 1:-1, 1:-1]
    return z


@pytest.mark.parametrize("cfg,prec,size", [(3, 0, (16, 16, 20)), (3, 1, (16, 16, 20)),
                                           (0, 0, (16, 16, 20)), (4, 1, (32, 16, 16)),
                                           (3, 0, (64, 64, 64)), (3, 0, (128, 128, 128)),
                                           (3, 1, (128, 128, 128))])
def test_adjoint_matches_oracle(gpu_available, cfg, prec, size):
    c = H.case(cfg, size=size)
    print(f"[adjoint] cfg={cfg} prec={prec} shape={c.shape}", flush=True)
    P = H.projector(c, precision=prec)
    out, st = P.project(H.fields(c), c.bc_flag, c.dt, po_coef=True)
    assert st.converged == 1
    rng = np.random.default_rng(5)
    W = [interior(rng.uniform(0.5, 1.5, c.halo_shape)) for _ in range(3)]
    g = [2 * Wq * interior(out[n]) for Wq, n in zip(W, ("u", "v", "w"))]
    res, ast = P.adjoint(*g, beta_grad=True)
    P.close()
    assert ast.converged == 1
    # the oracle's solve runs a fixed 120 iterations (the FP64 floor up to
    # 128^3: 200 iterations change its gradients by 3e-13 relative there);
    # 128^3 is the differentiation configuration at its full size
    o = orc.opts(max_iter=120, div_tol=0.0, coarse_scale=0.5)
    ref = orc.adjoint(c, out["po_coef"], out["p"], *g, o=o)
    for a, b in (("gu_star", "gus"), ("gv_star", "gvs"), ("gw_star", "gws"), ("gk_e", "gkx"),
                 ("gk_n", "gky"), ("gk_t", "gkz")):
        assert H.rel_inf(res[a], ref[b]) <= 1e-7, (a, H.rel_inf(res[a], ref[b]))
    # dL/dbeta = dL/dK * Dp/delta on the E,N,T faces
    I, J, K = c.shape
    dpe = c.dp_u[1:-1, 1:-1, 1:-1] / c.dxu[:, None, None]
    assert np.allclose(res["gbeta_e"], res["gk_e"] * dpe, rtol=1e-14, atol=0)


PCOT = [((1, 1, 1), None, 0), ((1, 1, 1), None, 1), ((0, 0, 0), None, 0), ((0, 0, 0), None, 1),
        ((4, 3, 5), None, 0), ((1, 1, 1), (0, 1, 1, 1, 1, 1), 0), ((1, 1, 1), (1, 1, 1, 1, 0, 0), 1)]


@pytest.mark.parametrize("pin,flags,prec", PCOT)
def test_pressure_cotangent_matches_oracle(gpu_available, pin, flags, prec):
    """A loss on the pressure too (L = sum W |u|^2 + sum Wp Pp^2 over every
    halo entry, ghosts included): the engine's gradients against the
    oracle's, in the pin gauge, the zero-mean gauge and with Dirichlet faces."""
    c = H.case(3, size=(16, 16, 20))
    bcv = None
    if flags is not None:
        c.bc_flag[:] = flags
        I, J, K = c.shape
        dims = [(J, K), (J, K), (I, K), (I, K), (I, J), (I, J)]
        rng0 = np.random.default_rng(4)
        bcv = [np.asfortranarray(rng0.uniform(-2.0, 2.0, dims[f])) if flags[f] == 0 else None
               for f in range(6)]
    print(f"[adjoint-p] pin={pin} flags={flags} prec={prec}", flush=True)
    P = H.projector(c, precision=prec, pin_i=pin[0], pin_j=pin[1], pin_k=pin[2])
    out, st = P.project(H.fields(c), c.bc_flag, c.dt, bc_val=bcv, po_coef=True)
    assert st.converged == 1
    rng = np.random.default_rng(8)
    W = [interior(rng.uniform(0, 1, c.halo_shape)) for _ in range(3)]
    Wp = rng.uniform(0, 1, c.halo_shape)
    g = [2 * Wq * interior(out[n]) for Wq, n in zip(W, ("u", "v", "w"))]
    gp = 2 * Wp * out["p"]
    res, ast = P.adjoint(*g, beta_grad=True, gp=gp)
    P.close()
    assert ast.converged == 1
    o = orc.opts(max_iter=120, div_tol=0.0, coarse_scale=0.5)
    ref = orc.adjoint(c, out["po_coef"], out["p"], *g, o=o, gp=gp, pin=pin)
    for a, b in (("gu_star", "gus"), ("gv_star", "gvs"), ("gw_star", "gws"), ("gk_e", "gkx"),
                 ("gk_n", "gky"), ("gk_t", "gkz")):
        assert H.rel_inf(res[a], ref[b]) <= 1e-7, (a, H.rel_inf(res[a], ref[b]))
    # the pressure term changes the gradients: without it they differ
    P0 = H.projector(c, precision=prec, pin_i=pin[0], pin_j=pin[1], pin_k=pin[2])
    P0.project(H.fields(c), c.bc_flag, c.dt, bc_val=bcv)
    r0, _ = P0.adjoint(*g)
    P0.close()
    assert H.rel_inf(r0["gu_star"], ref["gus"]) > 1e-4


def _device_setup(cfg, size, prec, div_tol):
    import torch
    from simcoflow3d_amd import projection as PR
    c = H.case(cfg, size=size)
    f = H.fields(c)
    geo = {}
    for n, a in f.items():
        if n in ("u_star", "v_star", "w_star"):
            continue
        geo[n] = PR.to_torch_halo(a) if a.ndim == 3 and a.shape == c.halo_shape \
            else PR.to_torch_cell(a)
    grid = (c.dxp, c.dyp, c.dzp, c.dxu, c.dyv, c.dzw)
    E = PR.DeviceProjection(c.shape, grid, geo, c.bc_flag, c.dt, precision=prec,
                            rho_l=c.rho_l, rho_g=c.rho_g, rho_ref=c.rho_l, div_tol=div_tol,
                            max_iter=400)
    us = [PR.to_torch_halo(f[n]) for n in ("u_star", "v_star", "w_star")]
    beta = E.gfm_beta(*us)
    return torch, PR, c, E, us, beta


@pytest.mark.parametrize("cfg,size,prec", [(3, (128, 128, 128), 0), (3, (64, 64, 64), 1),
                                           (1, (48, 48, 64), 0)])
def test_autograd_matches_finite_differences(gpu_available, cfg, size, prec):
    torch, PR, c, E, us, beta = _device_setup(cfg, size, prec, 1e-11)
    print(f"[autograd-fd] cfg={cfg} size={size} prec={prec}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(3)
    hs = us[0].shape
    mask = torch.zeros(hs, dtype=torch.float64, device="cuda")
    mask[1:-1, 1:-1, 1:-1] = 1.0
    W = [mask * (0.5 + torch.rand(hs, generator=gen, device="cuda", dtype=torch.float64))
         for _ in range(3)]

    def loss(u3, b3):
        u, v, w, _ = PR.project(E, *u3, b3)
        return sum((Wq * a * a).sum() for Wq, a in zip(W, (u, v, w)))

    leaves = [t.clone().requires_grad_(True) for t in us + beta]
    L = loss(leaves[:3], leaves[3:])
    L.backward()
    grads = [t.grad.detach() for t in leaves]
    assert E.last_stats.converged == 1
    worst = 0.0
    with torch.no_grad():
        for trial in range(32):
            # u*, v*, w* direction (interior faces)
            d = [m * torch.randn(hs, generator=gen, device="cuda", dtype=torch.float64)
                 for m in face_masks(hs, torch)]
            h = 1e-2
            lp = loss([a + h * b for a, b in zip(us, d)], beta)
            lm = loss([a - h * b for a, b in zip(us, d)], beta)
            fd = ((lp - lm) / (2 * h)).item()
            an = sum((a * b).sum() for a, b in zip(d, grads[:3])).item()
            worst = max(worst, abs(fd - an) / abs(an))
            assert abs(fd - an) <= 1e-6 * abs(an), ("ustar", trial, fd, an)
            # beta direction (relative perturbation of every face coefficient)
            db = [torch.randn(b.shape, generator=gen, device="cuda", dtype=torch.float64) * b
                  for b in beta]

            def central(h):
                lp = loss(us, [a + h * b for a, b in zip(beta, db)])
                lm = loss(us, [a - h * b for a, b in zip(beta, db)])
                return ((lp - lm) / (2 * h)).item()

            d1, d2 = central(1e-4), central(2e-4)
            fd = (4.0 * d1 - d2) / 3.0
            an = sum((a * b).sum() for a, b in zip(db, grads[3:])).item()
            worst = max(worst, abs(fd - an) / abs(an))
            assert abs(d1 - an) <= 1e-4 * abs(an), ("beta central", trial, d1, an)
            assert abs(fd - an) <= 1e-6 * abs(an), ("beta", trial, fd, an)
    print(f"[autograd-fd] worst relative error {worst:.3e}", flush=True)


def test_backward_after_interleaved_forward(gpu_available):
    """The engine keeps one forward pressure: a backward whose forward was
    followed by another forward recomputes its own first."""
    torch, PR, c, E, us, beta = _device_setup(0, (24, 24, 24), 0, 1e-11)
    print("[autograd-interleaved]", flush=True)
    a = [t.clone().requires_grad_(True) for t in us]
    u, v, w, _ = PR.project(E, *a, beta)
    (u * u).sum().backward()
    ref = a[0].grad.clone()
    b = [t.clone().requires_grad_(True) for t in us]
    u, v, w, _ = PR.project(E, *b, beta)
    other = [2.0 * t for t in us]
    PR.project(E, *other, beta)  # a second forward in between
    (u * u).sum().backward()
    assert torch.allclose(b[0].grad, ref, rtol=1e-9, atol=1e-12 * ref.abs().max().item())


@pytest.mark.parametrize("prec", (0, 1))
def test_autograd_pressure_loss_finite_differences(gpu_available, prec):
    """A loss on u and on the pressure (zero-mean gauge) through autograd:
    central differences in u* (exact for any step, the maps are linear) and
    the Richardson combination in beta, relative error <= 1e-6."""
    torch, PR, c, E, us, beta = _device_setup(3, (48, 40, 56), prec, 1e-11)
    print(f"[autograd-p] prec={prec}", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(12)
    hs = us[0].shape
    mask = torch.zeros(hs, dtype=torch.float64, device="cuda")
    mask[1:-1, 1:-1, 1:-1] = 1.0
    W = [mask * torch.rand(hs, generator=gen, device="cuda", dtype=torch.float64)
         for _ in range(3)]
    Wp = torch.rand(hs, generator=gen, device="cuda", dtype=torch.float64)

    def loss(u3, b3):
        u, v, w, p = PR.project(E, *u3, b3)
        return sum((Wq * a * a).sum() for Wq, a in zip(W, (u, v, w))) + (Wp * p * p).sum()

    leaves = [t.clone().requires_grad_(True) for t in us + beta]
    loss(leaves[:3], leaves[3:]).backward()
    grads = [t.grad.detach() for t in leaves]
    with torch.no_grad():
        for trial in range(6):
            d = [m * torch.randn(hs, generator=gen, device="cuda", dtype=torch.float64)
                 for m in face_masks(hs, torch)]
            h = 1e-2
            fd = ((loss([a + h * b for a, b in zip(us, d)], beta) -
                   loss([a - h * b for a, b in zip(us, d)], beta)) / (2 * h)).item()
            an = sum((a * b).sum() for a, b in zip(d, grads[:3])).item()
            assert abs(fd - an) <= 1e-6 * abs(an), ("ustar", trial, fd, an)
            db = [torch.randn(b.shape, generator=gen, device="cuda", dtype=torch.float64) * b
                  for b in beta]

            def central(h):
                lp = loss(us, [a + h * b for a, b in zip(beta, db)])
                lm = loss(us, [a - h * b for a, b in zip(beta, db)])
                return ((lp - lm) / (2 * h)).item()

            fd = (4.0 * central(1e-4) - central(2e-4)) / 3.0
            an = sum((a * b).sum() for a, b in zip(db, grads[3:])).item()
            assert abs(fd - an) <= 1e-6 * abs(an), ("beta", trial, fd, an)


def test_differentiation_configuration_as_stated(gpu_available):
    """The differentiation configuration at its size and with its stated
    check: 128^3, sphere solid, rho ratio 1000, FP64 forward and adjoint,
    L = sum w |u|^2 with w ~ U(0,1), central differences with eps = 1e-6 in
    64 seeded random directions of u* and 64 of the face coefficients beta,
    relative error <= 1e-6 in every direction.

    The difference L(x + eps d) - L(x - eps d) is summed term by term,
    sum w (u+ - u-)(u+ + u-), which equals it in exact arithmetic.  Summing
    the two losses separately and subtracting cancels about 8 of FP64's 16
    digits over these 6.3 M terms: the worst of 64 directions then reaches
    8e-6 (u*) and 1e-5 (beta) from the summation's rounding alone, whatever
    the forward's tolerance, against 1.5e-7 and 1e-8 term by term."""
    torch, PR, c, E, us, beta = _device_setup(3, (128, 128, 128), 0, 1e-10)
    print("[fd-as-stated] 128^3 fp64 eps=1e-6 64+64 directions", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(2024)
    hs = us[0].shape
    mask = torch.zeros(hs, dtype=torch.float64, device="cuda")
    mask[1:-1, 1:-1, 1:-1] = 1.0
    W = [mask * torch.rand(hs, generator=gen, device="cuda", dtype=torch.float64)
         for _ in range(3)]

    def velocities(u3, b3):
        u, v, w, _ = PR.project(E, *u3, b3)
        return u, v, w

    leaves = [t.clone().requires_grad_(True) for t in us + beta]
    sum((Wq * a * a).sum() for Wq, a in zip(W, velocities(leaves[:3], leaves[3:]))).backward()
    assert E.last_stats.converged == 1
    grads = [t.grad.detach() for t in leaves]
    eps = 1e-6

    def central(plus, minus):
        return (sum((Wq * (a - b) * (a + b)).sum() for Wq, a, b in zip(W, plus, minus)) /
                (2 * eps)).item()

    worst = [0.0, 0.0]
    with torch.no_grad():
        for trial in range(64):
            d = [m * torch.randn(hs, generator=gen, device="cuda", dtype=torch.float64)
                 for m in face_masks(hs, torch)]
            fd = central(velocities([a + eps * b for a, b in zip(us, d)], beta),
                         velocities([a - eps * b for a, b in zip(us, d)], beta))
            an = sum((a * b).sum() for a, b in zip(d, grads[:3])).item()
            worst[0] = max(worst[0], abs(fd - an) / abs(an))
            assert abs(fd - an) <= 1e-6 * abs(an), ("ustar", trial, fd, an)
        for trial in range(64):
            db = [torch.randn(b.shape, generator=gen, device="cuda", dtype=torch.float64) * b
                  for b in beta]
            fd = central(velocities(us, [a + eps * b for a, b in zip(beta, db)]),
                         velocities(us, [a - eps * b for a, b in zip(beta, db)]))
            an = sum((a * b).sum() for a, b in zip(db, grads[3:])).item()
            worst[1] = max(worst[1], abs(fd - an) / abs(an))
            assert abs(fd - an) <= 1e-6 * abs(an), ("beta", trial, fd, an)
    print(f"[fd-as-stated] worst relative error u* {worst[0]:.3e}, beta {worst[1]:.3e}",
          flush=True)
