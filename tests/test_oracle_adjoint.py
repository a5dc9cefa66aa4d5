"""Pins the oracle's adjoint of the projection.

Forward u = M(u* - K G p), A p = G^T a u*, A = G^T a K G; loss
L = sum w u^2 (+ sum wp Pp^2 with the pressure field, ghosts included).
Checks that do not restate the adjoint:
* the maps u* -> u and u* -> Pp are linear: <J v, w> = <v, J^T w>
  (dot-product tests, the pressure in the pin gauge, the zero-mean gauge and
  with a Dirichlet face, where no gauge applies);
* central finite differences of L in random directions of u* and of the
  face coefficients K = PoCoef(:,4..6).
"""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs


def forward(c, pc, ustar, pin=(1, 1, 1)):
    """Projection with given face coefficients; returns (u, v, w, Pp).  A
    Dirichlet pressure face holds the value 0, so the map stays linear."""
    for n, a in zip(("u_star", "v_star", "w_star"), ustar):
        c.fields[n] = np.asfortranarray(a)
    planes = [(c.J, c.K), (c.J, c.K), (c.I, c.K), (c.I, c.K), (c.I, c.J), (c.I, c.J)]
    bc_val = [np.zeros(planes[f], order="F") if c.bc_flag[f] == 0 else None for f in range(6)]
    b = orc.rhs(c, pc, bc_val)
    op = orc.sym_operator(c, pc)
    d, _ = orc.sym_diag(c, *op)
    o = orc.opts(max_iter=1000, div_tol=1e-300)
    x, st, sing = orc._pcg_to_floor(c, op, b, o, None)
    x = orc.gauge(c, d, sing, x, pin)
    Pp = orc.pressure_field(c, x)
    u, v, w = orc.correct(c, pc, Pp)
    return u, v, w, Pp


def interior(a):
    z = np.zeros_like(a)
    z[1:-1, 1:-1, 1:-1] = a[1:-1, 1:-1, 1:-1]
    return z


@pytest.fixture(scope="module")
def setup():
    c = configs.generate(3, size=(8, 8, 10))
    pc = orc.gfm_coef(c)
    rng = np.random.default_rng(11)
    W = [interior(rng.uniform(0, 1, c.halo_shape)) for _ in range(3)]
    us = [c.u_star.copy(), c.v_star.copy(), c.w_star.copy()]
    return c, pc, W, us, rng


def loss(c, pc, us, W):
    u, v, w, _ = forward(c, pc, us)
    return sum(np.sum(Wq * interior(a) ** 2) for Wq, a in zip(W, (u, v, w)))


def test_dot_product_linear_map(setup):
    c, pc, W, us, rng = setup
    # J v for v = random u*-perturbation on the faces the forward reads
    vdir = [interior(rng.standard_normal(c.halo_shape)) for _ in range(3)]
    Jv = forward(c, pc, vdir)[:3]
    wdir = [interior(rng.standard_normal(c.halo_shape)) for _ in range(3)]
    u, v, w, Pp = forward(c, pc, vdir)
    g = orc.adjoint(c, pc, Pp, *wdir)
    lhs = sum(np.sum(interior(a) * b) for a, b in zip(Jv, wdir))
    rhs = sum(np.sum(a * b) for a, b in zip(vdir, (g["gus"], g["gvs"], g["gws"])))
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_gradient_wrt_ustar_central_differences(setup):
    c, pc, W, us, rng = setup
    u, v, w, Pp = forward(c, pc, us)
    g = orc.adjoint(c, pc, Pp, *(2 * Wq * interior(a) for Wq, a in zip(W, (u, v, w))))
    for _ in range(4):
        d = [interior(rng.standard_normal(c.halo_shape)) for _ in range(3)]
        eps = 1e-6
        lp = loss(c, pc, [a + eps * b for a, b in zip(us, d)], W)
        lm = loss(c, pc, [a - eps * b for a, b in zip(us, d)], W)
        fd = (lp - lm) / (2 * eps)
        an = sum(np.sum(a * b) for a, b in zip(d, (g["gus"], g["gvs"], g["gws"])))
        assert abs(fd - an) <= 1e-6 * abs(an)


def test_gradient_wrt_face_coefficients_central_differences(setup):
    c, pc, W, us, rng = setup
    u, v, w, Pp = forward(c, pc, us)
    g = orc.adjoint(c, pc, Pp, *(2 * Wq * interior(a) for Wq, a in zip(W, (u, v, w))))
    gk = np.stack([g["gkx"], g["gky"], g["gkz"]], axis=-1)
    for _ in range(4):
        d = rng.standard_normal(pc[..., 3:].shape) * pc[..., 3:]
        eps = 1e-6
        pp, pm = pc.copy(order="F"), pc.copy(order="F")
        pp[..., 3:] += eps * d
        pm[..., 3:] -= eps * d
        fd = (loss(c, pp, us, W) - loss(c, pm, us, W)) / (2 * eps)
        an = np.sum(gk * d)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize("pin,dirichlet", [((1, 1, 1), False), ((0, 0, 0), False),
                                           ((5, 3, 7), False), ((1, 1, 1), True)])
def test_dot_product_pressure(pin, dirichlet):
    """<Jp v, wp> = <v, Jp^T wp> for Jp: u* -> Pp (every halo entry, so the
    face ghosts count); the velocity cotangents are zero."""
    c = configs.generate(3, size=(8, 8, 10))
    if dirichlet:
        c.bc_flag[1] = 0  # east face: Dirichlet pressure, the system is not singular
    pc = orc.gfm_coef(c)
    rng = np.random.default_rng(17)
    vdir = [interior(rng.standard_normal(c.halo_shape)) for _ in range(3)]
    Jv = forward(c, pc, vdir, pin)[3]
    wp = np.asfortranarray(rng.standard_normal(c.halo_shape))
    z = np.zeros(c.halo_shape, order="F")
    g = orc.adjoint(c, pc, Jv, z, z, z, gp=wp, pin=pin)
    lhs = np.sum(Jv * wp)
    rhs = sum(np.sum(a * b) for a, b in zip(vdir, (g["gus"], g["gvs"], g["gws"])))
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("pin", [(1, 1, 1), (0, 0, 0)])
def test_pressure_loss_central_differences(setup, pin):
    """L = sum W u^2 + sum Wp Pp^2: central differences in u* and in K."""
    c, pc, W, us, rng = setup
    Wp = rng.uniform(0, 1, c.halo_shape)

    def loss2(pcq, u3):
        u, v, w, Pp = forward(c, pcq, u3, pin)
        return (sum(np.sum(Wq * interior(a) ** 2) for Wq, a in zip(W, (u, v, w))) +
                np.sum(Wp * Pp ** 2))

    u, v, w, Pp = forward(c, pc, us, pin)
    g = orc.adjoint(c, pc, Pp, *(2 * Wq * interior(a) for Wq, a in zip(W, (u, v, w))),
                    gp=2 * Wp * Pp, pin=pin)
    eps = 1e-6
    for _ in range(3):
        d = [interior(rng.standard_normal(c.halo_shape)) for _ in range(3)]
        fd = (loss2(pc, [a + eps * b for a, b in zip(us, d)]) -
              loss2(pc, [a - eps * b for a, b in zip(us, d)])) / (2 * eps)
        an = sum(np.sum(a * b) for a, b in zip(d, (g["gus"], g["gvs"], g["gws"])))
        assert abs(fd - an) <= 1e-6 * abs(an), ("u*", fd, an)
    gk = np.stack([g["gkx"], g["gky"], g["gkz"]], axis=-1)
    for _ in range(3):
        dk = rng.standard_normal(pc[..., 3:].shape) * pc[..., 3:]
        pp, pm = pc.copy(order="F"), pc.copy(order="F")
        pp[..., 3:] += eps * dk
        pm[..., 3:] -= eps * dk
        fd = (loss2(pp, us) - loss2(pm, us)) / (2 * eps)
        an = np.sum(gk * dk)
        assert abs(fd - an) <= 1e-6 * abs(an), ("K", fd, an)


def test_pressure_cotangent_planted_gauge_error_is_caught():
    """The pin-gauge term matters: dropping it (the mean gauge's transpose
    applied to a pinned forward) breaks the dot-product identity."""
    c = configs.generate(3, size=(8, 8, 10))
    pc = orc.gfm_coef(c)
    rng = np.random.default_rng(23)
    vdir = [interior(rng.standard_normal(c.halo_shape)) for _ in range(3)]
    Jv = forward(c, pc, vdir, (1, 1, 1))[3]
    wp = np.asfortranarray(rng.standard_normal(c.halo_shape))
    z = np.zeros(c.halo_shape, order="F")
    g = orc.adjoint(c, pc, Jv, z, z, z, gp=wp, pin=(0, 0, 0))
    lhs = np.sum(Jv * wp)
    rhs = sum(np.sum(a * b) for a, b in zip(vdir, (g["gus"], g["gvs"], g["gws"])))
    assert abs(lhs - rhs) > 1e-6 * abs(lhs)
