"""Shared helpers for the parity tests: engine inputs from a generated case
and the oracle's reference path on the same inputs."""

import numpy as np

from simcoflow3d_amd import configs

FIELD_NAMES = ("u_star", "v_star", "w_star", "dp_u", "dp_v", "dp_w", "ee_area", "ne_area",
               "te_area", "vof", "vof_old", "vof_l", "phi_l", "cell_type_p", "cell_type_u",
               "cell_type_v", "cell_type_w", "moex_u", "moex_v", "moex_w")

SIZES = {0: (24, 24, 24), 1: (24, 24, 32), 2: (32, 32, 32), 3: (20, 20, 20), 4: (48, 24, 24)}


def case(cfg, size=None, seed=0):
    return configs.generate(cfg, size=size or SIZES[cfg], seed=seed)


def fields(c):
    return {n: c.fields[n] for n in FIELD_NAMES}


def projector(c, **kw):
    from simcoflow3d_amd.engine import Projector
    params = dict(rho_l=c.rho_l, rho_g=c.rho_g, rho_ref=c.rho_l, pin_i=1, pin_j=1, pin_k=1)
    params.update(kw)
    P = Projector(c.shape, **params)
    P.set_grid_1d(c.dxp, c.dyp, c.dzp, c.dxu, c.dyv, c.dzw)
    return P


def rel_inf(a, b):
    d = np.max(np.abs(a - b))
    s = max(np.max(np.abs(b)), 1e-300)
    return d / s


def run_engine(c, prec, device=False, pin=(1, 1, 1), bc_val=None, **kw):
    """One projection of case ``c`` through sp_project (host arrays) or
    sp_project_dev (device tensors).  Returns (outputs in the host's Fortran
    layouts: p, u, v, w halo arrays and po_coef (I, J, K, 6)), stats)."""
    P = projector(c, precision=prec, pin_i=pin[0], pin_j=pin[1], pin_k=pin[2], **kw)
    try:
        if not device:
            return P.project(fields(c), c.bc_flag, c.dt, bc_val=bc_val, po_coef=True)
        import torch
        from simcoflow3d_amd.projection import to_torch_halo
        dev = {n: to_torch_halo(a) for n, a in fields(c).items()}
        hs = (c.K + 2, c.J + 2, c.I + 2)
        out = {n: torch.zeros(hs, dtype=torch.float64, device="cuda") for n in ("p", "u", "v", "w")}
        out["po_coef"] = torch.zeros((6, c.K, c.J, c.I), dtype=torch.float64, device="cuda")
        bcv = None
        if bc_val is not None:
            bcv = [None if b is None else to_torch_halo(np.asarray(b, dtype=np.float64))
                   for b in bc_val]
        st = P.project_dev(dev, c.bc_flag, c.dt, out, bc_val=bcv)
        torch.cuda.synchronize()
        res = {n: np.asfortranarray(t.cpu().numpy().T) for n, t in out.items()}
        return res, st
    finally:
        P.close()


def assert_matches_oracle(c, out, st, ref, p_tol=1e-8, u_tol=1e-8):
    """The parity bars: PoCoef bit-identical, pressure within p_tol relative,
    corrected velocity on (1:I,1:J,1:K) within u_tol of the larger of max|u*|
    and max|u| of the oracle."""
    assert np.array_equal(out["po_coef"], ref.pocoef), \
        np.max(np.abs(out["po_coef"] - ref.pocoef))
    if np.max(np.abs(ref.Pp)) > 0:
        assert rel_inf(out["p"], ref.Pp) <= p_tol, rel_inf(out["p"], ref.Pp)
    umax = max(max(np.max(np.abs(a)) for a in (c.u_star, c.v_star, c.w_star)),
               max(np.max(np.abs(a)) for a in (ref.u, ref.v, ref.w)), 1e-300)
    for n, r in (("u", ref.u), ("v", ref.v), ("w", ref.w)):
        d = out[n][1:-1, 1:-1, 1:-1] - r[1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(d)) <= u_tol * umax, (n, np.max(np.abs(d)) / umax)
