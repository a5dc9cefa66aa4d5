"""The host-array entry point (sp_project) as a Fortran host drives it step
after step:

* sp_inputs.static_geometry = 1 reuses the device copies of the apertures,
  vof, vofO, cell types and MoExCell of the previous call: with a new u* the
  result equals that of a call that copies everything, and a context whose
  previous call did not copy them ignores the flag;
* only the entries (1:I,1:J,1:K) of u, v, w come back: the caller's other
  entries are kept (no upload of the output arrays is needed for that);
* a stale-face list that would overflow returns SP_ERR_CAPACITY, a status of
  its own, with nothing written.
"""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _new_velocity(c, seed):
    f = H.fields(c)
    g = dict(f)
    for q, n in enumerate(("u_star", "v_star", "w_star")):
        a = configs.uniform(seed, 50 + q, -1.0, 1.0, c.halo_shape)
        a[f[n] == 0.0] = 0.0  # keep the walls' and solid faces' zeros
        g[n] = np.asfortranarray(a)
    return g


@pytest.mark.parametrize("cfg,size,prec", [(0, (24, 24, 24), 1), (4, (40, 20, 24), 0)])
def test_static_geometry_reuses_device_copies(gpu_available, cfg, size, prec):
    c = H.case(cfg, size=size)
    print(f"[host-path] static geometry cfg={cfg} size={size} prec={prec}", flush=True)
    f1 = H.fields(c)
    f2 = _new_velocity(c, seed=9)
    P = H.projector(c, precision=prec)
    # the first call of a context copies everything whatever the flag says
    a1, s1 = P.project(f1, c.bc_flag, c.dt, static_geometry=True)
    a2, s2 = P.project(f2, c.bc_flag, c.dt, static_geometry=True)
    P.close()
    Q = H.projector(c, precision=prec)
    b2, t2 = Q.project(f2, c.bc_flag, c.dt)
    Q.close()
    assert s1.converged == 1 and s2.converged == 1 and t2.converged == 1
    assert s2.iterations == t2.iterations
    for n in ("p", "u", "v", "w"):
        assert np.array_equal(a2[n], b2[n]), n
    ref = orc.project(c, pin=(1, 1, 1))
    assert H.rel_inf(a1["p"], ref.Pp) <= 1e-8


def test_changed_geometry_is_copied_again(gpu_available):
    """A moved solid (static_geometry = 0) after a static call gives the
    oracle's result for the new geometry."""
    c = H.case(0, size=(24, 24, 24))
    d = H.case(0, size=(24, 24, 24), seed=3)
    print("[host-path] geometry change", flush=True)
    P = H.projector(c, precision=1)
    P.project(H.fields(c), c.bc_flag, c.dt)
    P.project(H.fields(c), c.bc_flag, c.dt, static_geometry=True)
    out, st = P.project(H.fields(d), d.bc_flag, d.dt, po_coef=True)
    P.close()
    ref = orc.project(d, pin=(1, 1, 1))
    assert st.converged == 1
    H.assert_matches_oracle(d, out, st, ref)


def test_halo_entries_of_outputs_are_kept(gpu_available):
    c = H.case(1, size=(20, 18, 26))
    print("[host-path] output halos", flush=True)
    hs = c.halo_shape
    rng = np.random.default_rng(5)
    out = {n: np.asfortranarray(rng.uniform(-9.0, 9.0, hs)) for n in ("u", "v", "w")}
    keep = {n: a.copy(order="F") for n, a in out.items()}
    out["p"] = np.zeros(hs, order="F")
    P = H.projector(c, precision=1)
    out, st = P.project(H.fields(c), c.bc_flag, c.dt, out=out)
    P.close()
    assert st.converged == 1
    inner = np.zeros(hs, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    ref = orc.project(c, pin=(1, 1, 1))
    for n, r in (("u", ref.u), ("v", ref.v), ("w", ref.w)):
        assert np.array_equal(out[n][~inner], keep[n][~inner]), n
        assert np.max(np.abs(out[n][inner] - r[inner])) <= 1e-8, n


def test_stale_face_overflow_has_its_own_status(gpu_available):
    """Every other cell solid (a checkerboard of vof = 0) gives each fluid
    cell six faces next to an unclassified cell: 6.3 M stale faces at 128^3,
    above the list's capacity of about 1.05 M."""
    from simcoflow3d_amd.engine import SP_ERR_CAPACITY, SpError
    c = H.case(2, size=(128, 128, 128))
    I, J, K = c.shape
    ii, jj, kk = np.meshgrid(np.arange(I + 2), np.arange(J + 2), np.arange(K + 2), indexing="ij")
    solid = ((ii + jj + kk) % 2 == 0)
    f = H.fields(c)
    f["vof"] = np.asfortranarray(np.where(solid, 0.0, 1.0))
    f["vof_old"] = f["vof"]
    ct = np.asfortranarray(solid[1:-1, 1:-1, 1:-1].astype(np.int32) * 2)
    f["cell_type_p"] = ct
    print("[host-path] stale-face overflow 128^3", flush=True)
    P = H.projector(c, precision=1)
    with pytest.raises(SpError) as ei:
        P.project(f, c.bc_flag, c.dt)
    P.close()
    assert ei.value.status == SP_ERR_CAPACITY


def test_static_geometry_survives_other_entry_points(gpu_available):
    """The device copies of the geometry that a static_geometry projection
    reuses (and the forward inputs a later adjoint reads) are not touched by
    the host-array calls in between: the momentum predictor, the geometry
    update and the liquid fields of the same grid size."""
    from simcoflow3d_amd import configs
    from tests import pred_cases as PC
    c = H.case(0, size=(24, 24, 24))
    print("[host-path] static geometry after other entry points", flush=True)
    f1 = H.fields(c)
    f2 = _new_velocity(c, seed=9)
    P = H.projector(c, precision=1)
    P.project(f1, c.bc_flag, c.dt, static_geometry=True)
    # other host-array entry points on the same context
    size = c.shape
    grid, geo, interp = PC.geometry(size, shape="bodies", seed=2)
    f = configs.predictor_fields(size, seed=4, itt=2)
    cells = configs.pred_cells(geo, interp)
    hs = tuple(n + 2 for n in size)
    pred = [np.zeros(hs, order="F") for _ in range(3)]
    dp = [np.zeros(hs, order="F") for _ in range(3)]
    ff = dict(f, flux_div_old=np.array(f["flux_div_old"], order="F", copy=True))
    for so in (1, 2):
        ff["params"] = dict(f["params"], space_order=so)
        P.predict_uvw(PC.p_areas(geo), cells, cells, ff, pred, dp)
    P.grid_preprocess(configs.geometry_case(size, "sphere", seed=1))
    a2, s2 = P.project(f2, c.bc_flag, c.dt, static_geometry=True)
    P.close()
    Q = H.projector(c, precision=1)
    b2, t2 = Q.project(f2, c.bc_flag, c.dt)
    Q.close()
    assert s2.converged == 1 and t2.converged == 1
    for n in ("p", "u", "v", "w"):
        assert np.array_equal(a2[n], b2[n]), n
