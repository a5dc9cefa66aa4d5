"""GPU parity of the projection with a moving-solid source and with the
zero-mean gauge.

Source: b += (vof - vofO) * dx*dy*dz / dt (ProjectionP.f90:527).  vofO is
the fluid fraction of the previous step; a solid that moved up by one cell
leaves vofO(k) = vof(k+1), which is what the geometry update after a solid
move hands the next projection.  The corrected field then keeps the source
minus its fluid mean as its divergence, so the engine must report
converged = 1 on the true residual while final_div stays at the oracle's
max |mres|/V.  Both entry points (host arrays and device arrays) and both
precisions run against the oracle with the bars of the configurations.

Gauge: pin (0,0,0) asks for a zero mean over the fluid cells instead of a
pinned cell; the Fortran host uses it.  The pressure is compared with the
oracle's zero-mean solution directly and modulo its mean.
"""

import numpy as np
import pytest

from oracle import orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


def moved_solid_case(cfg, size, axis=2, seed=0):
    """Configuration ``cfg`` with vofO of the solid one cell lower along
    ``axis``: vofO(.., k, ..) = vof(.., k+1, ..); the last interior plane
    and the ghosts keep vof."""
    c = H.case(cfg, size=size, seed=seed)
    vof = c.vof
    old = vof.copy(order="F")
    sl_dst = [slice(None)] * 3
    sl_src = [slice(None)] * 3
    sl_dst[axis] = slice(1, -2)
    sl_src[axis] = slice(2, -1)
    old[tuple(sl_dst)] = vof[tuple(sl_src)]
    c.fields["vof_old"] = np.asfortranarray(old)
    return c


SOURCE_CASES = [(0, (32, 32, 32), 2), (4, (48, 24, 24), 0), (1, (24, 24, 32), 1)]


@pytest.fixture(scope="module")
def source_refs():
    cache = {}

    def get(cfg, size, axis):
        key = (cfg, size, axis)
        if key not in cache:
            c = moved_solid_case(cfg, size, axis)
            assert np.max(np.abs(c.vof - c.vof_old)) > 0.5
            cache[key] = (c, orc.project(c, pin=(1, 1, 1)))
        return cache[key]
    return get


@pytest.mark.parametrize("cfg,size,axis", SOURCE_CASES)
@pytest.mark.parametrize("prec", (0, 1))
@pytest.mark.parametrize("device", (False, True))
def test_moving_solid_source_matches_oracle(gpu_available, source_refs, cfg, size, axis, prec,
                                            device):
    c, ref = source_refs(cfg, size, axis)
    print(f"[vof-source] cfg={cfg} size={size} axis={axis} prec={prec} device={device}",
          flush=True)
    out, st = H.run_engine(c, prec, device=device)
    # the corrected field keeps source - mean(source): converged on the true
    # residual, and the divergence reported is the oracle's
    assert st.converged == 1, st.as_dict()
    assert ref.div_inf > 1.0
    assert abs(st.final_div - ref.div_inf) <= 1e-8 * ref.div_inf, (st.final_div, ref.div_inf)
    H.assert_matches_oracle(c, out, st, ref)
    # the pressure is the source's, not the velocity's alone
    c0 = H.case(cfg, size=size)
    ref0 = orc.project(c0, pin=(1, 1, 1))
    assert H.rel_inf(ref.Pp, ref0.Pp) > 1e-3


GAUGE_CASES = [(0, (24, 24, 24), 0, False), (1, (24, 24, 32), 1, True),
               (4, (48, 24, 24), 1, False), (2, (32, 32, 32), 0, True)]


@pytest.mark.parametrize("cfg,size,prec,device", GAUGE_CASES)
def test_zero_mean_gauge_matches_oracle(gpu_available, cfg, size, prec, device):
    c = H.case(cfg, size=size)
    print(f"[gauge] cfg={cfg} size={size} prec={prec} device={device}", flush=True)
    ref = orc.project(c, pin=(0, 0, 0))
    out, st = H.run_engine(c, prec, device=device, pin=(0, 0, 0))
    assert st.converged == 1 and st.final_div <= 1e-10, st.as_dict()
    H.assert_matches_oracle(c, out, st, ref)
    act = (ref.d > 0).reshape(c.shape, order="F")
    pe = out["p"][1:-1, 1:-1, 1:-1][act]
    po = ref.Pp[1:-1, 1:-1, 1:-1][act]
    # zero mean over the fluid cells, and equal modulo the mean
    assert abs(pe.mean()) <= 1e-10 * np.max(np.abs(po))
    assert np.max(np.abs((pe - pe.mean()) - (po - po.mean()))) <= 1e-8 * np.max(np.abs(po))
    # the pinned gauge differs from it by a constant only
    refp = orc.project(c, pin=(1, 1, 1))
    dp = (refp.Pp - ref.Pp)[1:-1, 1:-1, 1:-1][act]
    assert np.ptp(dp) <= 1e-8 * np.max(np.abs(po))


@pytest.mark.parametrize("prec", (0, 1))
def test_source_with_zero_mean_gauge_and_dirichlet(gpu_available, prec):
    """A moved solid with an east Dirichlet pressure face (non-singular: the
    source is not mean-projected, so the corrected field's mres is the source
    itself)."""
    c = moved_solid_case(0, (24, 24, 24), axis=0)
    c.bc_flag[1] = 0
    I, J, K = c.shape
    bcv = [None, np.zeros((J, K), order="F"), None, None, None, None]
    print(f"[vof-source-dirichlet] prec={prec}", flush=True)
    ref = orc.project(c, pin=(0, 0, 0), bc_val=bcv)
    assert not ref.singular
    out, st = H.run_engine(c, prec, pin=(0, 0, 0), bc_val=bcv)
    assert st.converged == 1 and st.singular == 0, st.as_dict()
    assert abs(st.final_div - ref.div_inf) <= 1e-8 * ref.div_inf
    H.assert_matches_oracle(c, out, st, ref)
