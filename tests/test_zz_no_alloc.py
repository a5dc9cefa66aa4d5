"""The allocation-free NumberExternalCell (sp_number_external_cell_noalloc)
against the allocating entry point (sp_number_external_cell, device arrays)
and the oracle, on two calls in a row that share the GPU's static scratch:
the second on a smaller grid, so that a leftover of the first call would
show.  Bit-identical PosNu and ExtCell are required (integer work)."""

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs

pytestmark = pytest.mark.gpu

# (size, iu, iv, iw): the first call larger than the second
CALLS = [((96, 80, 200), 0, 0, 1), ((40, 30, 70), 1, 0, 0)]


def _cell_type(size, seed):
    cells = configs.geometry_case(size, "bodies", seed=seed, stretch=False)
    exp = orc.grid_preprocess(cells)
    return np.asfortranarray(exp[0]["cell_type"])


def test_noalloc_numbering_matches_allocating_entry_point(gpu_available):
    import torch
    from simcoflow3d_amd.engine import Projector
    for q, (size, iu, iv, iw) in enumerate(CALLS):
        print(f"[noalloc] call {q}: size={size} iu,iv,iw={iu},{iv},{iw}", flush=True)
        ct = _cell_type(size, seed=q + 3)
        assert np.any(ct == 2) and np.any(ct != 2)
        pos_ref, ext_ref = orc.number_external(ct, iu, iv, iw)
        I, J, K = size
        P = Projector(size)
        try:
            dct = torch.from_numpy(ct.T.copy()).cuda()
            shp = (K - iw, J - iv, I - iu)
            pos_a = torch.full(shp, -7, dtype=torch.int32, device="cuda")
            _, ext_a = P.number_external(dct, iu, iv, iw, posnu=pos_a, device=True)
            pos_b = torch.full(shp, -7, dtype=torch.int32, device="cuda")
            ext_b = torch.full((1,), -7, dtype=torch.int32, device="cuda")
            P.number_external_noalloc(dct, iu, iv, iw, pos_b, ext_b)
            a = pos_a.cpu().numpy().T
            b = pos_b.cpu().numpy().T
        finally:
            P.close()
        assert ext_a == ext_ref and int(ext_b.item()) == ext_ref, (ext_a, ext_b, ext_ref)
        assert np.array_equal(a, pos_ref)
        assert np.array_equal(b, pos_ref)


def test_noalloc_rejects_grids_beyond_the_scratch(gpu_available):
    """A numbering that needs more than the 2^22 int32 of static scratch is
    refused with SP_ERR_INVALID_ARG before any work is enqueued."""
    import torch
    from simcoflow3d_amd.engine import Projector, SpError, SP_ERR_INVALID_ARG
    size = (2100, 2100, 2)  # 4.41 M pieces of one chunk each
    print(f"[noalloc] oversize {size}", flush=True)
    P = Projector(size)
    try:
        ct = torch.zeros((1,), dtype=torch.int32, device="cuda")
        with pytest.raises(SpError) as ei:
            P.number_external_noalloc(ct, 0, 0, 0, ct, ct)
        assert ei.value.status == SP_ERR_INVALID_ARG
    finally:
        P.close()
