"""The engine's binary VTK XML writer (sp_write_vtk, PrintResultVTK of
PrintResult.f90:328-420) against the oracle's restatement, byte for byte:

* host cell arrays on a stretched grid, and device halo tensors;
* 160^3, where every cell array spans two base64 launches;
* the i8.8 name for an itt it cannot hold, the working folder for dir "",
  and the rejection of a missing field."""

import os

import numpy as np
import pytest

from oracle import vtk
from simcoflow3d_amd import engine

pytestmark = pytest.mark.gpu


def _grid(shape, seed):
    rng = np.random.default_rng(seed)
    d = [0.5 + rng.random(n) for n in shape]
    c = [np.cumsum(w) - w / 2 for w in d]
    return c, d


def _projector(shape, d):
    from simcoflow3d_amd.engine import Projector
    P = Projector(shape)
    P.set_grid_1d(*d, *d)
    return P


def _fields(shape, seed, halo=False):
    rng = np.random.default_rng(seed)
    I, J, K = shape
    s = (I + 2, J + 2, K + 2) if halo else shape
    f = {k: np.asfortranarray(rng.standard_normal(s) * 10.0 ** rng.integers(-4, 5))
         for k in engine.VTK_FIELDS}
    f["vof"] = np.asfortranarray(rng.random(s))
    f["p"][tuple(0 for _ in s)] = 0.0
    return f


def _args(c, d, f):
    a = dict(x=c[0], y=c[1], z=c[2], dx=d[0], dy=d[1], dz=d[2])
    a.update(f)
    return a


def _read(path):
    with open(path, "rb") as fh:
        return fh.read()


@pytest.mark.parametrize("shape", [(7, 5, 4), (2, 2, 2), (33, 17, 9)])
def test_vtk_host_and_device_match_oracle(gpu_available, tmp_path, shape):
    import torch
    print(f"[vtk] {shape}", flush=True)
    c, d = _grid(shape, sum(shape))
    f = _fields(shape, 3 + shape[0])
    expect = vtk.document(shape, *c, *d, f)
    P = _projector(shape, d)
    path = P.write_vtk(str(tmp_path) + "/", 12, _args(c, d, f))
    assert os.path.basename(path) == "PCell_Binary_3D00000012.vtr"
    assert _read(path) == expect
    # device tensors, the fields as halo arrays (0:I+1, ...)
    dev = {}
    for n, a in _args(c, d, f).items():
        if a.ndim == 3:
            h = np.zeros(tuple(s + 2 for s in a.shape), order="F")
            h[1:-1, 1:-1, 1:-1] = a
            a = h
        dev[n] = torch.from_numpy(np.asfortranarray(a).T.copy()).cuda()
    path2 = P.write_vtk(str(tmp_path / "dev") + "/", 12, dev, device=True)
    P.close()
    assert _read(path2) == expect


def test_vtk_two_launches_per_array(gpu_available, tmp_path):
    import torch
    shape = (160, 160, 160)
    print(f"[vtk] {shape}", flush=True)
    c, d = _grid(shape, 11)
    f = _fields(shape, 12)
    expect = vtk.document(shape, *c, *d, f)
    P = _projector(shape, d)
    dev = {n: torch.from_numpy(np.asfortranarray(a).T.copy()).cuda()
           for n, a in _args(c, d, f).items()}
    P.set_profiling(True)
    path = P.write_vtk(str(tmp_path) + "/", 3, dev, device=True)
    rows = [r for r in P.kernel_times() if r["name"] == "vtk_base64"]
    P.set_profiling(False)
    P.close()
    # 3 coordinate arrays in one launch each, 8 cell arrays in two
    assert rows and rows[0]["launches"] == 3 + 8 * 2
    assert _read(path) == expect


def test_vtk_name_and_folder(gpu_available, tmp_path, monkeypatch):
    shape = (4, 3, 2)
    c, d = _grid(shape, 1)
    f = _fields(shape, 2)
    P = _projector(shape, d)
    monkeypatch.chdir(tmp_path)
    path = P.write_vtk("", 123456789, _args(c, d, f))
    assert path == "PCell_Binary_3D********.vtr"
    assert os.path.exists(tmp_path / path)
    assert _read(tmp_path / path) == vtk.document(shape, *c, *d, f)
    args = _args(c, d, f)
    args["dz"] = None
    with pytest.raises(RuntimeError, match="null VTK"):
        P.write_vtk(str(tmp_path) + "/", 1, args)
    P.close()
