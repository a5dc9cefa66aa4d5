"""The Fortran side of the boundary: include/simco_proj.f90 (ISO_C_BINDING).

* The bind(C) types have the C structs' sizes and member offsets (checked
  against the ctypes mirror of include/simco_proj.h).
* On the GPU, a Fortran host program (tests/fortran/sp_drive.f90) fills
  sp_inputs with c_loc of arrays declared with the host's own bounds, calls
  sp_project, and reproduces the golden fixtures."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from simcoflow3d_amd import engine
from tests import golden_io
from tests import helpers as H

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "fortran", "_build")
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def _program(name):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        if not os.path.exists(FLANG):
            pytest.skip("flang not available and the Fortran programs are not built")
        from simcoflow3d_amd import build
        build.build_fortran()
    return path


def test_fortran_types_match_c_layout():
    out = subprocess.run([_program("sp_layout")], capture_output=True, text=True, check=True)
    got = dict(line.rsplit(" ", 1) for line in out.stdout.strip().splitlines())
    got = {k: int(v) for k, v in got.items()}
    structs = {"sp_params": engine.SpParams, "sp_inputs": engine.SpInputs,
               "sp_outputs": engine.SpOutputs, "sp_stats": engine.SpStats,
               "sp_adj_inputs": engine.SpAdjInputs, "sp_adj_outputs": engine.SpAdjOutputs,
               "sp_wall_bc": engine.SpWallBc, "sp_vtr_fields": engine.SpVtrFields,
               "sp_vtk_fields": engine.SpVtkFields, "sp_solid_body": engine.SpSolidBody,
               "sp_lvs_grid": engine.SpLvsGrid, "sp_clsvof_inputs": engine.SpClsvofInputs,
               "sp_clsvof_state": engine.SpClsvofState,
               "sp_cell_geom": engine.SpCellGeom, "sp_cell_interp": engine.SpCellInterp,
               "sp_liquid_field": engine.SpLiquidField, "sp_pred_inputs": engine.SpPredInputs,
               "sp_pred_outputs": engine.SpPredOutputs}
    for name, cls in structs.items():
        assert got[name] == ctypes.sizeof(cls), name
    for key, v in got.items():
        if "." not in key:
            continue
        sname, field = key.split(".")
        assert getattr(structs[sname], field).offset == v, key


@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_io.NAMES)
def test_fortran_host_matches_golden(gpu_available, tmp_path, name):
    cfg, c, exp = golden_io.load(name)
    print(f"[fortran] {name}", flush=True)
    I, J, K = c.shape
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as fh:
        fh.write(f"{I} {J} {K}\n{c.dt!r} {c.rho_l!r} {c.rho_g!r}\n1 1 1 0\n")
        fh.write(" ".join(str(int(v)) for v in c.bc_flag) + "\n")
    for n, a in H.fields(c).items():
        a = np.asarray(a)
        dt = np.int32 if a.dtype.kind == "i" else np.float64
        a.astype(dt).ravel(order="F").tofile(os.path.join(d, n + ".bin"))
    np.concatenate([c.dxp, c.dyp, c.dzp, c.dxu, c.dyv, c.dzw]).astype(np.float64).tofile(
        os.path.join(d, "grid.bin"))
    r = subprocess.run(["timeout", "-k", "5", "60", _program("sp_drive"), d],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    hs = c.halo_shape

    def rd(n, shape):
        return np.fromfile(os.path.join(d, n + ".bin"), dtype=np.float64).reshape(shape, order="F")

    pocoef = rd("out_pocoef", (I, J, K, 6))
    assert np.array_equal(pocoef, exp["pocoef"])
    assert H.rel_inf(rd("out_p", hs), exp["p"]) <= 1e-8
    umax = max(np.max(np.abs(c.u_star)), np.max(np.abs(c.v_star)), np.max(np.abs(c.w_star)))
    for n in ("u", "v", "w"):
        a = rd("out_" + n, hs)
        assert np.max(np.abs(a[1:-1, 1:-1, 1:-1] - exp[n][1:-1, 1:-1, 1:-1])) <= 1e-8 * umax
    it, conv, lev, div = open(os.path.join(d, "stats.txt")).read().split()
    assert int(conv) == 1 and float(div) <= 1e-10
