"""The golden VTR file tests/golden/vtr_case_7x5x4.vtr was written by the
reference's own writer (src/VTR_mod.f90 built by oracle/Makefile, called as
PrintResultVTR3D does, PrintResult.f90:131-146) on the fields of
vtr_case_7x5x4.npz.  These CPU checks pin what the GPU writer is compared
with: the document structure and the float32 values of every array."""

import os

import numpy as np

from tests import vtr_io

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "vtr_case_7x5x4.vtr")
INPUTS = os.path.join(HERE, "golden", "vtr_case_7x5x4.npz")


def expected_arrays(f):
    """real(., kind=4) of every written array, in the writer's element order
    (Fortran order, vector components interleaved)."""
    flat = {n: np.asarray(f[n], dtype=np.float64).ravel(order="F") for n in f.files if n != "itt"}
    vel = np.stack([flat["u"], flat["v"], flat["w"]], axis=1).ravel()
    out = {"X_COORDINATES": flat["x"], "Y_COORDINATES": flat["y"], "Z_COORDINATES": flat["z"],
           "Velocity": vel, "X_Velocity": flat["u"], "Y_Velocity": flat["v"],
           "Z_Velocity": flat["w"], "Pressure": flat["p"], "FluidLvs": flat["phi"],
           "LiquidLvs": flat["phi_l"], "FluidVof": flat["vof"], "LiquidVof": flat["vof_l"],
           "Mass error": flat["mres"]}
    return {k: v.astype(np.float32) for k, v in out.items()}


def test_golden_vtr_holds_float32_of_the_inputs():
    f = np.load(INPUTS)
    tags, arrays = vtr_io.parse(GOLD)
    exp = expected_arrays(f)
    assert list(arrays) == list(exp)
    for k, v in exp.items():
        assert np.array_equal(arrays[k].view(np.uint32), v.view(np.uint32)), k
    assert tags[0] == '<VTKFile type="RectilinearGrid" version="0.1" format="ascii">'
    assert tags[1] == ' <RectilinearGrid WholeExtent="1 7 1 5 1 4">'
    assert tags[-1] == "</VTKFile>"
