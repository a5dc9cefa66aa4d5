"""Writes the golden VTR fixture from the reference's own writer.

Builds oracle/_ref/vtr_ref_driver (oracle/Makefile target ``ref``: the
reference's src/VTR_mod.f90 and src/Precision.f90 with a host program of this
repository that calls them as PrintResultVTR3D does, PrintResult.f90:131-146)
and runs it on seeded fields of a 7x5x4 grid.  Commits the inputs
(vtr_case_7x5x4.npz) and the file it writes (vtr_case_7x5x4.vtr).  Run in a
container where the reference tree and amdflang exist:

    python tests/golden/make_vtr_golden.py
    python tests/golden/make_vtr_golden.py --time 128   # times the reference
                                                        # writer on a 128^3 case
"""

import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
SHAPE = (7, 5, 4)
ITT = 12
NAMES = ("u", "v", "w", "p", "phi", "phi_l", "vof", "vof_l", "mres")


def fields(seed=3):
    I, J, K = SHAPE
    rng = np.random.default_rng(seed)
    f = {"x": (np.arange(I) + 0.5) / I, "y": (np.arange(J) + 0.5) / J,
         "z": np.cumsum(rng.uniform(0.05, 0.3, K))}
    scales = {"u": 1.0, "v": 1.0, "w": 1.0, "p": 300.0, "phi": 0.5, "phi_l": 0.5, "vof": 1.0,
              "vof_l": 1.0, "mres": 1e-15}
    for n in NAMES:
        a = rng.standard_normal(SHAPE) * scales[n]
        if n in ("vof", "vof_l"):
            a = np.clip(np.abs(a), 0.0, 1.0)
        f[n] = np.asfortranarray(a)
    # values that exercise the number formatting: zero, one, tenths, float32
    # subnormals, powers of ten at the edges of the two-digit exponent, ties
    sp = [0.0, -0.0, 1.0, 0.1, -0.1, 1e-20, -1e-40, 1e-44, 3.0e38, -2.5e-38, 123456.789,
          0.999999999, 9.9999999e9, 1.0 / 3.0, 2.0 ** -24, 65504.0, -7.0e-3]
    flat = f["p"].reshape(-1, order="F")
    flat[:len(sp)] = sp
    f["p"] = np.asfortranarray(flat.reshape(SHAPE, order="F"))
    return f


def time_reference(n):
    """Wall time of the reference writer on seeded n^3 fields (CPU)."""
    import time
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    exe = os.path.join(REPO, "oracle", "_ref", "vtr_ref_driver")
    rng = np.random.default_rng(0)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "meta.txt"), "w") as fh:
            fh.write(f"{n} {n} {n} 1\n")
        for c in "xyz":
            ((np.arange(n) + 0.5) / n).tofile(os.path.join(d, c + ".bin"))
        for name in NAMES:
            rng.standard_normal(n ** 3).tofile(os.path.join(d, name + ".bin"))
        out = os.path.join(d, "out") + "/"
        os.makedirs(out + "Paraview")
        t0 = time.perf_counter()
        subprocess.run([exe, d, out], check=True)
        el = time.perf_counter() - t0
        size = os.path.getsize(os.path.join(out, "Paraview", "case_00000001.vtr"))
    vals = 12 * n ** 3 + 3 * n
    print(f"reference writer {n}^3: {el:.2f} s, {size / 1e6:.0f} MB, {vals / el / 1e6:.2f} M values/s")


def main():
    import sys
    if len(sys.argv) > 2 and sys.argv[1] == "--time":
        time_reference(int(sys.argv[2]))
        return
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    exe = os.path.join(REPO, "oracle", "_ref", "vtr_ref_driver")
    f = fields()
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "meta.txt"), "w") as fh:
            fh.write(f"{SHAPE[0]} {SHAPE[1]} {SHAPE[2]} {ITT}\n")
        for n, a in f.items():
            np.asarray(a, dtype=np.float64).ravel(order="F").tofile(os.path.join(d, n + ".bin"))
        out = os.path.join(d, "out") + "/"
        os.makedirs(out + "Paraview")
        subprocess.run([exe, d, out], check=True)
        name = f"case_{ITT:08d}.vtr"
        shutil.copy(os.path.join(out, "Paraview", name),
                    os.path.join(HERE, "vtr_case_7x5x4.vtr"))
    np.savez(os.path.join(HERE, "vtr_case_7x5x4.npz"), itt=ITT, **f)
    print("wrote", os.path.join(HERE, "vtr_case_7x5x4.vtr"))


if __name__ == "__main__":
    main()
