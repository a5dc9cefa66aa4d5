"""Pins the oracle's restatement of PrintResultVTK (oracle/vtk.py,
PrintResult.f90:328-420 through VTKFortran's binary XML writer) with the
known answers the reference's own libraries document and with what the
format fixes:

* base64: BeFoR64's doctest answers (befor64.F90) and, on random byte
  strings of every length 0..64, Python's base64 module;
* the packing of the byte count before the data: BeFoR64's pack_data
  doctest (befor64_pack_data_m.F90:485-499: last byte 63 for [0], [1.0]);
* PENF's signed integer strings (penf_stringify.F90 str_I4P doctests);
* the document: well-formed XML whose arrays decode to the int32 byte count
  and the float32 of every field in i-fastest order, face coordinates equal
  to the exact faces of a grid of binary-fraction widths, and the file name
  of the i8.8 edit descriptor.
"""

import base64
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from oracle import vtk


@pytest.mark.parametrize("data,code", [
    (np.array([0.0, -32.12], dtype="<f4").tobytes(), b"AAAAAOF6AMI="),       # :651-654
    (np.array([2023, -24], dtype="<i4").tobytes(), b"5wcAAOj///8="),          # :701-704
    (np.array([-203, -10], dtype="<i2").tobytes(), b"Nf/2/w=="),              # :726-729
    (np.array([120, -1], dtype="<i1").tobytes(), b"eP8="),                    # :751-754
    (np.array([1.0, 2.0], dtype="<f8").tobytes(), b"AAAAAAAA8D8AAAAAAAAAQA=="),  # :626-629
    (np.array([23, 324, 25456656, 2], dtype="<i8").tobytes(),
     b"FwAAAAAAAABEAQAAAAAAABBwhAEAAAAAAgAAAAAAAAA="),                       # :676-679
    (np.array([0.0], dtype="<f4").tobytes(), b"AAAAAA=="),                    # :459-464
    (np.array([23], dtype="<i8").tobytes(), b"FwAAAAAAAAA="),                 # :483-487
    (b"hello", b"aGVsbG8="),                                                  # :576-579
    (b"helloworld", b"aGVsbG93b3JsZA=="),                                     # :776-779
])
def test_base64_befor64_doctests(data, code):
    assert vtk.b64_encode(data) == code


def test_base64_matches_library_on_every_length():
    rng = np.random.default_rng(5)
    for n in range(65):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert vtk.b64_encode(b) == base64.b64encode(b), n


def test_pack_data_order():
    p = vtk.pack_i4_r4([0], [1.0])
    assert len(p) == 8 and p[-1] == 63


def test_penf_str():
    assert vtk.penf_str(-1) == "-1"
    assert vtk.penf_str(-1, no_sign=True) == "1"
    assert vtk.penf_str(0) == "+0"
    assert vtk.penf_str(129) == "+129"
    assert vtk.penf_str(1, no_sign=True) == "1"


def test_file_name():
    assert vtk.file_name(12) == "PCell_Binary_3D00000012.vtr"
    assert vtk.file_name(0) == "PCell_Binary_3D00000000.vtr"
    assert vtk.file_name(99999999) == "PCell_Binary_3D99999999.vtr"
    assert vtk.file_name(100000000) == "PCell_Binary_3D********.vtr"
    assert vtk.file_name(-3) == "PCell_Binary_3D********.vtr"


def parse(doc):
    """The document's tags and decoded arrays {name: (byte count, float32)}."""
    root = ET.fromstring(doc)
    arrays = {}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text.strip())
        arrays[da.get("Name")] = (int(np.frombuffer(raw[:4], "<i4")[0]),
                                  np.frombuffer(raw[4:], "<f4"))
    return root, arrays


def fields_for(shape, seed, halo=False):
    rng = np.random.default_rng(seed)
    I, J, K = shape
    s = (I + 2, J + 2, K + 2) if halo else shape
    out = {k: rng.standard_normal(s) * 10.0 ** rng.integers(-3, 4) for k, _ in vtk.CELL_ARRAYS}
    out["vof"] = rng.random(s)
    return out


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 2, 5), (7, 5, 4), (6, 6, 6)])
def test_document_structure_and_values(shape):
    I, J, K = shape
    # widths that are binary fractions: the faces are exact sums
    dx = np.full(I, 0.25)
    dy = 0.125 * (1 + np.arange(J) % 3)
    dz = np.full(K, 0.5)
    xc = np.cumsum(dx) - dx / 2
    yc = np.cumsum(dy) - dy / 2
    zc = 1.0 + np.cumsum(dz) - dz / 2
    f = fields_for(shape, 1 + I)
    doc = vtk.document(shape, xc, yc, zc, dx, dy, dz, f)
    assert doc.endswith(b"</VTKFile>\n")
    assert doc.startswith(b'<?xml version="1.0"?>\n<VTKFile type="RectilinearGrid" version="1.0" '
                          b'byte_order="LittleEndian">\n')
    root, arr = parse(doc)
    ext = f"+1 +{I + 1} +1 +{J + 1} +1 +{K + 1}"
    grid = root.find("RectilinearGrid")
    assert grid.get("WholeExtent") == ext
    assert grid.find("Piece").get("Extent") == ext
    names = [d.get("Name") for d in root.iter("DataArray")]
    assert names == ["X", "Y", "Z", "p", "u", "v", "w", "LiquidLvs", "FluidLvs", "LiquidVof",
                     "FluidVof"]
    for d in root.iter("DataArray"):
        assert d.get("type") == "Float32" and d.get("format") == "binary"
        assert d.get("NumberOfComponents") == "1"
    faces = {"X": np.concatenate([[0.0], np.cumsum(dx)]),
             "Y": np.concatenate([[0.0], np.cumsum(dy)]),
             "Z": 1.0 + np.concatenate([[0.0], np.cumsum(dz)])}
    for n, exact in faces.items():
        nb, v = arr[n]
        assert nb == 4 * exact.size
        assert np.array_equal(v, exact.astype(np.float32)), n
    for key, name in vtk.CELL_ARRAYS:
        nb, v = arr[name]
        assert nb == 4 * I * J * K
        # i fastest: element (i, j, k) at i + I (j + J k)
        expect = np.array([np.float32(f[key][i, j, k]) for k in range(K) for j in range(J)
                           for i in range(I)], dtype=np.float32)
        assert np.array_equal(v.view(np.uint32), expect.view(np.uint32)), name


def test_halo_arrays_read_the_interior():
    shape = (4, 3, 2)
    h = fields_for(shape, 7, halo=True)
    c = {k: a[1:-1, 1:-1, 1:-1].copy() for k, a in h.items()}
    one = np.ones(4), np.ones(3), np.ones(2)
    g = (np.arange(4) + 0.5, np.arange(3) + 0.5, np.arange(2) + 0.5)
    assert vtk.document(shape, *g, *one, h) == vtk.document(shape, *g, *one, c)
