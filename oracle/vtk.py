"""CPU restatement of PrintResultVTK (PrintResult.f90:328-420), the binary
VTK XML RectilinearGrid file the reference writes through VTKFortran's
binary-local writer.  Test infrastructure only: the tests and the bench's
CPU leg use it as the checker; the engine never imports it.

The document, byte for byte:
* `<?xml version="1.0"?>` and the VTKFile tag on their own lines
  (vtk_fortran_vtk_file_xml_writer_abstract.f90:776-794, little endian);
* the topology and Piece start tags with WholeExtent / Extent
  "1 I+1 1 J+1 1 K+1" printed by PENF's str, which prefixes a '+' to every
  non-negative integer (penf_stringify.F90 str_I4P; :845-900, :1084-1106);
* Coordinates X, Y, Z (:1162-1181): cell faces as float32, from the cell
  centres and widths (PrintResult.f90:347-361);
* CellData (:968-1029) with p, u, v, w, LiquidLvs, FluidLvs, LiquidVof,
  FluidVof (PrintResult.f90:374-381), each packed i fastest
  (PrintResult.f90:396-419);
* every DataArray tag (:903-933) laid out by FoXy's xml_tag (start tag
  indented, content on its own line indented two more, end tag indented;
  foxy_xml_tag.F90:293-342), its content the base64 code of the int32 byte
  count followed by the float32 data (vtk_fortran_dataarray_encoder.f90:
  1047-1062; BeFoR64's encode_bits, befor64.F90, standard alphabet, '='
  padding);
* start tags indent by two spaces per level and every record ends with
  char(10).
The file is PCell_Binary_3D<itt as i8.8>.vtr; i8.8 prints eight asterisks for
an itt it cannot hold (negative or above 99999999).
"""

import numpy as np

_ALPHABET = np.frombuffer(
    b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/", dtype=np.uint8)

CELL_ARRAYS = (("p", "p"), ("u", "u"), ("v", "v"), ("w", "w"), ("phi_l", "LiquidLvs"),
               ("phi", "FluidLvs"), ("vof_l", "LiquidVof"), ("vof", "FluidVof"))


def b64_encode(data):
    """BeFoR64's encode_bits: each 3 bytes become four 6-bit indexes into the
    alphabet, high bits first; a short last group is zero-filled and its
    missing characters replaced by '='."""
    b = np.frombuffer(bytes(data), dtype=np.uint8)
    n = b.size
    pad = (3 - n % 3) % 3
    g = np.concatenate([b, np.zeros(pad, np.uint8)]).reshape(-1, 3).astype(np.uint32)
    word = (g[:, 0] << 16) | (g[:, 1] << 8) | g[:, 2]
    idx = np.stack([(word >> s) & 63 for s in (18, 12, 6, 0)], axis=1)
    out = _ALPHABET[idx].reshape(-1)
    if pad:
        out[-pad:] = ord("=")
    return out.tobytes()


def penf_str(n, no_sign=False):
    """PENF's str for integers: '+' before every non-negative value unless
    no_sign."""
    s = str(int(n))
    if n >= 0:
        s = "+" + s
    return s[1:] if no_sign else s


def file_name(itt):
    """'PCell_Binary_3D'//trim(curd)//'.vtr' with curd written as i8.8."""
    curd = f"{int(itt):08d}" if 0 <= int(itt) <= 99999999 else "*" * 8
    return f"PCell_Binary_3D{curd}.vtr"


def face_coordinates(centres, widths):
    """x(i) = xc(i) - dx(i)/2 for i = 1..n and x(n+1) = xc(n) + dx(n)/2, in
    double, then stored as float32 (PrintResult.f90:347-361)."""
    c = np.asarray(centres, dtype=np.float64)
    d = np.asarray(widths, dtype=np.float64)
    x = np.empty(c.size + 1)
    x[:-1] = c - d / 2.0
    x[-1] = c[-1] + d[-1] / 2.0
    return x.astype(np.float32)


def pack_i4_r4(a1, a2):
    """BeFoR64's pack_data for int32 a1 and float32 a2: the bytes of a1, then
    those of a2, in the machine's (little-endian) order
    (befor64_pack_data_m.F90:485-509)."""
    return (np.ascontiguousarray(a1, dtype="<i4").tobytes()
            + np.ascontiguousarray(a2, dtype="<f4").tobytes())


def encode_float32(values):
    """encode_binary_dataarray1_rank1_R4P (vtk_fortran_dataarray_encoder.f90:
    1047-1062): base64 of pack_data([int32 4 n], x)."""
    v = np.ascontiguousarray(values, dtype="<f4").reshape(-1)
    return b64_encode(pack_i4_r4([4 * v.size], v))


def _cells(a, shape):
    """The field on (1:I, 1:J, 1:K), i fastest, as float32; a is a cell array
    (I, J, K) or a halo array (I+2, J+2, K+2), both indexed [i, j, k]."""
    a = np.asarray(a, dtype=np.float64)
    if a.shape == tuple(s + 2 for s in shape):
        a = a[1:-1, 1:-1, 1:-1]
    assert a.shape == tuple(shape), a.shape
    return a.reshape(-1, order="F").astype(np.float32)


def document(shape, xc, yc, zc, dx, dy, dz, fields):
    """The bytes of the file PrintResultVTK writes; fields maps p, u, v, w,
    phi_l, phi, vof_l, vof to cell or halo arrays indexed [i, j, k]."""
    I, J, K = shape
    nl = "\n"
    out = []

    def rec(indent, text):
        out.append(" " * indent + text + nl)

    def data_array(indent, name, values):
        attrs = (f'type="Float32" NumberOfComponents="{penf_str(1, True)}" Name="{name}" '
                 f'format="binary"')
        code = encode_float32(values).decode()
        out.append(" " * indent + f"<DataArray {attrs}>" + nl + " " * (indent + 2) + code + nl
                   + " " * indent + "</DataArray>" + nl)

    ext = " ".join(penf_str(v) for v in (1, I + 1, 1, J + 1, 1, K + 1))
    out.append('<?xml version="1.0"?>' + nl)
    out.append('<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian">' + nl)
    rec(2, f'<RectilinearGrid WholeExtent="{ext}">')
    rec(4, f'<Piece Extent="{ext}">')
    rec(6, "<Coordinates>")
    for name, c, d in (("X", xc, dx), ("Y", yc, dy), ("Z", zc, dz)):
        data_array(8, name, face_coordinates(c, d))
    rec(6, "</Coordinates>")
    rec(6, "<CellData>")
    for key, name in CELL_ARRAYS:
        data_array(8, name, _cells(fields[key], shape))
    rec(6, "</CellData>")
    rec(4, "</Piece>")
    rec(2, "</RectilinearGrid>")
    rec(0, "</VTKFile>")
    return "".join(out).encode()
