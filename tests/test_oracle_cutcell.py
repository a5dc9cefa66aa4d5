"""Pins of the oracle's cut-cell geometry (orc_face_geo, orc_grid_preprocess;
GridPreprocess, CutCell.f90:62-325, 883-981) by checks that do not retype
its formulas:

* FaceGeoCal against an independent computation of the clipped area and
  centroid: the positive part of the face is integrated column by column
  (its height is piecewise linear in x, so Simpson's rule on each piece is
  exact), plus the closed forms of full, empty, half and corner-triangle
  faces;
* the reference's treatment of corners lying exactly on the line (dropped:
  a diagonal through two corners gives area 0), a quirk kept on purpose;
* GridPreprocess on planar solids: the shared face seen from both cells
  agrees with the direct clipping at the face, WlLh of a cut cell equals
  the exact length of the line in the cell times the cell depth, U/V/W
  faces through a cut P centre equal the exact mid-plane fraction, faces
  next to solid cells are zero, the DefineCell classes and vofL scaling,
  the stretched-spacing face-centre rule, and the reference's FCT(:,:,:,2)
  sum, which weighs the bottom face's first centroid component.
"""

import math

import numpy as np
import pytest

from oracle import orc
from simcoflow3d_amd import configs


def clipped_face(d1, d2, n1, n2, c):
    """Area fraction and centroid of {n1 x + n2 y + c > 0} in
    [-d1/2,d1/2] x [-d2/2,d2/2], by exact piecewise integration over x."""
    a, b = d1 / 2.0, d2 / 2.0

    def interval(x):  # the fluid part [lo, hi] of the column at x
        if n2 == 0.0:
            return (-b, b) if n1 * x + c > 0.0 else (0.0, 0.0)
        y0 = (-c - n1 * x) / n2
        if n2 > 0.0:
            lo = min(max(y0, -b), b)
            return lo, b
        hi = min(max(y0, -b), b)
        return -b, hi

    cuts = [-a, a]
    if n2 != 0.0:
        for yb in (-b, b):
            if n1 != 0.0:
                cuts.append((-c - n2 * yb) / n1)
    elif n1 != 0.0:
        cuts.append(-c / n1)
    cuts = sorted(x for x in cuts if -a <= x <= a)
    A = Mx = My = 0.0
    for x0, x1 in zip(cuts[:-1], cuts[1:]):
        if x1 <= x0:
            continue
        xs = (x0, 0.5 * (x0 + x1), x1)
        L, XL, YL = [], [], []
        for x in xs:
            # with n2 = 0 the column height is a step: take it inside the piece
            lo, hi = interval(x if n2 != 0.0 else xs[1])
            ln = hi - lo
            L.append(ln)
            XL.append(x * ln)
            YL.append(0.5 * (hi * hi - lo * lo))
        w = (x1 - x0) / 6.0
        A += w * (L[0] + 4 * L[1] + L[2])
        Mx += w * (XL[0] + 4 * XL[1] + XL[2])
        My += w * (YL[0] + 4 * YL[1] + YL[2])
    if A <= 0.0:
        return 0.0, 0.0, 0.0
    return A / (d1 * d2), Mx / A, My / A


def test_face_geo_against_exact_clipping():
    rng = np.random.default_rng(7)
    for _ in range(2000):
        d1, d2 = rng.uniform(0.2, 3.0, 2)
        th = rng.uniform(0, 2 * math.pi)
        n1, n2 = math.cos(th), math.sin(th)
        n3, d3 = rng.uniform(-1, 1), rng.uniform(-1, 1)
        phi = rng.uniform(-1.5, 1.5) * max(d1, d2)
        area, cent = orc.face_geo(d1, d2, n1, n2, n3, d3, phi)
        ea, ex, ey = clipped_face(d1, d2, n1, n2, n3 * d3 + phi)
        assert area == pytest.approx(ea, abs=1e-12)
        if ea > 1e-9:
            assert cent[0] == pytest.approx(ex, abs=1e-10 * max(d1, d2))
            assert cent[1] == pytest.approx(ey, abs=1e-10 * max(d1, d2))
        assert cent[2] == d3


def test_face_geo_closed_forms():
    # full and empty faces
    a, c = orc.face_geo(2.0, 0.5, 0.3, 0.4, 0.0, 0.0, 10.0)
    assert a == 1.0 and c[0] == 0.0 and c[1] == 0.0
    a, c = orc.face_geo(2.0, 0.5, 0.3, 0.4, 0.0, 0.7, -10.0)
    assert a == 0.0 and list(c) == [0.0, 0.0, 0.7]
    # half face x > 0: centroid at d1/4
    a, c = orc.face_geo(2.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0)
    assert a == 0.5 and c[0] == pytest.approx(0.5) and c[1] == pytest.approx(0.0)
    # the third component enters through n3*d3: x + 0.5*0.4 > 0 on [-1,1]
    a, c = orc.face_geo(2.0, 1.0, 1.0, 0.0, 0.5, 0.4, 0.0)
    assert a == pytest.approx(0.6) and c[0] == pytest.approx(0.4) and c[2] == 0.4
    # corner triangle at (+a,+b) with legs L along both axes
    d1, d2, L = 2.0, 1.0, 0.3
    s = 1 / math.sqrt(2)
    phi = -s * (d1 / 2 + d2 / 2 - L)
    a, c = orc.face_geo(d1, d2, s, s, 0.0, 0.0, phi)
    assert a == pytest.approx(L * L / 2 / (d1 * d2))
    assert c[0] == pytest.approx(d1 / 2 - L / 3) and c[1] == pytest.approx(d2 / 2 - L / 3)
    # orientation: swapping the normal's sign complements the area
    a1, _ = orc.face_geo(1.3, 0.7, 0.6, -0.8, 0.2, 0.1, 0.05)
    a2, _ = orc.face_geo(1.3, 0.7, -0.6, 0.8, -0.2, 0.1, -0.05)
    assert a1 + a2 == pytest.approx(1.0)


def test_face_geo_drops_corners_on_the_line():
    """Corners with value exactly 0 are neither inside nor crossings
    (CutCell.f90:905-956 test > 0 and a product < 0): the diagonal through
    two corners keeps only the single positive corner, area 0."""
    a, c = orc.face_geo(1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
    assert a == 0.0 and list(c[:2]) == [0.0, 0.0]
    # a line through one corner only still gives the exact triangle
    a, _ = orc.face_geo(1.0, 1.0, 1.0, 2.0, 0.0, 0.0, -1.5)
    ea, _, _ = clipped_face(1.0, 1.0, 1.0, 2.0, -1.5)
    assert a == pytest.approx(ea)


def _plane_case(size, theta, d, stretch=False):
    I, J, K = size
    dxp = np.full(I, 1.0 / I)
    dyp = np.full(J, 1.0 / J)
    dzp = np.full(K, 1.0 / K)
    if stretch:
        dzp = 1.02 ** np.arange(K)
        dzp /= dzp.sum()
    grids = configs.staggered_spacings(dxp, dyp, dzp)
    n1, n2 = math.cos(theta), math.sin(theta)
    rng = np.random.default_rng(11)
    cells = []
    for g in grids:
        phi = n1 * g["x"] + n2 * g["y"] - d
        vof = np.empty_like(phi)
        for idx in np.ndindex(phi.shape):
            ea, _, _ = clipped_face(g["dx"][idx], g["dy"][idx], n1, n2, phi[idx])
            vof[idx] = ea
        c = dict(dx=g["dx"], dy=g["dy"], dz=g["dz"], phi=np.asfortranarray(phi),
                 nx=np.full(phi.shape, n1, order="F"), ny=np.full(phi.shape, n2, order="F"),
                 nz=np.zeros(phi.shape, order="F"), vof=np.asfortranarray(vof),
                 vof_l=np.asfortranarray(0.5 * vof),
                 cell_type=np.zeros(phi.shape, np.int32, order="F"),
                 delh=np.asfortranarray(rng.uniform(size=phi.shape)))
        for m in ("fce", "fcn", "fct"):
            c[m] = np.asfortranarray(rng.uniform(-1, 1, (I + 2, J + 2, K + 2, 3)))
        cells.append(c)
    return cells, grids, (n1, n2)


def _segment_length(a, b, n1, n2, c):
    """Length of {n1 x + n2 y + c = 0} inside [-a,a] x [-b,b] (Liang-Barsky)."""
    p0 = np.array([-c * n1, -c * n2])     # closest point to the origin (|n| = 1)
    dvec = np.array([-n2, n1])
    t0, t1 = -1e300, 1e300
    for k, lim in ((0, a), (1, b)):
        if dvec[k] == 0.0:
            if abs(p0[k]) > lim:
                return 0.0
            continue
        ta, tb = (-lim - p0[k]) / dvec[k], (lim - p0[k]) / dvec[k]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return max(0.0, t1 - t0)


@pytest.mark.parametrize("theta,d", [(0.3, 0.43), (2.1, -0.2), (4.0, -0.61)])
def test_planar_solid_geometry(theta, d):
    size = (9, 8, 5)
    cells, grids, (n1, n2) = _plane_case(size, theta, d)
    inp = [dict(c) for c in cells]
    P, U, V, W = orc.grid_preprocess(cells)
    I, J, K = size
    # DefineCell classes
    for c0, o in zip(inp, (P, U, V, W)):
        vof = c0["vof"]
        assert np.all(o["cell_type"][vof >= 1 - 1e-3] == 0)
        assert np.all(o["vof"][vof >= 1 - 1e-3] == 1.0)
        assert np.array_equal(o["vof_l"][vof >= 1 - 1e-3],
                              c0["vof_l"][vof >= 1 - 1e-3] / vof[vof >= 1 - 1e-3])
        assert np.all(o["cell_type"][vof <= 1e-3] == 2) and np.all(o["vof"][vof <= 1e-3] == 0)
        mid = (vof > 1e-3) & (vof < 1 - 1e-3)
        assert np.all(o["cell_type"][mid] == 1)
    g = grids[0]
    typ = P["cell_type"]
    for i in range(1, I):
        for j in range(1, J + 1):
            for k in range(1, K + 1):
                q = (i - 1, j - 1, k - 1)
                # the face between P cells i and i+1, clipped directly at the face
                if typ[q] != 2 and typ[i, j - 1, k - 1] != 2:
                    phi = inp[0]["phi"][q]
                    ea, ey, _ = clipped_face(g["dy"][q], g["dz"][q], n2, 0.0,
                                             phi + n1 * g["dx"][q] / 2)
                    assert P["ee_area"][i, j, k] == pytest.approx(ea, abs=1e-12)
                    if ea > 1e-6:
                        assert P["fce"][i, j, k, 1] == pytest.approx(ey, abs=1e-12)
                else:
                    assert P["ee_area"][i, j, k] == 0.0
    # WlLh of cut P cells (no zero face next to them) = segment length * dz
    for idx in zip(*np.nonzero(typ == 1)):
        i, j, k = idx
        faces = [P["ee_area"][i, j + 1, k + 1], P["ee_area"][i + 1, j + 1, k + 1],
                 P["ne_area"][i + 1, j, k + 1], P["ne_area"][i + 1, j + 1, k + 1]]
        if min(faces) == 0.0:
            continue
        c = inp[0]["phi"][idx]
        L = _segment_length(g["dx"][idx] / 2, g["dy"][idx] / 2, n1, n2, c)
        assert P["wllh"][idx] == pytest.approx(L * g["dz"][idx], rel=1e-9, abs=1e-14)
    # U face through the centre of a cut P cell (i,j,k): UCell%EEArea(i-1,j,k)
    for idx in zip(*np.nonzero((inp[0]["vof"] > 1e-3) & (inp[0]["vof"] < 1 - 1e-3))):
        i, j, k = idx
        phi = inp[0]["phi"][idx]
        ea, _, _ = clipped_face(g["dy"][idx], g["dz"][idx], n2, 0.0, phi)
        assert U["ee_area"][i, j + 1, k + 1] == pytest.approx(ea, abs=1e-12)
        ea, _, _ = clipped_face(g["dz"][idx], g["dx"][idx], 0.0, n1, phi)
        assert V["ne_area"][i + 1, j, k + 1] == pytest.approx(ea, abs=1e-12)
    # faces next to solid cells are zero on every grid
    for o in (P, U, V, W):
        for idx in zip(*np.nonzero(o["cell_type"] == 2)):
            i, j, k = (x + 1 for x in idx)
            for a, s in (("ee_area", (i - 1, j, k)), ("ee_area", (i, j, k)),
                         ("ne_area", (i, j - 1, k)), ("ne_area", (i, j, k)),
                         ("te_area", (i, j, k - 1)), ("te_area", (i, j, k))):
                assert o[a][s] == 0.0


def test_demotion_only_on_p_and_small_faces():
    cells = configs.geometry_case((10, 9, 8), "bodies", seed=4)
    P, U, V, W = orc.grid_preprocess(cells)
    I, J, K = 10, 9, 8
    for o in (P,):
        for i in range(2, I + 1):
            for j in range(2, J + 1):
                for k in range(2, K + 1):
                    mf = max(o["ee_area"][i - 1, j, k], o["ee_area"][i, j, k],
                             o["ne_area"][i, j - 1, k], o["ne_area"][i, j, k],
                             o["te_area"][i, j, k - 1], o["te_area"][i, j, k])
                    if o["cell_type"][i - 1, j - 1, k - 1] != 2:
                        assert mf >= 1e-2
    # U/V/W keep their DefineCell class
    for c, o in zip(cells[1:], (U, V, W)):
        v = c["vof"]
        assert np.all(o["cell_type"][(v > 1e-3) & (v < 1 - 1e-3)] == 1)


def test_stretched_face_centre_rule():
    size = (6, 5, 9)
    cells, grids, _ = _plane_case(size, 0.3, 0.43, stretch=True)
    P, U, V, W = orc.grid_preprocess(cells)
    I, J, K = size
    wt = W["cell_type"]
    for i in range(1, I + 1):
        for j in range(1, J + 1):
            for k in range(1, K):
                q = (i - 1, j - 1, k - 1)
                if wt[q] != 1 and wt[i - 1, j - 1, k] != 1:
                    # dz differs between k and k+1 everywhere on this grid
                    assert W["fce"][i, j, k, 2] == 0.0 and W["fcn"][i, j, k, 2] == 0.0
            assert W["fce"][i, j, K, 2] == 0.0 and W["fcn"][i, j, K, 2] == 0.0
    # the last U cell of each line has zero transverse centre components
    assert np.all(U["fcn"][I, 1:J + 1, 1:K + 1, 0] == 0.0)
    assert np.all(U["fct"][I, 1:J + 1, 1:K + 1, 0] == 0.0)


def test_fct_second_component_uses_bottom_first_component():
    """CutCell.f90:541-543: FCT(i,j,k-1,2) sums CeB(i,j,k,1), not CeB(..,2).
    Checked on the U grid, whose FCT GridPreprocess leaves as CellGeoCal
    wrote it (the P-grid FCT is overwritten by :225-323)."""
    cells = configs.geometry_case((5, 4, 6), "sphere", seed=2)
    P = orc.grid_preprocess(cells)[1]
    c = cells[1]
    hit = 0
    for i in range(1, 6):
        for j in range(1, 5):
            for k in range(1, 6):
                q, qp = (i - 1, j - 1, k - 1), (i - 1, j - 1, k)
                if P["cell_type"][q] == 2 or P["cell_type"][qp] == 2:
                    continue
                args = lambda s, r: (c["dx"][s], c["dy"][s], c["nx"][s], c["ny"][s],
                                     c["nz"][s], r * c["dz"][s] / 2, c["phi"][s])
                te, ft = orc.face_geo(*args(q, 1))
                be, fb = orc.face_geo(*args(qp, -1))
                den = be + te + 1e-20
                exp1 = (fb[0] * be + ft[1] * te) / den
                assert P["fct"][i, j, k, 1] == exp1
                if abs(fb[0] - fb[1]) > 1e-9 and be > 1e-3:
                    hit += 1
    assert hit > 0


def test_define_moex_and_numbering_restated_over_arrays():
    """DefineMomentumExchangeCell and NumberExternalCell against vectorised
    restatements: shifted-array masks, and a cumulative count of the
    non-solid cells in C order of the (i, j, k) array (k fastest)."""
    rng = np.random.default_rng(3)
    I, J, K = 7, 6, 5
    cts = [np.asfortranarray(rng.integers(0, 3, (I, J, K)).astype(np.int32)) for _ in range(4)]
    ms = [np.asfortranarray(rng.integers(0, 2, (I, J, K)).astype(np.int32)) for _ in range(3)]
    got = orc.define_moex(*cts, *ms)
    ctp = cts[0]
    for ax, (ct, m0, g) in enumerate(zip(cts[1:], ms, got)):
        solid = ctp == 2
        nb = np.zeros_like(solid)
        sl = [slice(None)] * 3
        sl2 = [slice(None)] * 3
        sl[ax] = slice(0, -1)
        sl2[ax] = slice(1, None)
        nb[tuple(sl)] = solid[tuple(sl2)]
        exp = np.where((ct == 1) & (solid | nb), 1, m0)
        assert np.array_equal(g, exp), ax
    for iu, iv, iw in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)):
        ct = cts[iu + 2 * iv + 3 * iw]
        pos, ext = orc.number_external(ct, iu, iv, iw)
        sub = ct[:I - iu, :J - iv, :K - iw]
        flat = (sub != 2).ravel(order="C")
        num = np.where(flat, np.cumsum(flat) - 1, -1).reshape(sub.shape, order="C")
        assert np.array_equal(pos, num)
        assert ext == int(flat.sum()) - 1
    pos, ext = orc.number_external(np.full((3, 3, 3), 2, np.int32), 0, 0, 0)
    assert ext == -1 and np.all(pos == -1)


def _after_preprocess(cells):
    out = orc.grid_preprocess(cells)
    I, J, K = cells[0]["dx"].shape
    interp = [dict(moex=np.zeros((I, J, K), np.int32, order="F")) for _ in range(3)]
    return out, interp


def test_new_cell_face_without_solid_closed_forms():
    """No solid on a stretched grid: U/V/W centres stay at their cells'
    centres, every face interpolates at Eta = the fraction of the distance
    between the two centres where the face lies, with Al = 1, DAl = 0."""
    size = (6, 5, 7)
    cells, grids, _ = _plane_case(size, 0.0, -10.0, stretch=True)   # solid far outside
    out, interp = _after_preprocess(cells)
    cs, xs = orc.new_cell_face(out, interp)
    I, J, K = size
    for m in range(3):
        assert np.all(cs[1 + m]["cell_cent"] == 0.0)
        assert np.array_equal(cs[1 + m]["delh"], np.abs(cells[1 + m]["phi"]))
    # the W grid along z: face k of W sits on P centre k+1, W centres at P faces
    W = grids[3]
    for k in range(1, K):
        dzu = W["z"][0, 0, k] - W["z"][0, 0, k - 1]
        eta = cells[0]["dz"][0, 0, k] / 2.0 / dzu   # the face is half a P cell ahead
        assert xs[2]["szt"][3, 3, k] == pytest.approx(dzu, rel=1e-14)
        assert xs[2]["etat"][3, 3, k] == pytest.approx(eta, rel=1e-12)
    for x in xs:
        for n in ("ale", "aln", "alt"):
            assert np.all(x[n][1:I, 1:J, 1:K] == 1.0)
        for n in ("dale", "daln", "dalt"):
            assert np.all(x[n][1:I, 1:J, 1:K] == 0.0)
    # boundary copies: plane 0 equals plane 1, the last entry the one before
    assert np.array_equal(xs[0]["sxe"][0], xs[0]["sxe"][1])
    assert np.array_equal(xs[0]["sxe"][I, 1:J + 1, 1:K + 1], xs[0]["sxe"][I - 1, 1:J + 1, 1:K + 1])


@pytest.mark.parametrize("theta,d", [(0.4, 0.52), (2.5, -0.3)])
def test_new_cell_face_interpolation_in_global_frame(theta, d):
    """Planar solid (fluid where n.x > d): Al is the ratio of the distances
    of the face centroid and of the interpolation point to the plane, and
    DAl the transverse offset of the two centroids along the normal over
    S * distance, all computed from global coordinates."""
    size = (10, 9, 4)
    cells, grids, (n1, n2) = _plane_case(size, theta, d)
    out, interp = _after_preprocess(cells)
    cs, xs = orc.new_cell_face(out, interp)
    I, J, K = size
    checked = 0
    for m in range(3):  # U, V, W
        g, c, x = grids[1 + m], cs[1 + m], xs[m]
        for i in range(1, I):
            for j in range(1, J + 1):
                for k in range(1, K + 1):
                    q, qn = (i - 1, j - 1, k - 1), (i, j - 1, k - 1)
                    if c["cell_type"][q] == 2 or c["cell_type"][qn] == 2:
                        continue
                    p0 = np.array([g["x"][q], g["y"][q], g["z"][q]]) + c["cell_cent"][q]
                    p1 = np.array([g["x"][qn], g["y"][qn], g["z"][qn]]) + c["cell_cent"][qn]
                    S = p1[0] - p0[0]
                    assert x["sxe"][i, j, k] == pytest.approx(S, rel=1e-12)
                    eta = abs((c["fce"][i, j, k, 0] - c["cell_cent"][q][0]) / S)
                    if eta > 1.0:
                        continue
                    xf = (1 - eta) * p0 + eta * p1
                    dist = lambda p: n1 * p[0] + n2 * p[1] - d
                    sy, sz = p1[1] - p0[1], p1[2] - p0[2]
                    if not (abs(sy) > 1e-4 * g["dy"][q] or abs(sz) > 1e-4 * g["dz"][q]):
                        assert x["ale"][i, j, k] == 1.0 and x["dale"][i, j, k] == 0.0
                        continue
                    dh = dist(xf)
                    fc = np.array([g["x"][q], g["y"][q], g["z"][q]]) + c["fce"][i, j, k]
                    dhec = dist(fc)
                    if dh < 1e-3 and c["ee_area"][i, j, k] > 0.0:
                        dh = dhec
                    if abs(dh) < 1e-6:
                        continue
                    assert x["ale"][i, j, k] == pytest.approx(abs(dhec) / abs(dh), rel=1e-8)
                    assert x["dale"][i, j, k] == pytest.approx(sy * n2 / (S * abs(dh)),
                                                               rel=1e-8, abs=1e-10)
                    checked += 1
    assert checked > 10


def test_cell_linking_chains_follow_the_loop_order():
    """CellLinking in the reference's order: a minus-direction link reads
    the master's already linked centre, a plus-direction link the master's
    centre before linking."""
    size = (8, 5, 5)
    cells = configs.geometry_case(size, "sphere", seed=1)
    out, interp = _after_preprocess(cells)
    U = out[1]
    j, k = 1, 1   # 0-based (j, k) = Fortran (2, 2)
    for i in (2, 3, 4):         # Fortran i = 3, 4, 5: normal -x, chained
        U["nx"][i, j, k], U["ny"][i, j, k], U["nz"][i, j, k] = -1.0, 0.0, 0.0
        interp[0]["moex"][i, j, k] = 1
    for i in (5, 6):            # Fortran i = 6, 7: normal +x
        U["nx"][i, j, k], U["ny"][i, j, k], U["nz"][i, j, k] = 1.0, 0.0, 0.0
        interp[0]["moex"][i, j, k] = 1
    base, _ = orc.new_cell_face(out, [dict(moex=np.zeros(size, np.int32, order="F"))] * 3)
    orig = base[1]["cell_cent"]
    cs, xs = orc.new_cell_face(out, interp)
    cent, dx = cs[1]["cell_cent"], U["dx"]
    tol = 1e-8
    prev = orig[1, j, k].copy()
    for i in (2, 3, 4):
        exp = prev.copy()
        exp[0] = prev[0] - dx[i - 1, j, k] + tol * dx[i - 1, j, k]
        assert np.array_equal(cent[i, j, k], exp), i
        assert list(xs[0]["msce"][i, j, k]) == [i, j + 1, k + 1]   # Fortran i-1, j, k
        prev = exp
    for i in (5, 6):
        exp = orig[i + 1, j, k].copy()
        exp[0] = orig[i + 1, j, k][0] + dx[i + 1, j, k] - tol * dx[i + 1, j, k]
        assert np.array_equal(cent[i, j, k], exp), i
        assert list(xs[0]["msce"][i, j, k]) == [i + 2, j + 1, k + 1]
