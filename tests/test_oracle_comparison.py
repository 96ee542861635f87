"""CPU pins of the experimental-comparison oracle (oracle/comparison.py) and of the
host planning of the device version (pad column map), by closed forms:

* bilinear interpolation reproduces linear functions exactly, so linear_polar of
  img[i, j] = a i + b j + c equals a ys + b xs + c at every sample inside the image
  and the constant value outside, with ys = o0 + rs sin(ts), xs = o1 + rs cos(ts);
  polar_linear of P[t, r] = a t + b r + c likewise at its sample positions;
* add_pad on an image whose columns are numbered duplicates exactly the columns the
  reference's repeated np.insert duplicates (hand-derived case), and the product's
  column map gives the same image;
* griddata 'linear' reproduces a linear function inside the hull and is NaN
  outside; a masked source pixel masks every target point whose triangle touches it;
* optimize_scale_offset recovers an exact affine relation;
* slabmaker on a two-atom cubic cell keeps exactly the translates of the centred box,
  cell atom fastest, then x, y, z copies (half-integer coordinates, exact); on a
  two-element hexagonal cell its output equals, as a set, a brute-force enumeration of
  the translates inside the box, with zero minima on every axis.
"""

import numpy as np
import pytest

from oracle import comparison as C
from giwaxsim_amd.tools.comparison import pad_columns


def test_linear_polar_of_linear_image_is_exact():
    h, w = 9, 12
    i, j = np.mgrid[:h, :w].astype(float)
    a, b, c = 0.7, -1.3, 5.0
    img = a * i + b * j + c
    o, r = (4, 5), 6.0
    out = C.linear_polar(img, o=o, r=r, cont=-99.0)
    oh, ow = out.shape
    assert (oh, ow) == (int(round(r * 2 * np.pi)), int(round(r)))
    rs = np.linspace(0, r, ow)
    ts = np.linspace(0, 2 * np.pi, oh)
    xs = rs * np.cos(ts[:, None]) + o[1]
    ys = rs * np.sin(ts[:, None]) + o[0]
    inside = (ys >= 0) & (ys <= h - 1) & (xs >= 0) & (xs <= w - 1)
    want = np.where(inside, a * ys + b * xs + c, -99.0)
    assert inside.any() and (~inside).any()
    assert np.allclose(out, want, rtol=0, atol=1e-12)


def test_polar_linear_of_linear_polar_image_is_exact():
    ph, pw = 40, 15
    t, rr = np.mgrid[:ph, :pw].astype(float)
    P = 0.25 * t - 0.5 * rr + 2.0
    out = C.polar_linear(P, o=(6, 7), r=None, output=(13, 16))
    ys, xs = np.mgrid[:13, :16] - np.array([6, 7])[:, None, None]
    rs = np.sqrt(ys ** 2 + xs ** 2)
    ts = np.arctan2(ys, xs)
    ts[ts < 0] += 2 * np.pi
    rs = rs / pw * (pw - 1)
    ts = ts / (2 * np.pi) * (ph - 1)
    inside = (ts <= ph - 1) & (rs <= pw - 1)
    assert np.allclose(out, np.where(inside, 0.25 * ts - 0.5 * rs + 2.0, 0.0), rtol=0, atol=1e-12)


def test_add_pad_column_duplication_by_hand():
    # pixel width 0.1, pad 0.2 -> 2 columns; range 0.5..1.5 -> start 5, end 15, spacing 3:
    # copy of column 5 inserted after it, then of (shifted) column 10 = original 9
    W = 20
    r_axis = np.arange(W) * 0.1
    img = np.tile(np.arange(W, dtype=float), (3, 1))
    out = C.add_pad(img, r_axis, 0.2, (0.5, 1.5))
    want_cols = [0, 1, 2, 3, 4, 5, 5, 6, 7, 8, 9, 9, 10, 11, 12, 13, 14, 15, 16, 17]
    assert np.array_equal(out[0], np.array(want_cols, dtype=float))
    src = pad_columns(W, r_axis, 0.2, (0.5, 1.5))
    assert np.array_equal(src, want_cols)
    assert np.array_equal(img[:, src], out)
    assert pad_columns(W, r_axis, 0.01, (0.5, 1.5)) is None  # rounds to 0 pixels
    with pytest.raises(AssertionError):
        pad_columns(W, r_axis, 0.3, (0.5, 0.6))  # range too small for the pad


def test_rebin_linear_function_and_mask():
    qxy2 = np.linspace(-0.1, 2.0, 41)
    qz2 = np.linspace(-0.05, 1.5, 33)
    X2, Z2 = np.meshgrid(qxy2, qz2)
    m2 = 3.0 * X2 - 2.0 * Z2 + 0.5
    mask2 = np.zeros_like(m2)
    mask2[10, 20] = 1
    qxy1 = np.linspace(0.0, 2.1, 37)   # last columns outside the source hull
    qz1 = np.linspace(0.0, 1.4, 29)
    r, mk = C.rebin_qmap_and_mask(qxy1, qz1, m2, qxy2, qz2, mask2)
    X1, Z1 = np.meshgrid(qxy1, qz1)
    inside = X1 <= 2.0
    assert np.allclose(r[inside], (3.0 * X1 - 2.0 * Z1 + 0.5)[inside], rtol=0, atol=1e-12)
    assert np.all(np.isnan(r[~inside]))
    # targets within one source cell of the masked pixel are masked, far ones are not
    near = (np.abs(X1 - qxy2[20]) < (qxy2[1] - qxy2[0])) & (np.abs(Z1 - qz2[10]) < (qz2[1] - qz2[0]))
    assert mk[near].any() and not mk[~near].any()


def test_optimize_scale_offset_exact():
    rng = np.random.default_rng(0)
    sim = rng.normal(size=(20, 30))
    ref = 2.5 * sim - 0.75
    mask = np.zeros_like(sim, dtype=int)
    mask[3:5] = 1
    ref[3:5] = 1e6  # masked pixels do not count
    s, b = C.optimize_scale_offset(sim, ref, mask)
    assert abs(s - 2.5) < 1e-12 and abs(b + 0.75) < 1e-12


def test_slabmaker_cubic_lattice_by_hand():
    """a = b = c = 1, C at the origin and O at (0.5, 0.5, 0), sizes (2.3, 3.7, 1.2):
    5, 8 and 3 copies give extents 5.5, 8.5 and 3; the buffers 1.6, 2.4 and 0.9 keep
    C at copies kx in {2, 3}, ky in {3..6}, kz in {1, 2} and O at kx in {2, 3},
    ky in {2..5}, kz in {1, 2}; minima (2, 2.5, 1).  The reference's order is cell atom
    fastest, then kx, then ky, then kz (all values exact multiples of 0.5)."""
    eye = np.eye(3)
    cell = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0]])
    out, el = C.slabmaker(cell, np.array(['C', 'O']), 2.3, 3.7, 1.2, eye[0], eye[1], eye[2])
    want, want_el = [], []
    for kz in range(4):
        for ky in range(9):
            for kx in range(6):
                if kx in (2, 3) and 3 <= ky <= 6 and kz in (1, 2):
                    want.append((kx - 2, ky - 2.5, kz - 1))
                    want_el.append('C')
                if kx in (2, 3) and 2 <= ky <= 5 and kz in (1, 2):
                    want.append((kx + 0.5 - 2, ky + 0.5 - 2.5, kz - 1))
                    want_el.append('O')
    assert np.array_equal(out, np.array(want)) and list(el) == want_el
    # without the final shift (standalone slabmaker.py): the first shift's minima are
    # (0, 0, 0), so the kept atoms sit at their lattice positions
    cut, el2 = C.slabmaker(cell, np.array(['C', 'O']), 2.3, 3.7, 1.2, eye[0], eye[1], eye[2], shift_again=False)
    assert np.array_equal(cut, np.array(want) + np.array([2.0, 2.5, 1.0])) and list(el2) == want_el


def test_slabmaker_hexagonal_cell_brute_force():
    a, c = 2.5, 6.5
    av = np.array([a, 0.0, 0.0])
    bv = np.array([a * np.cos(np.radians(120)), a * np.sin(np.radians(120)), 0.0])
    cv = np.array([0.0, 0.0, c])
    cell = np.array([[0.0, 0.0, 0.0], [0.3, 1.1, 3.2]])
    cel = np.array(['C', 'O'])
    size = (7.0, 6.0, 9.0)
    out, el = C.slabmaker(cell.copy(), cel.copy(), *size, av, bv, cv)
    assert np.all(out.min(axis=0) == 0.0) and np.all(out.max(axis=0) <= np.array(size) + 1e-9)
    nx, ny, nz = (int(np.ceil(2 * s / v)) for s, v in zip(size, (av[0], bv[1], cv[2])))
    pts, els = [], []
    for i in range(2):
        for kx in range(nx + 1):
            for ky in range(ny + 1):
                for kz in range(nz + 1):
                    pts.append(cell[i] + kx * av + ky * bv + kz * cv)
                    els.append(cel[i])
    pts, els = np.array(pts), np.array(els)
    ext = pts.max(axis=0) - pts.min(axis=0)
    sh = pts - pts.min(axis=0)
    buf = (ext - np.array(size)) / 2
    keep = np.all((sh >= buf - 1e-9) & (sh <= ext - buf + 1e-9), axis=1)
    want = sh[keep] - sh[keep].min(axis=0)
    assert len(out) == int(keep.sum()) > 0

    def canon(p, e):
        order = np.lexsort((np.round(p[:, 0], 6), np.round(p[:, 1], 6), np.round(p[:, 2], 6), e))
        return p[order], e[order]
    p1, e1 = canon(out, el)
    p2, e2 = canon(want, els[keep])
    assert np.array_equal(e1, e2) and np.max(np.abs(p1 - p2)) < 1e-9
