"""CPU tests of the notebook-level map helpers of giwaxsim_amd.tools.comparison
(reference tools/comparison.py), by hand-derived cases and closed forms:

* mirror_qmap_fix_detgap (322-378) and mirror_qmap_positive_qxy_only_old (239-266) on
  a 2 x 6 map worked by hand (gap fill, reflection, the qxy = 0 column last / doubled);
* the missing wedge of mask_forbidden_pixels / create_wedge_mask (73-159): a pixel is
  masked iff |qxy| is below the in-plane distance to the Ewald circle of wavelength
  12400 / E through the origin at incidence alpha_i, i.e. b(qz) = k cos a -
  sqrt(k^2 - (qz - k sin a)^2); rows the circle does not reach mask nothing;
* rebin_and_combine_qmaps (193-237) equals scipy griddata + the percentile or position
  scale, spliced at the qxy = 0 column, bit for bit;
* find_nearest, match_qmap_index, normalize_qmap(_position), scale_offset_mask_qmap by
  hand; add_f0_q_dependence (380-404) in place, exactly (a0 + a2 + a4 + a6 + a8) / Z at
  q = 0 and the Cromer-Mann f0(q) / Z elsewhere.
"""

import math

import numpy as np
import pytest

from giwaxsim_amd.tools import comparison as C
from giwaxsim_amd.tools.ptable_dict import aff_dict, ptable

QXY = np.array([-2.0, -1.0, 0.0, 1.0, 2.0, 3.0])
QZ = np.array([0.1, 0.2])
MAP = np.array([[5.0, 1.0, 7.0, 2.0, 3.0, 4.0],
                [1.0, 9.0, 8.0, 8.5, 1.0, 6.0]])


def test_mirror_fix_detgap_by_hand():
    m, q, z = C.mirror_qmap_fix_detgap(MAP.copy(), QXY, QZ)
    # gap-filled positive side [[2, 5, 4], [8.5, 1, 6]] reflected, then the qxy = 0 column
    assert np.array_equal(q, [-3.0, -2.0, -1.0, 1.0, 2.0, 3.0, 0.0])
    assert np.array_equal(m, [[4.0, 5.0, 2.0, 2.0, 5.0, 4.0, 7.0], [6.0, 1.0, 8.5, 8.5, 1.0, 6.0, 8.0]])
    assert z is QZ
    m2, q2, _ = C.mirror_qmap_fix_detgap(MAP[:, [0, 1, 3, 4, 5]].copy(), QXY[[0, 1, 3, 4, 5]], QZ)
    assert np.array_equal(q2, [-3.0, -2.0, -1.0, 1.0, 2.0, 3.0]) and m2.shape == (2, 6)


def test_mirror_old_by_hand():
    m, q, _ = C.mirror_qmap_positive_qxy_only_old(MAP, QXY, QZ)
    assert np.array_equal(q, [-3.0, -2.0, -1.0, 0.0, 0.0, 1.0, 2.0, 3.0])
    assert np.array_equal(m, [[4.0, 3.0, 2.0, 7.0, 7.0, 2.0, 3.0, 4.0], [6.0, 1.0, 8.5, 8.0, 8.0, 8.5, 1.0, 6.0]])


@pytest.mark.parametrize('alpha, energy', [(0.15, 12700.0), (0.3, 8000.0)])
def test_missing_wedge_is_the_ewald_circle(alpha, energy):
    k = 2 * np.pi / (12400 / energy)
    a = np.deg2rad(alpha)
    qxy = np.linspace(-1.5, 1.5, 601)
    qz = np.concatenate((np.linspace(0.0, 2.5, 120), [2 * k + 1.0]))   # last row beyond the circle
    img = np.arange(len(qz) * len(qxy), dtype=np.float64).reshape(len(qz), len(qxy)) + 1.0
    masked = C.mask_forbidden_pixels(img, qxy, qz, alpha, energy)
    mask = C.create_wedge_mask(img, qxy, qz, alpha, energy)
    assert mask.dtype.kind == 'i' and set(np.unique(mask)) <= {0, 1}
    assert np.array_equal(np.isnan(masked), mask == 1)
    assert np.array_equal(masked[mask == 0], img[mask == 0]) and not np.isnan(img).any()
    for i, z in enumerate(qz):
        h = (z - k * np.sin(a)) ** 2
        if h > k ** 2:
            assert not mask[i].any()
            continue
        b = abs(k * np.cos(a) - math.sqrt(k ** 2 - h))
        clear = np.abs(np.abs(qxy) - b) > 1e-9
        assert np.array_equal(mask[i][clear] == 1, (np.abs(qxy) < b)[clear]), i
    assert mask[1:-1].any() and not mask[-1].any()


@pytest.mark.parametrize('pos', [0, (0.6, 0.9)])
def test_rebin_and_combine_vs_griddata(pos):
    from scipy.interpolate import griddata
    rng = np.random.default_rng(5)
    qxy1, qz1 = np.linspace(-1.2, 1.7, 47), np.linspace(-0.05, 1.6, 41)
    qxy2, qz2 = np.linspace(-1.0, 1.9, 60), np.linspace(0.0, 1.7, 55)
    q1 = rng.random((41, 47)) + 0.5
    q2 = rng.random((55, 60)) * 3
    got = C.rebin_and_combine_qmaps(q1, qxy1, qz1, q2, qxy2, qz2, pos=pos)
    g1x, g1z = np.meshgrid(qxy1, qz1)
    g2x, g2z = np.meshgrid(qxy2, qz2)
    rb = griddata(np.vstack((g2x.ravel(), g2z.ravel())).T, q2.ravel(), (g1x, g1z), method='linear')
    if pos:
        ix, iz = np.argmin(np.abs(qxy1 - pos[0])), np.argmin(np.abs(qz1 - pos[1]))
        rb = rb * (q1[iz, ix] / rb[iz, ix])
    else:
        rb = rb * (np.nanpercentile(q1, 15) / np.nanpercentile(rb, 75))
    zero = np.abs(qxy1).argmin()
    want = q1.copy()
    want[:, zero:] = rb[:, zero:]
    assert np.isnan(want).any()   # rows of qmap1 below qmap2's qz range
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok], want[ok]) and np.array_equal(got[:, :zero], q1[:, :zero])


def test_small_helpers_by_hand():
    assert C.find_nearest([0.0, 1.0, 2.0, 3.0], 1.6) == 2 and C.find_nearest([0.0, 1.0, 2.0], 0.5) == 0
    assert C.match_qmap_index(1.1, 0.18, QXY, QZ) == (1, 3)
    a = np.array([[1.0, np.nan], [2.0, 1.0]])
    assert np.array_equal(C.normalize_qmap(a, np.array([[8.0, 0.0], [np.nan, 0.0]])), a * 2.0, equal_nan=True)
    b = np.array([[2.0, 4.0], [6.0, 8.0]])
    assert np.array_equal(C.normalize_qmap_position(b, b * 3.0 + 1.0, np.array([0.0, 1.0]), np.array([0.0, 1.0]),
                                                    0.9, 0.1), b * (13.0 / 4.0))
    mask = np.array([[1, 0], [0, 1]])
    assert np.array_equal(C.scale_offset_mask_qmap(b, mask, scale=2.0, offset=-1.0), [[0.0, 7.0], [11.0, 0.0]])


@pytest.mark.parametrize('element', ['C', 'O', 'S'])
def test_add_f0_q_dependence(element):
    h = np.linspace(-2.0, 2.0, 9)
    v = np.linspace(0.0, 2.0, 5)
    img = np.full((5, 9), 3.0)
    out = C.add_f0_q_dependence(img, h, v, element)
    assert out is img
    a = aff_dict[element]
    Z = ptable[element]
    assert out[0, 4] == 3.0 * (((((a[0] + a[2]) + a[4]) + a[6]) + a[8]) / Z)
    for i, qz in enumerate(v):
        for j, qxy in enumerate(h):
            q2 = qxy * qxy + qz * qz
            f0 = sum(a[2 * t] * math.exp(-a[2 * t + 1] * q2 / (16 * math.pi ** 2)) for t in range(4)) + a[8]
            assert out[i, j] == pytest.approx(3.0 * f0 / Z, rel=1e-14)
