"""Pins of the oracle's detector path (tools/detector.py, tools/comparison.py:790-882)
and of the host planner's rotation matrices against it."""

from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from oracle import giwaxs_oracle as OG
from giwaxsim_amd import engine

GOLDEN = __import__('os').path.join(__import__('os').path.dirname(__file__), 'golden')


def test_make_detector():
    x, y, z, h, v = O.make_detector(2.0, 5, 1.0, 3)
    assert np.array_equal(h, [-2, -1, 0, 1, 2]) and np.array_equal(v, [-1, 0, 1])
    assert np.all(x == 0) and np.array_equal(y[1], h) and np.array_equal(z[:, 2], v)


def test_rotations_closed_form():
    x, y, z, h, v = O.make_detector(2.0, 9, 2.0, 9)
    xr, yr, zr = O.rotate_about_normal(x, y, z, 90.0)         # about +x
    assert np.allclose(yr, -z, atol=1e-15) and np.allclose(zr, y, atol=1e-15)
    ph = 30.0
    xr, yr, zr = O.rotate_about_vertical(x, y, z, ph)         # about +z
    assert np.allclose(xr, -y * np.sin(np.radians(ph)), atol=1e-15)
    assert np.allclose(yr, y * np.cos(np.radians(ph)), atol=1e-15)
    assert np.array_equal(zr, z)
    th = 20.0
    xr, yr, zr = O.rotate_about_horizontal(x, y, z, th)       # about +y
    assert np.allclose(xr, z * np.sin(np.radians(th)), atol=1e-15)
    assert np.allclose(zr, z * np.cos(np.radians(th)), atol=1e-15)


def test_intersect_exact_floor():
    rng = np.random.default_rng(11)
    q = np.linspace(-2.0, 2.0, 41)
    iq = rng.uniform(size=(41, 41, 41))
    pts = rng.uniform(-2.5, 2.5, size=(3, 500))
    got = O.intersect_detector(iq, q, q, q, pts[0][None], pts[1][None], pts[2][None])[0]
    dq = Fraction(float(q[1] - q[0]))
    want = []
    for a, b, c in pts.T:
        idx = []
        for val in (a, b, c):
            t = (Fraction(float(val - q[0]))) / dq
            k = t.numerator // t.denominator
            idx.append(min(max(k, 0), 40))
        want.append(iq[idx[1], idx[0], idx[2]])
    assert np.array_equal(got, np.array(want))


def test_mirror_cases():
    q = np.arange(9.0).reshape(3, 3)
    m = O.mirror_vertical_horizontal(q)
    # odd x odd: centre row/col doubled sums, centre x4
    assert m[1, 1] == 16.0
    assert np.array_equal(m[1], [(3 + 5) * 2, (4 + 4) * 2 if False else 16.0, (5 + 3) * 2])
    assert np.array_equal(m[:, 1], [(1 + 7) * 2, 16.0, (7 + 1) * 2])
    assert m[0, 0] == 0 + 2 + 6 + 8
    q2 = np.arange(4.0).reshape(2, 2)
    assert np.array_equal(O.mirror_vertical_horizontal(q2), np.full((2, 2), 6.0))
    q3 = np.arange(6.0).reshape(2, 3)
    m3 = O.mirror_vertical_horizontal(q3)
    assert np.array_equal(m3[:, 1], [(1 + 4) * 2, (4 + 1) * 2])
    assert m3[0, 0] == 0 + 2 + 3 + 5


def test_detector_constant_iq():
    q = np.linspace(-3, 3, 31)
    iq = np.full((31, 31, 31), 7.0)
    psis = np.linspace(0, 180, 7)
    one = np.array([0.0])
    d, h, v = O.detectormaker(iq, q, q, q, 16, 2.0, (0, 0, 0), ('None',) * 3, psis, None, one,
                              None, one, None, mirror=True)
    assert np.allclose(d, 4 * 7.0 * 1e-6, rtol=1e-14)


def test_planner_matrices_match_oracle():
    """Host planner matrices (corner tracking) == oracle (full-grid rotation)."""
    n, mq = 17, 2.0
    init_vals, init_axs = (0.15, 90.0, 12.0), ('psi', 'phi', 'theta')
    psis = np.linspace(0, 180, 5)
    phis = np.linspace(0, 30, 3)
    thetas = np.linspace(0, 1, 2)
    w = lambda a: np.ones_like(a) / len(a)
    plan = engine.detector_plan(n, mq, init_vals, init_axs, psis, w(psis), phis, w(phis), thetas,
                                w(thetas))
    x, y, z, h, v = O.make_detector(mq, n, mq, n)
    x, y, z = OG._init_rotations(x, y, z, init_vals, init_axs)
    k = 0
    for ps in psis:
        Rp = OG.normal_matrix(x, y, z, ps)
        x1, y1, z1 = OG._apply(Rp, x, y, z)
        for ph in phis:
            Rv = OG.vertical_matrix(x1, y1, z1, ph)
            x2, y2, z2 = OG._apply(Rv, x1, y1, z1)
            for th in thetas:
                Rh = OG.horizontal_matrix(x2, y2, z2, th)
                assert np.array_equal(plan.orient_R[k], np.concatenate([Rp.ravel(), Rv.ravel(),
                                                                        Rh.ravel()]))
                k += 1
    assert k == len(plan.orient_w)
    assert len(plan.init_R) == 3


def test_scale_offset_closed_form():
    rng = np.random.default_rng(12)
    s = rng.uniform(0, 5, size=(40, 40))
    y = 2.5 * s + 0.75 + rng.normal(0, 0.01, s.shape)
    m = (rng.uniform(size=s.shape) < 0.2).astype(int)
    a, b = O.optimize_scale_offset(s, y, m)
    ss, yy = s[m == 0], y[m == 0]
    sc = np.sum((ss - ss.mean()) * (yy - yy.mean())) / np.sum((ss - ss.mean()) ** 2)
    assert a == pytest.approx(sc, rel=1e-12) and b == pytest.approx(yy.mean() - sc * ss.mean(), rel=1e-12)


def test_pm6_fixture_axes():
    """The reference's golden run saves det_h / det_v = linspace(-2, 2, 500)
    restricted to >= 0 (simulate_GIWAXS.py:172-180)."""
    dh = np.load(f'{GOLDEN}/pm6_det_h.npy', allow_pickle=False)
    dv = np.load(f'{GOLDEN}/pm6_det_v.npy', allow_pickle=False)
    _, _, _, h, v = O.make_detector(2.0, 500, 2.0, 500)
    assert np.array_equal(dh, h[h >= 0]) and np.array_equal(dv, v[v >= 0])
