"""CPU tests of the host side of the comparison and fit scripts
(giwaxsim_amd.tools.comparison, giwaxsim_amd.fit_slabsize), by hand-derived cases:

* mirror_qmap_positive_qxy_only (tools/comparison.py:268-319) on a 2 x 6 map: the
  qxy > 0 columns kept, a pixel replaced by its mirrored qxy < 0 partner only where
  that one is brighter than 1.1x, the missing partner of the last column taken as 0,
  the qxy == 0 column put first;
* make_linecut (tools/comparison.py:977-989): rows with lower < chi < upper averaged,
  pixels below 0.1 ignored, q = 0 and empty columns dropped;
* rebin_weights, applied in the device kernel's order (0 + c0 v0 + c1 v1 + c2 v2),
  equals scipy's griddata 'linear' bit for bit, for values and for a 0/1 mask (whose
  "> 0" test sees the sign of rounding-level weights), on a generic sub-grid and on a
  target grid that coincides with the source grid (points on vertices and edges);
* minimize_nelder: lmfit's arcsine transform round-trips and maps the whole real line
  into [min, max]; a two-parameter quadratic residual converges to its minimum with
  a fixed third parameter untouched; the evaluation cap stops the fit and keeps the
  best parameters seen; the callback sees every evaluation in order.
"""

import numpy as np
import pytest

from giwaxsim_amd.fit_slabsize import FitLogger, _Param, minimize_nelder
from giwaxsim_amd.tools.comparison import make_linecut, mirror_qmap_positive_qxy_only, rebin_weights


def test_mirror_qmap_by_hand():
    qxy = np.array([-2.0, -1.0, 0.0, 1.0, 2.0, 3.0])
    qz = np.array([0.1, 0.2])
    qmap = np.array([[5.0, 1.0, 7.0, 2.0, 3.0, 4.0],
                     [1.0, 9.0, 8.0, 8.5, 1.0, 6.0]])
    m, q, z = mirror_qmap_positive_qxy_only(qmap.copy(), qxy, qz)
    # positive side [1, 2, 3]; mirrored negative side [-1, -2] -> [col 1, col 0], padded with 0
    # row 0: 2 vs 1 (keep 2), 3 vs 5 (3.3 < 5: take 5), 4 vs 0 (keep)
    # row 1: 8.5 vs 9 (9.35 > 9: keep), 1 vs 1 (1.1 > 1: keep), 6 vs 0
    assert np.array_equal(q, [0.0, 1.0, 2.0, 3.0])
    assert np.array_equal(m, [[7.0, 2.0, 5.0, 4.0], [8.0, 8.5, 1.0, 6.0]])
    assert z is qz
    # no qxy == 0 column: nothing prepended
    m2, q2, _ = mirror_qmap_positive_qxy_only(qmap[:, [0, 1, 3, 4, 5]].copy(), qxy[[0, 1, 3, 4, 5]], qz)
    assert np.array_equal(q2, [1.0, 2.0, 3.0]) and np.array_equal(m2, [[2.0, 5.0, 4.0], [8.5, 1.0, 6.0]])


def test_make_linecut_by_hand():
    polar = np.array([[1.0, 2.0, 0.05, 4.0],
                      [3.0, 4.0, 0.01, 8.0],
                      [100.0, 100.0, 100.0, 100.0]])
    chi = np.array([10.0, 20.0, 50.0])
    r = np.array([0.0, 0.5, 1.0, 1.5])
    qs, ints = make_linecut(polar.copy(), chi, r, 0, 30)
    # rows 0, 1; column 0 has q = 0, column 2 only has pixels below 0.1
    assert np.array_equal(qs, [0.5, 1.5]) and np.array_equal(ints, [3.0, 6.0])


def test_bounded_transform():
    p = _Param(60.0, 25.0, 180.0, True)
    assert abs(p.from_internal(p.to_internal()) - 60.0) < 1e-12
    u = np.linspace(-20, 20, 101)
    v = p.from_internal(u)
    assert np.all(v >= 25.0) and np.all(v <= 180.0)
    assert p.from_internal(np.pi / 2) == 180.0 and p.from_internal(-np.pi / 2) == 25.0


def test_minimize_nelder_quadratic():
    params = {'x_size': _Param(60.0, 25.0, 180.0, True), 'y_size': _Param(40.0, 25.0, 120.0, True),
              'z_size': _Param(33.0, 25.0, 99.0, False)}
    seen_z = set()

    def resid(v):
        seen_z.add(v['z_size'])
        return np.array([v['x_size'] - 47.0, 2.0 * (v['y_size'] - 71.0), 0.5])

    log = FitLogger()
    res = minimize_nelder(resid, params, 400, log.callback)
    assert abs(res.params['x_size'].value - 47.0) < 1e-2 and abs(res.params['y_size'].value - 71.0) < 1e-2
    assert res.params['z_size'].value == 33.0 and seen_z == {33.0}
    assert abs(res.chisqr - 0.25) < 1e-3 and res.nvarys == 2 and res.ndata == 3
    assert log.iterations == list(range(1, res.nfev + 1))
    assert not res.aborted and 'x_size' in res.report()


def test_minimize_nelder_evaluation_cap():
    params = {'x_size': _Param(60.0, 25.0, 180.0, True)}
    norms = []

    def resid(v):
        r = np.array([v['x_size'] - 47.0])
        norms.append(abs(r[0]))
        return r

    res = minimize_nelder(resid, params, 5, None)
    assert res.aborted and res.nfev == 5 and len(norms) == 5
    assert abs(res.params['x_size'].value - 47.0) == pytest.approx(min(norms))


def _csr_apply_host(row_ptr, col, val, x):
    """gx_csr_apply's arithmetic: per row, 0 + sum in stored order; NaN for empty rows."""
    out = np.full(len(row_ptr) - 1, np.nan)
    for r in range(len(row_ptr) - 1):
        a, b = row_ptr[r], row_ptr[r + 1]
        if b > a:
            acc = 0.0
            for k in range(a, b):
                acc = acc + val[k] * x[col[k]]
            out[r] = acc
    return out


@pytest.mark.parametrize('grids', ['generic', 'coincident'])
def test_rebin_weights_bit_exact_vs_griddata(grids):
    from scipy.interpolate import griddata
    rng = np.random.default_rng(3)
    if grids == 'generic':
        q2x, q2z = np.arange(-20, 77) * 0.025, np.linspace(0, 1.85, 80)
        q1x = np.linspace(-2, 2, 64)
        q1x = q1x[(q1x >= -0.4) & (q1x <= 1.9)]
        q1z = np.linspace(-2, 2, 64)
        q1z = q1z[(q1z >= -0.1) & (q1z <= 1.85)]   # some rows below the source grid: NaN
    else:
        q2x = np.linspace(0.03125, 1.90625, 30)
        q2z = np.linspace(0.0, 1.8125, 29)
        q1x, q1z = q2x, q2z
    vals = rng.random((len(q2z), len(q2x)))
    mask = np.zeros_like(vals)
    mask[:, :3] = 1
    mask[rng.random(mask.shape) < 0.02] = 1
    row_ptr, col, val, shape = rebin_weights(q1x, q1z, q2x, q2z)
    g1x, g1z = np.meshgrid(q1x, q1z)
    g2x, g2z = np.meshgrid(q2x, q2z)
    pts = np.vstack((g2x.ravel(), g2z.ravel())).T
    for v in (vals, mask):
        want = griddata(pts, v.ravel(), (g1x, g1z), method='linear').ravel()
        got = _csr_apply_host(row_ptr, col, val, v.ravel())
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok], want[ok])
    assert shape == g1x.shape
    if grids == 'generic':
        assert np.isnan(want).any()
