"""Pins of the oracle's geometry: driver sizes (tools/comparison.py:705-731),
crop (tools/voxelgrids.py:16-48), rectangle path lengths
(tools/voxelgrids.py:238-309), rotations (tools/utilities.py:222-245,
303-317) and the host planner's agreement with the oracle."""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from giwaxsim_amd import engine

# (r, q_voxel, max_q) -> N, q_num, frames, crop edge, computed independently of
# this code from the reference's formulas (tools/comparison.py:705-731 for the grid
# and frames, tools/voxelgrids.py:16-48 for the crop).
SIZES = [
    ((0.25, 2 * np.pi / (0.25 * 512) * (1 + 1e-9), 2.5), (512, 145, 453, 103)),
    ((0.2, 2 * np.pi / (0.2 * 1024) * (1 + 1e-9), 2.0), (1024, 185, 581, 131)),
    ((0.15, 2 * np.pi / (0.15 * 2048) * (1 + 1e-9), 3.0), (2048, 417, 1307, 295)),
    ((0.3, 0.02, 2.0), (1048, 285, 892, 203)),
]


@pytest.mark.parametrize("args,want", SIZES)
def test_plan_sizes(args, want):
    r, qv, mq = args
    coords = np.array([[0, 0, 0], [10.0, 10.0, 10.0]])
    p = O.voxelgrid_plan(coords, r, qv, mq)
    small, ax, _, _ = O.downselect_voxelgrid(np.zeros((p['q_num'],) * 3), p['q_axis'], p['q_axis'],
                                            p['q_axis'], mq)
    assert (p['grid_size'], p['q_num'], len(p['phis']), len(ax)) == want
    # the product planner agrees exactly
    g = engine.qgrid(qv, mq)
    assert np.array_equal(g.q_axis, p['q_axis'])
    assert np.array_equal(g.phis, p['phis'])
    assert g.c_n == len(ax) and np.array_equal(g.crop_axis, ax)
    assert engine.check_physical(r, qv, mq, 10, 10, 10) == p['grid_size']


def test_plan_errors():
    coords = np.array([[0, 0, 0], [10.0, 10.0, 10.0]])
    with pytest.raises(Exception, match='non-physical'):
        O.voxelgrid_plan(coords, 1.0, 0.05, 5.0)
    with pytest.raises(Exception, match='non-physical'):
        engine.check_physical(1.0, 0.05, 5.0, 10, 10, 10)
    big = np.array([[0, 0, 0], [500.0, 500.0, 500.0]])
    with pytest.raises(Exception, match='bounds smaller'):
        O.voxelgrid_plan(big, 0.25, 0.05, 2.0)
    with pytest.raises(Exception, match='bounds smaller'):
        engine.check_physical(0.25, 0.05, 2.0, 500, 500, 500)


def chord_length(x, hor, ver, phi_deg):
    """Independent geometry: vertical chord at x through the rectangle
    [0,hor]x[0,ver] rotated by phi (CCW) and shifted so min x = 0."""
    p = np.deg2rad(phi_deg)
    c, s = np.cos(p), np.sin(p)
    pts = np.array([[0, 0], [hor, 0], [hor, ver], [0, ver]], dtype=float)
    rot = pts @ np.array([[c, s], [-s, c]])
    rot[:, 0] -= rot[:, 0].min()
    ys = []
    for k in range(4):
        (x0, y0), (x1, y1) = rot[k], rot[(k + 1) % 4]
        if x0 == x1:
            continue
        t = (x - x0) / (x1 - x0)
        if 0 <= t <= 1:
            ys.append(y0 + t * (y1 - y0))
    return (max(ys) - min(ys)) if ys else 0.0


@pytest.mark.parametrize("phi", [0.0, 3.7, 17.2, 45.0, 63.1, 89.4, 90.0, 91.3, 120.0, 178.6])
@pytest.mark.parametrize("hor,ver", [(50.0, 50.0), (100.0, 40.0), (43.0, 176.0)])
def test_collapse_lengths_geometry(phi, hor, ver):
    x = np.linspace(0.0, 260.0, 1301)
    got = O.rectangular_collapse_lengths(x, hor, ver, phi)
    want = np.array([chord_length(v, hor, ver, phi) for v in x])
    # the reference returns 0 at x == 0 exactly; elsewhere it is the true chord
    # (boundary samples of the phi == 0 / 90 branches use strict '<')
    mask = x > 0
    assert np.allclose(got[mask], want[mask], rtol=1e-9, atol=1e-9) or phi in (0.0, 90.0)
    if phi in (0.0, 90.0):
        w = hor if phi == 0.0 else ver
        h = ver if phi == 0.0 else hor
        assert np.array_equal(got, np.where(x < w, h, 0.0))
    # area under the chord = rectangle area
    assert np.trapezoid(got, x) == pytest.approx(hor * ver, rel=5e-3)


def test_collapse_record_matches_oracle():
    """The device evaluates the branch constants from the host record with the
    same operations; emulate that evaluation in numpy and compare bit-exactly."""
    x = np.arange(2048) * 0.15
    for phi in np.linspace(0, 179.86, 97):
        for hor, ver in ((np.float64(100.0), np.float64(100.0)), (np.float64(37.5), np.float64(141.2))):
            rec = engine.collapse_record(phi, hor, ver)
            mode, H, V, s1, s2, s12, mid, A, tp, vc, tt, cp, ct = rec
            if mode == 1:
                dev = np.where(x < H, V, 0.0)
            elif mode == 2:
                dev = np.where(x < V, H, 0.0)
            else:
                with np.errstate(all='ignore'):
                    r1 = (A + x * tp) - ((vc - x) * tt)
                    r3 = (H - ((x - vc) / cp)) / ct
                dev = np.select([x == 0, x < s1, x <= s2, x < s12], [0.0, r1, mid, r3], 0.0)
            assert np.array_equal(dev, O.rectangular_collapse_lengths(x, hor, ver, phi))


def test_collapse_records_equal_per_frame_records():
    """The vectorised records of a whole phi list (what the frame plans use) equal the
    per-frame collapse_record bit for bit, on the phi lists of the voxel-grid q grids,
    a 1-degree list with 90 and 180, and random phis, for equal, unequal and random
    bounds (either order)."""
    rng = np.random.default_rng(11)
    lists = [engine.qgrid(0.02, 2.0).phis, engine.qgrid(2 * np.pi / (0.25 * 256) * (1 + 1e-9), 2.0).phis,
             np.arange(0, 181, 1.0), np.sort(rng.uniform(0, 180, 300)), np.linspace(0, 179.86, 97)]
    bounds = [(100.0, 100.0), (37.5, 141.2), (186.0, 43.0)] + [tuple(rng.uniform(5, 400, 2)) for _ in range(6)]
    for phis in lists:
        for hor, ver in bounds:
            got = engine.collapse_records(phis, hor, ver)
            want = np.array([engine.collapse_record(phi, hor, ver) for phi in phis], dtype=np.float64)
            assert got.shape == want.shape
            assert np.array_equal(got.view(np.int64), want.view(np.int64)), (hor, ver)


def test_rotate_coords_z_y():
    rng = np.random.default_rng(3)
    c = rng.uniform(-50, 50, size=(1000, 3))
    assert np.array_equal(O.rotate_coords_z_y(c, 0.0), c[:, 1])
    for phi in (13.0, 90.0, 137.25):
        ref = Rotation.from_euler('z', phi, degrees=True).apply(c)[:, 1]
        assert np.allclose(O.rotate_coords_z_y(c, phi), ref, rtol=0, atol=1e-12)


def test_rotation_matrix_vs_scipy():
    rng = np.random.default_rng(4)
    for _ in range(50):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        th = rng.uniform(-np.pi, np.pi)
        R = O.rotation_matrix(u, th)
        assert np.allclose(R, Rotation.from_rotvec(u * th).as_matrix(), atol=1e-14)
        assert np.array_equal(R, engine.rotation_matrix(u, th))


def test_frame_axes_closed_form():
    N, r = 512, 0.25
    for phi in (0.0, 33.3, 90.0, 150.0):
        hx, hy, vz, qh = O.frame_axes(N, r, phi)
        R = (N // 2 - 1) * 2 * np.pi / (N * r)
        assert qh.max() == pytest.approx(R, rel=1e-15)
        a = np.deg2rad(-phi)
        assert hy[0] == pytest.approx(-R * np.cos(a)) and hy[-1] == pytest.approx(R * np.cos(a))
        assert hx[0] == pytest.approx(R * np.sin(a), abs=1e-15)
        assert np.allclose(np.diff(hy), 2 * R * np.cos(a) / (N - 1), atol=1e-12)
    # the horizontal axis is a symmetric linspace of N points over [-R, R], not the
    # FFT grid (spacing 2R/(N-1) instead of 2 pi/(N r))
    hx, hy, vz, qh = O.frame_axes(N, r, 0.0)
    assert not np.allclose(hy, qh)
    assert np.array_equal(vz, qh)
