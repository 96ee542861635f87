"""Pins of the oracle's per-frame path: scatter (tools/voxelgrids.py:337-341),
fill/smooth (343-379), FFT + |F|^2 (381-392), binning (464-503), mean/crop/f0
(tools/comparison.py:765-788, tools/voxelgrids.py:828-857)."""

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter1d

import oracle as O
from giwaxsim_amd import engine


def test_scatter_brute_force_and_sum():
    rng = np.random.default_rng(5)
    n, N, r = 300, 32, 0.5
    coords = rng.uniform(0, 12, size=(n, 3))
    f = rng.normal(6, 1, n) + 1j * rng.normal(0, 0.1, n)
    g, st = O.frame_grid(coords, f, 37.0, N, r, 0.0, 0, 0, False, 0, return_stats=True)
    # brute force: exact per-atom rotation with scipy-free formula and python loop
    p = np.radians(37.0)
    want = np.zeros((N, N), complex)
    yr = coords[:, 0] * np.sin(p) + coords[:, 1] * np.cos(p)
    yr = yr - yr.min()
    zr = coords[:, 2] - coords[:, 2].min()
    fvalid = 0
    for k in range(n):
        yi, zi = int(yr[k] // r), int(zr[k] // r)
        if 0 <= yi < N and 0 <= zi < N:
            want[zi, yi] += f[k]
            fvalid += f[k]
    assert np.array_equal(g, want)
    assert np.isclose(g.sum(), fvalid)
    assert len(st['y_idx']) < n   # some atoms rotate out of the grid in this case


def test_single_atom_flat_and_two_atom_fringes():
    N = 64
    g = np.zeros((N, N), complex)
    g[10, 20] = 6.0 + 0.01j
    I = O.frame_iq(g)
    assert np.allclose(I, abs(6.0 + 0.01j) ** 2, rtol=1e-13)
    g[10, 20 + 8] = 6.0 + 0.01j   # second atom 8 voxels along y
    I = O.frame_iq(g)
    ky = np.fft.fftshift(np.fft.fftfreq(N)) * N
    want = 2 * abs(6.0 + 0.01j) ** 2 * (1 + np.cos(2 * np.pi * ky * 8 / N))
    assert np.allclose(I, want[None, :], rtol=1e-12, atol=1e-9)


def test_parseval():
    rng = np.random.default_rng(6)
    g = rng.normal(size=(128, 128)) + 1j * rng.normal(size=(128, 128))
    assert np.isclose(O.frame_iq(g).sum(), 128 * 128 * np.sum(np.abs(g) ** 2), rtol=1e-12)


def test_lattice_bragg_peak():
    N, r, d = 256, 0.25, 3.35
    g = np.zeros((N, N), complex)
    for k in range(18):
        g[int((k * d) // r), 100] = 6.0
    I = O.frame_iq(g)
    qv = np.fft.fftshift(np.fft.fftfreq(N, d=r) * 2 * np.pi)
    col = I[:, N // 2]
    peaks = qv[np.argsort(col)[-3:]]
    # the strongest non-DC peak sits at 2 pi / d along qz
    nz = [q for q in peaks if abs(q) > 0.5]
    assert min(abs(abs(q) - 2 * np.pi / d) for q in nz) < 2 * np.pi / (N * r)


def test_fill_regions_and_nm():
    rng = np.random.default_rng(7)
    coords = rng.uniform(0, 20, size=(400, 3))
    f = np.full(400, 6.0 + 0j)
    N, r = 128, 0.5
    xb = np.ptp(coords[:, 0]); yb = np.ptp(coords[:, 1]); zb = np.ptp(coords[:, 2])
    avg = O.avg_voxel_f_of(f, xb, yb, zb, r)
    g, st = O.frame_grid(coords, f, 30.0, N, r, avg, xb, yb, True, 0, return_stats=True)
    mv = st['max_voxels']
    A = avg * mv
    assert np.all(g[:st['z_min'], :] == A)
    assert np.all(g[st['z_max']:, :] == A)
    assert np.all(g[st['z_min']:st['z_max'] + 1, :st['y_min'] + 1] == A)
    assert np.all(g[st['z_min']:st['z_max'] + 1, st['y_max']:] == A)
    # interior: atoms + num_missing * avg with num_missing from the chord lengths
    L = O.rectangular_collapse_lengths(np.arange(N) * r, yb, xb, 30.0)
    nm = (mv - np.floor_divide(L, r)).astype(int)
    z = (st['z_min'] + st['z_max']) // 2
    y = np.arange(st['y_min'] + 1, st['y_max'])
    g0 = O.frame_grid(coords, f, 30.0, N, r, avg, xb, yb, False, 0)
    assert np.allclose(g[z, y], g0[z, y] + nm[y] * avg, rtol=1e-15, atol=1e-15)


def test_smoothing_blend():
    rng = np.random.default_rng(8)
    coords = rng.uniform(0, 20, size=(400, 3))
    f = np.full(400, 6.0 + 0.02j)
    N, r, sig = 128, 0.5, 3
    xb = np.ptp(coords[:, 0]); yb = np.ptp(coords[:, 1]); zb = np.ptp(coords[:, 2])
    avg = O.avg_voxel_f_of(f, xb, yb, zb, r)
    g, st = O.frame_grid(coords, f, 10.0, N, r, avg, xb, yb, False, sig, return_stats=True)
    g0 = O.frame_grid(coords, f, 10.0, N, r, avg, xb, yb, False, 0)
    my = np.zeros(N); my[st['y_min'] + sig:st['y_max'] - sig] = 1
    mz = np.zeros(N); mz[st['z_min'] + sig:st['z_max'] - sig] = 1
    # textbook Gaussian (radius 4 sigma, normalised), circular convolution
    rad = int(4 * sig + 0.5)
    w = np.exp(-0.5 * (np.arange(-rad, rad + 1) / sig) ** 2)
    w /= w.sum()
    sy = sum(w[k + rad] * np.roll(my, -k) for k in range(-rad, rad + 1))
    sz = sum(w[k + rad] * np.roll(mz, -k) for k in range(-rad, rad + 1))
    m = np.outer(sz, sy)
    want = g0 * m + avg * st['max_voxels'] * (1 - m)
    assert np.allclose(g, want, rtol=1e-13, atol=1e-13)
    # the product's half kernel is scipy's
    hw, hr = engine.gauss_half_kernel(sig)
    assert hr == rad and np.allclose(hw, w[rad:], rtol=1e-15)


def test_bin_frame_brute_force():
    rng = np.random.default_rng(9)
    N = 24
    q = np.linspace(-1.0, 1.0, 11)
    iq = rng.uniform(0, 10, size=(N, N))
    hx = np.linspace(-1.3, 1.3, N)
    hy = np.linspace(0.9, -0.9, N)
    vz = np.fft.fftshift(np.fft.fftfreq(N, 0.4) * 2 * np.pi) / 4
    vg = np.zeros((11,) * 3)
    vc = np.zeros((11,) * 3)
    O.bin_frame(iq, hx, hy, vz, q, q, q, vg, vc)
    wg = np.zeros((11,) * 3)
    wc = np.zeros((11,) * 3)
    dq = q[1] - q[0]
    from fractions import Fraction
    def fl(a, b):
        x = Fraction(float(a)) / Fraction(float(b))
        return x.numerator // x.denominator
    for i in range(N):
        for j in range(N):
            if -1 <= hx[j] <= 1 and -1 <= hy[j] <= 1 and -1 <= vz[i] <= 1:
                b = (fl(hy[j] - q[0], dq), fl(hx[j] - q[0], dq), fl(vz[i] - q[0], dq))
                wg[b] += iq[i, j]
                wc[b] += 1
    assert np.array_equal(vc, wc)
    assert np.allclose(vg, wg, rtol=1e-14)


def test_f0_factor():
    q = np.array([0.0, 1.0, 3.0])
    iq = np.ones((3, 3, 3))
    out = O.add_f0_q_3d(iq, q, q, q)
    # q = 0: (sum a + c)/Z = 5.9992/6
    assert out[0, 0, 0] == pytest.approx((5.9992 / 6) ** 2, rel=1e-12)
    assert np.all(np.diff(out[0, 0, :]) < 0)


def test_voxelgrid_counts_invariant():
    """sum(count) = number of in-range pixels over all frames."""
    rng = np.random.default_rng(10)
    coords = rng.uniform(0, 15, size=(200, 3))
    f = np.full(200, 6.0 + 0j)
    r, qv, mq = 0.5, 2 * np.pi / (0.5 * 64) * (1 + 1e-9), 1.5
    res = O.voxelgridmaker(coords, f, r, qv, mq, phis=None)
    p = res['plan']
    total = 0
    for phi in p['phis']:
        hx, hy, vz, _ = O.frame_axes(p['grid_size'], r, phi)
        q = p['q_axis']
        hm = (hx <= q.max()) & (hx >= q.min()) & (hy <= q.max()) & (hy >= q.min())
        vm = (vz <= q.max()) & (vz >= q.min())
        total += hm.sum() * vm.sum()
    assert res['count'].sum() == total
