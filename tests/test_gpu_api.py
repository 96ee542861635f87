"""GPU parity of the reference-compatible helpers (tools/voxelgrids.py,
tools/utilities.py, tools/comparison.py) and of the config-file entry point
against the FP64 oracle."""

import os
from multiprocessing import shared_memory

import numpy as np
import pytest

import oracle as O
from giwaxsim_amd import engine
from giwaxsim_amd.tools import comparison as C
from giwaxsim_amd.tools import utilities as U
from giwaxsim_amd.tools import voxelgrids as V

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def test_collapse_lengths_bit_exact():
    x = np.arange(1024) * 0.2
    for phi in (0.0, 0.31, 45.0, 89.99, 90.0, 133.7):
        got = V.rectangular_collapse_lengths(x, 100.0, 37.5, phi)
        assert np.array_equal(got, O.rectangular_collapse_lengths(x, 100.0, 37.5, phi))


@pytest.mark.parametrize("element", ['C', 'H', 'O', 'S', 'Cl', 'Br', 'Na', 'Pb'])
def test_add_f0(element):
    from giwaxsim_amd.tools.ptable_dict import aff_dict, ptable
    rng = np.random.default_rng(0)
    q = np.linspace(-2.1, 2.1, 23)
    iq = rng.uniform(size=(23, 23, 23))
    want = O.add_f0_q_3d(iq, q, q, q, aff=aff_dict[element], Z=ptable[element])
    assert np.allclose(V.add_f0_q_3d(iq, q, q, q, element), want, rtol=1e-14, atol=0)


def test_avg_voxel_f_closed_form():
    """The mean f per voxel of a rectangular lattice of n_x n_y n_z atoms, spacing a:
    f n_x n_y n_z r^3 / ((n_x - 1)(n_y - 1)(n_z - 1) a^3) (unrotated bounding box)."""
    nx, ny, nz, a, r = 9, 7, 5, 2.0, 0.25
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    coords = np.stack([i.ravel() * a, j.ravel() * a, k.ravel() * a], axis=1).astype(np.float64)
    elements = np.full(len(coords), 'C')
    res = engine.run_voxelgrid(coords, elements, r, 2 * np.pi / (r * 256) * (1 + 1e-9), 2.0, 12700.0,
                               fill_bkg=True, frames=(0, 2))
    f = engine.element_table(['C'], 12700.0)[1][0]
    want = f * nx * ny * nz * r ** 3 / ((nx - 1) * (ny - 1) * (nz - 1) * a ** 3)
    assert res['avg'] == pytest.approx(want, rel=1e-13)


def _shm(shape):
    shm = shared_memory.SharedMemory(create=True, size=int(np.prod(shape)) * 8)
    a = np.ndarray(shape, dtype=np.float64, buffer=shm.buf)
    a[...] = 0
    return shm, a


def test_process_file2():
    rng = np.random.default_rng(1)
    N = 64
    q = np.linspace(-1.5, 1.5, 31)
    iq = rng.uniform(0, 100, size=(N, N))
    hx, hy, vz, _ = O.frame_axes(N, 0.5, 23.0)
    s1, g = _shm((31,) * 3)
    s2, c = _shm((31,) * 3)
    try:
        g[...] = rng.uniform(0, 1, size=g.shape)   # binning adds to what the grids hold
        c[...] = 3.0
        wg, wc = g.copy(), c.copy()
        V.process_file2(iq, hx, hy, vz, q, q, q, s1.name, s2.name)
        O.bin_frame(iq, hx, hy, vz, q, q, q, wg, wc)
        # sorted axes: each bin's pixels are added in np.add.at's order, bit for bit
        assert np.array_equal(c, wc)
        assert np.array_equal(g, wg)
        # axes in no order (a caller's own arrays): atomic adds, same counts, sums to rounding
        perm = rng.permutation(N)
        g[...] = 0.0
        c[...] = 0.0
        wg[...] = 0.0
        wc[...] = 0.0
        V.process_file2(iq[:, perm], hx[perm], hy[perm], vz, q, q, q, s1.name, s2.name)
        O.bin_frame(iq[:, perm], hx[perm], hy[perm], vz, q, q, q, wg, wc)
        assert np.array_equal(c, wc)
        assert np.allclose(g, wg, rtol=1e-14, atol=0)
    finally:
        del g, c
        for s in (s1, s2):
            s.close()
            s.unlink()


def test_frames_from_files(tmp_path):
    """process_file (tools/voxelgrids.py:418-462) equals process_file2 on the same saved
    frame, and frames_to_iq_parallel (508-533) over three saved frames equals the
    oracle's binning of each (sum / count, 0 where no pixel lands)."""
    rng = np.random.default_rng(4)
    N = 64
    q = np.linspace(-1.5, 1.5, 31)
    paths = []
    wg = np.zeros((31,) * 3)
    wc = np.zeros((31,) * 3)
    for k, phi in enumerate((0.0, 23.0, 61.5)):
        iq = rng.uniform(0, 100, size=(N, N))
        hx, hy, vz, _ = O.frame_axes(N, 0.5, phi)
        base = str(tmp_path / f'frame{k}')
        for name, arr in (('iq', iq), ('h_qx', hx), ('h_qy', hy), ('v_qz', vz)):
            np.save(f'{base}_{name}.npy', arr)
        paths.append(base)
        O.bin_frame(iq, hx, hy, vz, q, q, q, wg, wc)
    got = V.frames_to_iq_parallel(paths, 31, q, q, q)
    want = np.divide(wg, wc, out=np.zeros_like(wg), where=wc != 0)
    assert np.array_equal(got == 0, want == 0) and (want == 0).any()
    assert np.array_equal(got, want)
    s1, g1 = _shm((31,) * 3)
    s2, c1 = _shm((31,) * 3)
    s3, g2 = _shm((31,) * 3)
    s4, c2 = _shm((31,) * 3)
    try:
        V.process_file(paths[1], 31, q, q, q, s1.name, s2.name)
        V.process_file2(*(np.load(f'{paths[1]}_{n}.npy') for n in ('iq', 'h_qx', 'h_qy', 'v_qz')), q, q, q,
                        s3.name, s4.name)
        assert np.array_equal(c1, c2) and np.array_equal(g1, g2) and c1.sum() > 0
    finally:
        del g1, c1, g2, c2
        for s in (s1, s2, s3, s4):
            s.close()
            s.unlink()


@pytest.mark.parametrize("fill,smooth", [(False, 0), (True, 2)])
def test_rotate_project_fft_coords(fill, smooth):
    rng = np.random.default_rng(2)
    coords = rng.uniform(0, 20, size=(600, 3))
    f = np.where(rng.uniform(size=600) < 0.7, 6.003 + 0.002j, 8.01 + 0.006j)
    r, N = 0.25, 256
    q = engine.qgrid(2 * np.pi / (r * N) * (1 + 1e-9), 2.0).q_axis
    xb, yb, zb = [np.ptp(coords[:, i]) for i in range(3)]
    avg = O.avg_voxel_f_of(f, xb, yb, zb, r)
    phi = 31.4
    s1, g = _shm((len(q),) * 3)
    s2, c = _shm((len(q),) * 3)
    try:
        V.rotate_project_fft_coords((coords, f, phi, N, r, avg, xb, yb, zb, fill, smooth, q, q, q,
                                     s1.name, s2.name))
        wg = np.zeros_like(g)
        wc = np.zeros_like(c)
        grid = O.frame_grid(coords, f, phi, N, r, avg, xb, yb, fill, smooth)
        hx, hy, vz, _ = O.frame_axes(N, r, phi)
        O.bin_frame(O.frame_iq(grid), hx, hy, vz, q, q, q, wg, wc)
        assert np.array_equal(c, wc)
        m = wg > 1e-6 * wg.max()
        assert np.max(np.abs(g[m] - wg[m]) / wg[m]) < 1e-5
    finally:
        del g, c
        for s in (s1, s2):
            s.close()
            s.unlink()


def test_rotate_coords_z():
    rng = np.random.default_rng(3)
    c = rng.uniform(-40, 40, size=(5000, 3))
    for phi in (0.0, 17.3, 90.0):
        got = U.rotate_coords_z(c, phi)
        assert np.array_equal(got[:, 1], O.rotate_coords_z_y(c, phi))
        assert np.array_equal(got[:, 2], c[:, 2])


def test_optimize_scale_offset():
    rng = np.random.default_rng(4)
    s = rng.uniform(0, 5, size=(64, 64))
    y = 1.7 * s - 0.2 + rng.normal(0, 0.05, s.shape)
    m = (rng.uniform(size=s.shape) < 0.3).astype(int)
    a, b = C.optimize_scale_offset(s, y, m)
    a2, b2 = O.optimize_scale_offset(s, y, m)
    assert a == pytest.approx(a2, rel=1e-12) and b == pytest.approx(b2, rel=1e-10, abs=1e-12)


def test_voxelgridmaker_fitting_api():
    coords, elements = U.load_xyz(f'{GOLDEN}/graphite_medium.xyz')
    r, qv, mq, E = 0.25, 2 * np.pi / (0.25 * 256) * (1 + 1e-9), 2.0, 12700.0
    iq, qx, qy, qz = C.voxelgridmaker_fitting(coords, elements, r, qv, mq, E)
    ids, ftab, _, _ = engine.element_table(elements, E)
    ref = O.voxelgridmaker(coords, ftab[ids], r, qv, mq)
    assert np.array_equal(qx, ref['qx']) and iq.shape == ref['iq'].shape
    m = ref['iq'] > 1e-6 * ref['iq'].max()
    assert np.max(np.abs(iq[m] - ref['iq'][m]) / ref['iq'][m]) <= 1e-5


def test_simulate_giwaxs_config(tmp_path):
    from giwaxsim_amd import simulate_GIWAXS as S
    cfg = tmp_path / 'cfg.txt'
    out = tmp_path / 'out'
    cfg.write_text(
        f"input_filepath={GOLDEN}/graphite_medium.xyz\nx_size=20\ny_size=20\nz_size=14\n"
        "a=2.456\nb=2.456\nc=6.696\nalpha=90\nbeta=90\ngamma=120\n"
        f"r_voxel_size=0.25\nq_voxel_size={2 * np.pi / (0.25 * 256) * (1 + 1e-9)!r}\nmax_q=2\n"
        "energy=12700\nfill_bkg=True\nsmooth=2\nnum_pixels=64\n"
        "angle_init_val1=0.15\nangle_init_ax1=psi\n"
        "psi_start=0\npsi_end=180\npsi_num=19\nphi_start=0\nphi_end=0\nphi_num=1\n"
        "theta_start=0\ntheta_end=0\ntheta_num=1\nmirror=True\n"
        f"save_folder={out}\n")
    config = U.parse_config_file(str(cfg))
    res = S.main(config)
    det = np.load(out / 'graphite_medium' / 'det_sum.npy', allow_pickle=False)
    dh = np.load(out / 'graphite_medium' / 'det_h.npy', allow_pickle=False)
    assert det.shape == (32, 32) and np.all(dh >= 0)
    # the same pipeline through the oracle
    coords, el = C.slabmaker_fitting(f'{GOLDEN}/graphite_medium.xyz', 20.0, 20.0, 14.0, 2.456, 2.456,
                                     6.696, 90.0, 90.0, 120.0)
    ids, ftab, _, _ = engine.element_table(el, 12700.0)
    ref = O.voxelgridmaker(coords, ftab[ids], 0.25, float(config['q_voxel_size']), 2.0,
                           fill_bkg=True, smooth=2)
    psis = np.linspace(0, 180, 19)
    one = np.array([0.0])
    d, h, v = O.detectormaker(ref['iq'], ref['qx'], ref['qy'], ref['qz'], 64, 2.0, (0.15, 0, 0),
                              ('psi', 'None', 'None'), psis, None, one, None, one, None, mirror=True)
    d = d[np.ix_(np.where(v >= 0)[0], np.where(h >= 0)[0])]
    m = d > 1e-6 * d.max()
    assert np.max(np.abs(det[m] - d[m]) / d[m]) <= 1e-5


def test_detector_uses_the_array_it_is_given():
    """detectormaker_fitting gathers from the iq it is passed, not from a device copy of
    an earlier result: arrays derived from voxelgridmaker_fitting's iq (scaled, flipped,
    copied, edited in place) give the oracle's image of that same array, bit for bit."""
    coords, elements = U.load_xyz(f'{GOLDEN}/graphite_medium.xyz')
    r, qv, mq, E = 0.25, 2 * np.pi / (0.25 * 256) * (1 + 1e-9), 2.0, 12700.0
    iq, qx, qy, qz = C.voxelgridmaker_fitting(coords, elements, r, qv, mq, E)
    psis = np.linspace(0, 180, 13)
    one = np.array([0.0])
    args = (qx, qy, qz, 64, 2.0, (0.15, 0, 0), ('psi', 'None', 'None'), psis, None, one, None, one,
            None)

    def check(a, what):
        got, _, _ = C.detectormaker_fitting(a, *args, mirror=True)
        want, _, _ = O.detectormaker(np.asarray(a), *args, mirror=True)
        assert np.array_equal(got, want), what
        return got

    base = check(iq, 'as returned')
    assert not np.array_equal(check(iq * 2, 'iq * 2'), base)
    assert not np.array_equal(check(np.flip(iq, 0), 'flipped along qy'), base)
    check(iq.copy(), 'copy')
    iq[:, :, : iq.shape[2] // 2] = 0.0
    assert not np.array_equal(check(iq, 'edited in place'), base)
