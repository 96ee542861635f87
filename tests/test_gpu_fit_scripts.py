"""GPU tests of the comparison and fit scripts on the graphite test cell
(20 x 20 x 14 A slab, N = 256, 64 px detector, psi 0..180 in 19 steps):

* simulate_GIWAXS.main with img_path (experimental branch, simulate_GIWAXS.py:123-169)
  vs the oracle pipeline (oracle voxel grid and detector, mirror, trim, griddata
  rebin, shift_peak, scale / offset): trimmed axes and rebin_map identical,
  sim_comp_map to 1e-4 of the map's scale (the image carries the
  FP32-transform error, 1e-5 above 1e-6 of its maximum);
* fit_slabsize.main with num_evals = 0 (direct generation: the maps of the initial
  sizes equal evaluate_fit's and the oracle comparison's) and with a short
  Nelder-Mead fit of x and z against a map made from a slab of other sizes: every
  output file written, sizes inside their bounds, and the best residual norm no
  larger than the initial one.
"""

import os
import sys

import numpy as np
import pytest

from giwaxsim_amd import engine
from oracle import comparison as C

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
QV = 2 * np.pi / (0.25 * 256) * (1 + 1e-9)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _base_config(out):
    return {'input_filepath': f'{GOLDEN}/graphite_medium.xyz', 'a': '2.456', 'b': '2.456', 'c': '6.696',
            'alpha': '90', 'beta': '90', 'gamma': '120', 'r_voxel_size': '0.25', 'q_voxel_size': repr(QV),
            'max_q': '2.0', 'energy': '12700', 'fill_bkg': 'True', 'smooth': '2', 'num_pixels': '64',
            'angle_init_val1': '0.15', 'angle_init_ax1': 'psi', 'psi_start': '0', 'psi_end': '180',
            'psi_num': '19', 'phi_start': '0', 'phi_end': '0', 'phi_num': '1', 'theta_start': '0',
            'theta_end': '0', 'theta_num': '1', 'mirror': 'True', 'save_folder': str(out)}


def _exp_files(tmp_path, seed=0):
    """Experimental map on its own grid: qxy from -0.5 (an exact 0 column) to 1.9."""
    rng = np.random.default_rng(seed)
    qxy = np.arange(-20, 77) * 0.025
    qz = np.linspace(0.0, 1.85, 80)
    X, Z = np.meshgrid(qxy, qz)
    img = 1e-4 * (1.5 + np.cos(6 * np.hypot(X, Z))) + 1e-6 * rng.random(X.shape)
    img[:, 22:25] = 0.0   # a detector gap on the positive side, filled from the mirror
    mask = np.zeros((len(qz), int(np.sum(qxy > 0)) + 1))
    mask[:, :5] = 1
    paths = {k: str(tmp_path / f'{k}') for k in ('img.npy', 'mask.npy', 'qxy.txt', 'qz.txt')}
    np.save(paths['img.npy'], img)
    np.save(paths['mask.npy'], mask)
    np.savetxt(paths['qxy.txt'], qxy)
    np.savetxt(paths['qz.txt'], qz)
    return paths, img, qxy, qz, mask


def _oracle_image(size):
    import oracle as O
    from giwaxsim_amd.tools.comparison import slabmaker_fitting
    coords, el = slabmaker_fitting(f'{GOLDEN}/graphite_medium.xyz', *size, 2.456, 2.456, 6.696, 90.0, 90.0, 120.0)
    ids, ftab, _, _ = engine.element_table(el, 12700.0)
    ref = O.voxelgridmaker(coords, ftab[ids], 0.25, QV, 2.0, fill_bkg=True, smooth=2)
    one = np.array([0.0])
    return O.detectormaker(ref['iq'], ref['qx'], ref['qy'], ref['qz'], 64, 2.0, (0.15, 0, 0),
                           ('psi', 'None', 'None'), np.linspace(0, 180, 19), None, one, None, one, None,
                           mirror=True)


def _oracle_maps(d, h, v, img, qxy, qz, mask, pad, fit_scale_offset, zero_sim_masked):
    from giwaxsim_amd.tools.comparison import mirror_qmap_positive_qxy_only
    e, q2, z2 = mirror_qmap_positive_qxy_only(img.copy(), qxy, qz)
    sim, ht, vt = C.trim_sim_data(d, h, v, q2, z2)
    rm, rk = C.rebin_qmap_and_mask(ht, vt, e, q2, z2, mask)
    rm[rk == 1] = 0
    if zero_sim_masked:
        sim[rk == 1] = 0
    sp = C.shift_peak(sim.copy(), ht, vt, *pad)
    if fit_scale_offset:
        s, b = C.optimize_scale_offset(sp, rm, rk)
        sp = s * sp + b
        sp[rk == 1] = 0
    return rm, sp, ht, vt


def test_simulate_experimental_branch(tmp_path):
    from giwaxsim_amd import simulate_GIWAXS as S
    paths, img, qxy, qz, mask = _exp_files(tmp_path)
    cfg = _base_config(tmp_path / 'out')
    cfg.update({'x_size': '20', 'y_size': '20', 'z_size': '14', 'img_path': paths['img.npy'],
                'mask_path': paths['mask.npy'], 'qxy_path': paths['qxy.txt'], 'qz_path': paths['qz.txt'],
                'fit_scale_offset': 'True', 'pad_width': '0.05', 'pad_range_min': '1.0', 'pad_range_max': '1.6'})
    _say("[gpu-fit] simulate_GIWAXS experimental branch, graphite 20 x 20 x 14 A")
    S.main(cfg)
    sub = tmp_path / 'out' / 'graphite_medium'
    got = {k: np.load(sub / f'{k}.npy', allow_pickle=False) for k in ('det_h', 'det_v', 'rebin_map', 'sim_comp_map')}
    for f in ('fit_result.png', 'fit_result_linecut.png', 'config.txt'):
        assert (sub / f).exists(), f
    d, h, v = _oracle_image((20.0, 20.0, 14.0))
    rm, sc, ht, vt = _oracle_maps(d, h, v, img, qxy, qz, mask, (0.05, (1.0, 1.6)), True, True)
    assert np.array_equal(got['det_h'], ht) and np.array_equal(got['det_v'], vt)
    e_rm = np.max(np.abs(got['rebin_map'] - rm)) / np.max(np.abs(rm))
    e_sc = np.max(np.abs(got['sim_comp_map'] - sc)) / np.max(np.abs(rm))
    _say(f"  rebin_map {e_rm:.2e}, sim_comp_map {e_sc:.2e} of max|map|")
    assert e_rm == 0.0 and e_sc <= 1e-4


def test_fit_slabsize_main(tmp_path):
    from giwaxsim_amd import fit_slabsize as F
    from giwaxsim_amd.tools.comparison import trim_sim_data
    # "experimental" map: the oracle image of a 24 x 20 x 17 A slab on the trimmed positive grid
    d, h, v = _oracle_image((24.0, 20.0, 17.0))
    img, qxy, qz = trim_sim_data(d, h, v, np.array([0.01, 1.9]), np.array([0.0, 1.85]))
    paths = {k: str(tmp_path / k) for k in ('img.npy', 'mask.npy', 'qxy.txt', 'qz.txt')}
    np.save(paths['img.npy'], img)
    mask = np.zeros(img.shape)
    mask[:, :2] = 1
    np.save(paths['mask.npy'], mask)
    np.savetxt(paths['qxy.txt'], qxy)
    np.savetxt(paths['qz.txt'], qz)
    cfg = _base_config(tmp_path / 'direct')
    cfg.update({'x_size_init': '20', 'y_size_init': '20', 'z_size_init': '14', 'img_path': paths['img.npy'],
                'mask_path': paths['mask.npy'], 'qxy_path': paths['qxy.txt'], 'qz_path': paths['qz.txt'],
                'pad_width': '0.05', 'pad_range_min': '1.0', 'pad_range_max': '1.6', 'num_evals': '0'})
    _say("[gpu-fit] fit_slabsize direct generation 20 x 20 x 14 A")
    res0 = F.main(dict(cfg))
    d0, h0, v0 = _oracle_image((20.0, 20.0, 14.0))
    rm, sc, ht, vt = _oracle_maps(d0, h0, v0, img, qxy, qz, mask, (0.05, (1.0, 1.6)), True, False)
    assert np.array_equal(res0['det_h'], ht) and np.array_equal(res0['det_v'], vt)
    e_sc = np.max(np.abs(res0['sim_comp_map'] - sc)) / np.max(np.abs(rm))
    _say(f"  sim_comp_map {e_sc:.2e} of max|map|")
    assert e_sc <= 1e-4
    for f in ('det_h.npy', 'det_v.npy', 'rebin_map.npy', 'sim_comp_map.npy', 'fit_result.png',
              'fit_result_linecut.png', 'fit_config.txt'):
        assert (tmp_path / 'direct' / f).exists(), f

    cfg.update({'save_folder': str(tmp_path / 'fit'), 'num_evals': '8', 'fit_x': 'True', 'fit_z': 'True',
                'x_size_init': '30', 'y_size_init': '26', 'z_size_init': '30'})
    _say("[gpu-fit] fit_slabsize Nelder-Mead, 8 evaluations of x and z from 30 x 26 x 30 A")
    res = F.main(cfg)
    r, log = res['result'], res['logger']
    x, y, z = res['best_params']
    _say(f"  best sizes {x:.3f} {y:.3f} {z:.3f}; chi2 {r.chisqr:.4e} (initial {log.residual_norms[0] ** 2:.4e}); "
         f"nfev {r.nfev}")
    assert 25 <= x <= 90 and 25 <= z <= 90 and y == 26.0
    assert r.nfev == 8 and log.iterations == list(range(1, 9))
    assert abs(r.chisqr - min(log.residual_norms) ** 2) <= 1e-12 * r.chisqr
    # chisqr is sum(r * r), the logged norms np.linalg.norm(r): equal to rounding
    assert r.chisqr <= log.residual_norms[0] ** 2 * (1 + 1e-12)
    for f in ('fit_result.txt', 'fit_residuals.png', 'fit_x_size_evolution.png', 'fit_z_size_evolution.png',
              'fit_result.png', 'fit_result_linecut.png', 'fit_config.txt'):
        assert (tmp_path / 'fit' / f).exists(), f
