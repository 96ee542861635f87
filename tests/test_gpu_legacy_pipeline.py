"""GPU tests of the legacy two-stage pipeline (giwaxsim_amd.voxelgridmaker and
giwaxsim_amd.detectormaker, reference old_modules/voxelgridmaker.py and
old_modules/detectormaker.py) on the graphite test cell (N = 256), against the oracle:

* stage 1, aff_num_qs = 1: the saved {gen_name}_iq.npy equals the oracle's voxel grid
  (frames, mean, crop, (f0_C/Z)^2 of the most common element 'C') to 1e-5 above 1e-6
  of its maximum; the saved axes are the oracle's crop axes exactly;
* stage 1, aff_num_qs = 3: the last q shell's grid (f = f0(q_shell) + f1 + i f2,
  no final f0 factor) with the shell's voxels doubled, to the same bar;
* generate_voxel_grid_low_mem (tools/voxelgrids.py:535-722) on the full q_num^3 grid:
  the oracle's sum / count (0 where no frame reaches) to the same bar, zeros in the same
  voxels, the full q axis, cropped by downselect_voxelgrid equal to the stage-1 grid bit
  for bit, and the output_dir form writing the four arrays;
* stage 2: the image read back from the handoff, summed over 19 psi orientations and
  mirrored, equals the oracle detector fed the same I(q) bit for bit (the legacy floor
  keeps 1e-6 and skips the 1e-6 rescale, so legacy * 1e-6 == oracle exactly); a second
  run writes a new {gen_name}_det_sum1 folder.
"""

import os
import sys

import numpy as np
import pytest

from giwaxsim_amd import engine

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
QV = 2 * np.pi / (0.25 * 256) * (1 + 1e-9)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _stage1_config(out, aff_num_qs):
    return {'input_filepath': f'{GOLDEN}/graphite_medium.xyz', 'gen_name': f'gr{aff_num_qs}', 'r_voxel_size': '0.25',
            'q_voxel_size': repr(QV), 'aff_num_qs': str(aff_num_qs), 'energy': '12700', 'max_q': '2.0',
            'output_dir': str(out), 'fill_bkg': 'True', 'smooth': '2'}


def _close(got, want):
    m = want > 1e-6 * want.max()
    return float(np.max(np.abs(got[m] - want[m]) / want[m]))


@pytest.mark.parametrize('aff_num_qs', [1, 3])
def test_voxelgridmaker_stage(tmp_path, aff_num_qs):
    import oracle as O
    from giwaxsim_amd import voxelgridmaker as V
    from giwaxsim_amd.tools.f1f2 import f1_f2
    from giwaxsim_amd.tools.utilities import load_xyz
    _say(f"[gpu-legacy] voxelgridmaker stage, graphite, aff_num_qs={aff_num_qs}")
    path = V.main(_stage1_config(tmp_path, aff_num_qs))
    g = f'gr{aff_num_qs}'
    iq = np.load(f'{path}/{g}_iq.npy', allow_pickle=False)
    qx = np.load(f'{path}/{g}_qx.npy', allow_pickle=False)
    assert os.path.exists(f'{path}/{g}_config.txt')
    coords, el = load_xyz(f'{GOLDEN}/graphite_medium.xyz')
    ids, ftab, _, uniq = engine.element_table(el, 12700.0)
    if aff_num_qs == 1:
        ref = O.voxelgridmaker(coords, ftab[ids], 0.25, QV, 2.0, fill_bkg=True, smooth=2)
        want = ref['iq']
    else:
        grid = engine.qgrid(float(QV), 2.0)
        step = grid.max_q_diag / 3
        q_val = 0.5 * step + 2 * step
        f_shell = np.array([V.shell_f0(q_val, str(e)) + complex(*f1_f2(str(e), 12700.0)) for e in uniq])
        ref = O.voxelgridmaker(coords, f_shell[ids], 0.25, QV, 2.0, fill_bkg=True, smooth=2)
        mean = np.divide(ref['sum'], ref['count'], out=np.zeros_like(ref['sum']), where=ref['count'] != 0)
        small, ax, _, _ = O.downselect_voxelgrid(mean, ref['plan']['q_axis'], ref['plan']['q_axis'],
                                                 ref['plan']['q_axis'], 2.0)
        X, Y, Z = np.meshgrid(ax, ax, ax)
        qr = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
        shell = (qr <= q_val + step / 2) & (qr > q_val - step / 2)
        assert shell.any()
        want = small + np.where(shell, small, 0)
    assert np.array_equal(qx, ref['qx'])
    err = _close(iq, want)
    _say(f"  iq {iq.shape}: max rel err {err:.2e}")
    assert iq.shape == want.shape and err <= 1e-5


def test_detectormaker_stage(tmp_path):
    import oracle as O
    from giwaxsim_amd import detectormaker as D
    from giwaxsim_amd import voxelgridmaker as V
    _say("[gpu-legacy] voxelgridmaker + detectormaker stages, graphite, 64 px, psi 0..180 x 19, mirror")
    path = V.main(_stage1_config(tmp_path, 1))
    cfg = {'iq_output_folder': path, 'gen_name': 'gr1', 'max_q': '2.0', 'num_pixels': '64',
           'angle_init_val1': '0.15', 'angle_init_ax1': 'psi', 'psi_start': '0', 'psi_end': '180', 'psi_num': '19',
           'phi_start': '0', 'phi_end': '0', 'phi_num': '1', 'theta_start': '0', 'theta_end': '0', 'theta_num': '1',
           'mirror': 'True'}
    out = D.main(dict(cfg))
    assert os.path.basename(out) == 'gr1_det_sum'
    det = np.load(f'{out}/gr1_det_sum.npy', allow_pickle=False)
    h = np.load(f'{out}/gr1_det_h.npy', allow_pickle=False)
    for f in ('gr1_det_v.npy', 'gr1_det_sum_log.png', 'gr1_det_sum_lin.png', 'gr1_config.txt'):
        assert os.path.exists(f'{out}/{f}'), f
    iq = np.load(f'{path}/gr1_iq.npy', allow_pickle=False)
    ax = np.load(f'{path}/gr1_qx.npy', allow_pickle=False)
    one = np.array([0.0])
    want, wh, _ = O.detectormaker(iq, ax, ax, ax, 64, 2.0, (0.15, 0, 0), ('psi', 'None', 'None'),
                                  np.linspace(0, 180, 19), None, one, None, one, None, mirror=True)
    assert np.array_equal(h, wh)
    assert np.all(det > 0) and not np.isnan(det).any()
    assert np.array_equal(det * 1e-6, want)
    _say(f"  image {det.shape} bit-exact vs oracle (x 1e-6)")
    assert os.path.basename(D.main(dict(cfg))) == 'gr1_det_sum1'


def test_generate_voxel_grid_low_mem_full_grid(tmp_path):
    import oracle as O
    from giwaxsim_amd import voxelgridmaker as V
    from giwaxsim_amd.tools.utilities import load_xyz
    from giwaxsim_amd.tools.voxelgrids import downselect_voxelgrid, generate_voxel_grid_low_mem
    _say("[gpu-legacy] generate_voxel_grid_low_mem, graphite, full grid")
    path = f'{GOLDEN}/graphite_medium.xyz'
    iq, qx, qy, qz = generate_voxel_grid_low_mem(path, 0.25, QV, 2.0, 1, 12700, 'gr', fill_bkg=True, smooth=2)
    coords, el = load_xyz(path)
    ids, ftab, _, _ = engine.element_table(el, 12700.0)
    ref = O.voxelgridmaker(coords, ftab[ids], 0.25, QV, 2.0, fill_bkg=True, smooth=2)
    want = np.divide(ref['sum'], ref['count'], out=np.zeros_like(ref['sum']), where=ref['count'] != 0)
    assert iq.shape == want.shape == (len(qx),) * 3
    assert np.array_equal(qx, ref['plan']['q_axis']) and np.array_equal(qy, qx) and np.array_equal(qz, qx)
    assert np.array_equal(iq == 0, want == 0)
    err = _close(iq, want)
    _say(f"  iq {iq.shape}: max rel err {err:.2e}, {int(np.sum(want == 0))} unreached voxels")
    assert err <= 1e-5 and np.sum(want == 0) > 0
    small, ax, _, _ = downselect_voxelgrid(iq, qx, qy, qz, 2.0)
    crop, cax, _, _ = V.voxel_grid_cropped(path, 0.25, QV, 2.0, 1, 12700, fill_bkg=True, smooth=2)
    assert np.array_equal(ax, cax) and np.array_equal(small, crop)
    assert generate_voxel_grid_low_mem(path, 0.25, QV, 2.0, 1, 12700, 'gr', output_dir=str(tmp_path),
                                       fill_bkg=True, smooth=2) is None
    saved = np.load(tmp_path / 'gr_output_files' / 'gr_iq.npy', allow_pickle=False)
    assert np.array_equal(saved, iq)
    for k in ('qx', 'qy', 'qz'):
        assert np.array_equal(np.load(tmp_path / 'gr_output_files' / f'gr_{k}.npy', allow_pickle=False), qx)
