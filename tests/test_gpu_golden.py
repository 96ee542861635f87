"""End to end against the reference's own golden run: the example config
(config_templates/simulate_GIWAXS_config.txt == output_data/PM6_sample/config.txt)
on test_input_files/PM6_sample.pdb, whose saved det_sum / det_h / det_v are
committed under tests/golden/.  N = ceil(2 pi / (0.02 * 0.3)) = 1048 is not a
power of two.

The comparison is loose by construction: the reference's f1/f2 come from
xraydb's Chantler tables (absent here; an offline approximation is used), its
rotations round through BLAS and its detector sum is thread-order dependent.
Axes must match exactly; the image must match in shape (log-intensity
correlation) and level (median ratio)."""

import os

import numpy as np
import pytest

from giwaxsim_amd import simulate_GIWAXS as S
from giwaxsim_amd.tools.utilities import parse_config_file

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def test_pm6_golden(tmp_path):
    cfg = parse_config_file(os.path.join(GOLDEN, 'pm6_config.txt'))
    cfg['input_filepath'] = os.path.join(GOLDEN, 'PM6_sample.pdb')
    cfg['save_folder'] = str(tmp_path)
    print("[gpu-golden] PM6 example config, N=1048, 892 frames, 2880 orientations", flush=True)
    S.main(cfg)
    out = tmp_path / 'PM6_sample'
    got = np.load(out / 'det_sum.npy', allow_pickle=False)
    dh = np.load(out / 'det_h.npy', allow_pickle=False)
    dv = np.load(out / 'det_v.npy', allow_pickle=False)
    want = np.load(os.path.join(GOLDEN, 'pm6_det_sum.npy'), allow_pickle=False)
    assert np.array_equal(dh, np.load(os.path.join(GOLDEN, 'pm6_det_h.npy'), allow_pickle=False))
    assert np.array_equal(dv, np.load(os.path.join(GOLDEN, 'pm6_det_v.npy'), allow_pickle=False))
    assert got.shape == want.shape
    lg, lw = np.log(got), np.log(want)
    corr = np.corrcoef(lg.ravel(), lw.ravel())[0, 1]
    ratio = np.median(got / want)
    print(f"  log-image correlation {corr:.5f}, median ratio {ratio:.4f}", flush=True)
    assert corr > 0.99
    assert 0.9 < ratio < 1.1
