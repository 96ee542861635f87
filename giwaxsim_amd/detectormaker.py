"""Second stage of the reference's legacy two-stage pipeline (old_modules/detectormaker.py).

    python -m giwaxsim_amd.detectormaker --config detectormaker_config.txt

Reads the same config keys (old_modules/detectormaker.py:16-42) and the stage handoff
``{iq_output_folder}/{gen_name}_iq|qx|qy|qz.npy`` written by
``giwaxsim_amd.voxelgridmaker``, sums the weighted orientations on the GPU (the
detector path of tools/detector.py), mirrors when asked, floors NaN and values <= 0 to
1e-6 (no 1e-6 rescale, unlike detectormaker_fitting) and writes, into a new folder
``{gen_name}_det_sum``, ``_det_sum1``, ... (detectormaker.py:54-59), the axes, the image,
its log / linear PNGs and the config (detectormaker.py:84-85, 136-166).
"""

from __future__ import annotations

import argparse
import os
import time

import numpy as np

from .tools.comparison import detector_image
from .tools.utilities import parse_config_file, save_config_to_txt, str_to_bool

__all__ = ['main']


def _pngs(det_sum_path, gen_name, det_sum, det_h, det_v):
    try:
        import matplotlib
        matplotlib.use('Agg')
        import matplotlib.pyplot as plt
    except ImportError:
        return
    extent = (np.min(det_h), np.max(det_h), np.min(det_v), np.max(det_v))
    for name, norm in (('log', matplotlib.colors.LogNorm(vmin=np.percentile(det_sum, 10),
                                                         vmax=np.percentile(det_sum, 99.9))),
                       ('lin', matplotlib.colors.Normalize(vmin=np.percentile(det_sum, 10),
                                                           vmax=np.percentile(det_sum, 99.9)))):
        fig, ax = plt.subplots()
        im = ax.imshow(det_sum, norm=norm, extent=extent, cmap='turbo', origin='lower')
        ax.set_xlabel('q horizontal (1/Å)')
        ax.set_ylabel('q vertical (1/Å)')
        ax.set_ylim(bottom=0)
        fig.colorbar(im, ax=ax)
        plt.tight_layout()
        plt.savefig(f'{det_sum_path}/{gen_name}_det_sum_{name}.png', dpi=300)
        plt.close(fig)


def main(config):
    save_path = config.get('iq_output_folder')
    gen_name = config.get('gen_name')
    max_q = float(config.get('max_q', 2.5))
    num_pixels = int(config.get('num_pixels', 500))
    angle_init_vals = tuple(float(config.get(f'angle_init_val{k}', 0)) for k in (1, 2, 3))
    angle_init_axs = tuple(config.get(f'angle_init_ax{k}', 'None') for k in (1, 2, 3))
    psis = np.linspace(float(config.get('psi_start')), float(config.get('psi_end')), num=int(config.get('psi_num')))
    phis = np.linspace(float(config.get('phi_start')), float(config.get('phi_end')), num=int(config.get('phi_num')))
    thetas = np.linspace(float(config.get('theta_start')), float(config.get('theta_end')),
                         num=int(config.get('theta_num')))
    psi_wp = config.get('psi_weights_path', None)
    phi_wp = config.get('phi_weights_path', None)
    theta_wp = config.get('theta_weights_path', None)
    mirror = str_to_bool(config.get('mirror', 'False'))
    if not os.path.exists(save_path):
        raise Exception(f'Path does not exist: {save_path}')
    iq = np.load(f'{save_path}/{gen_name}_iq.npy', allow_pickle=False)
    qx = np.load(f'{save_path}/{gen_name}_qx.npy', allow_pickle=False)
    qy = np.load(f'{save_path}/{gen_name}_qy.npy', allow_pickle=False)
    qz = np.load(f'{save_path}/{gen_name}_qz.npy', allow_pickle=False)
    det_sum_path = f'{save_path}/{gen_name}_det_sum'
    i = 0
    while os.path.exists(det_sum_path):
        i += 1
        det_sum_path = f'{save_path}/{gen_name}_det_sum{i}'
    os.mkdir(det_sum_path)
    img, det_h, det_v = detector_image(iq, qx, qy, qz, num_pixels, max_q, angle_init_vals, angle_init_axs, psis,
                                       psi_wp, phis, phi_wp, thetas, theta_wp, mirror, floor_scale=False)
    np.save(f'{det_sum_path}/{gen_name}_det_h.npy', det_h)
    np.save(f'{det_sum_path}/{gen_name}_det_v.npy', det_v)
    det_sum = img.cpu().numpy()
    det_sum[det_sum != det_sum] = 1e-6
    det_sum[det_sum <= 0] = 1e-6
    np.save(f'{det_sum_path}/{gen_name}_det_sum.npy', det_sum)
    _pngs(det_sum_path, gen_name, det_sum, det_h, det_v)
    save_config_to_txt(config, f'{det_sum_path}/{gen_name}_config.txt')
    return det_sum_path


if __name__ == "__main__":
    start = time.time()
    parser = argparse.ArgumentParser(description="Process a configuration file.")
    parser.add_argument('--config', type=str, required=True, help='Path to the configuration file')
    args = parser.parse_args()
    main(parse_config_file(args.config))
    print(f'\nTotal Time: {str(time.time() - start)}')
