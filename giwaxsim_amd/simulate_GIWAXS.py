"""Config-file entry point, mirroring the reference's ``simulate_GIWAXS.py``.

    python -m giwaxsim_amd.simulate_GIWAXS --config cfg.txt

Reads the same ``key=value`` config (keys and defaults of
simulate_GIWAXS.py:18-82), builds each slab, runs the GPU voxel-grid and
detector paths and saves ``det_h.npy``, ``det_v.npy``, ``det_sum.npy``
(+ PNGs when matplotlib is available) and ``config.txt`` per input file,
like simulate_GIWAXS.py:96-218.  The experimental-comparison branch
(``img_path``) needs ``fabio`` and the rebinning of
tools/comparison.py:406-445, which this package does not provide.
"""

from __future__ import annotations

import argparse
import glob
import os
import time

import numpy as np

from .tools.comparison import detectormaker_fitting, slabmaker_fitting, voxelgridmaker_fitting
from .tools.utilities import load_pdb_cell_params, parse_config_file, save_config_to_txt, str_to_bool


def main(config):
    input_folder = config.get('input_folder', None)
    input_path = config.get('input_filepath', None)
    filetype = config.get('filetype', None)
    x_size = float(config.get('x_size', 0))
    y_size = float(config.get('y_size', 0))
    z_size = float(config.get('z_size', 0))
    a = float(config.get('a', 0))
    b = float(config.get('b', 0))
    c = float(config.get('c', 0))
    alpha = float(config.get('alpha', 0))
    beta = float(config.get('beta', 0))
    gamma = float(config.get('gamma', 0))

    r_voxel_size = float(config.get('r_voxel_size', 0.3))
    q_voxel_size = float(config.get('q_voxel_size', 0.04))
    max_q = float(config.get('max_q', 2.5))
    energy = float(config.get('energy', 10000))
    fill_bkg = str_to_bool(config.get('fill_bkg', 'False'))
    smooth = int(config.get('smooth', 0))

    num_pixels = int(config.get('num_pixels', max_q / q_voxel_size))
    angle_init_vals = (float(config.get('angle_init_val1', 0)), float(config.get('angle_init_val2', 0)),
                       float(config.get('angle_init_val3', 0)))
    angle_init_axs = (config.get('angle_init_ax1', 'None'), config.get('angle_init_ax2', 'None'),
                      config.get('angle_init_ax3', 'None'))
    psis = np.linspace(float(config.get('psi_start')), float(config.get('psi_end')),
                       int(config.get('psi_num')))
    psi_weights_path = config.get('psi_weights_path', None)
    phis = np.linspace(float(config.get('phi_start')), float(config.get('phi_end')),
                       int(config.get('phi_num')))
    phi_weights_path = config.get('phi_weights_path', None)
    thetas = np.linspace(float(config.get('theta_start')), float(config.get('theta_end')),
                         int(config.get('theta_num')))
    theta_weights_path = config.get('theta_weights_path', None)
    mirror = str_to_bool(config.get('mirror', 'False'))

    img_path = config.get('img_path', None)
    save_folder = config.get('save_folder', os.getcwd())
    os.makedirs(save_folder, exist_ok=True)

    if input_folder:
        if not filetype:
            print(filetype)
            raise Exception('filetype must be specified')
        input_paths = glob.glob(f'{input_folder}/*{filetype}')
    elif input_path:
        input_paths = [input_path]
    else:
        raise Exception('Either input_folder or input_path must be specified')
    gen_names = [os.path.splitext(os.path.basename(p))[0] for p in input_paths]

    outputs = {}
    for i, path in enumerate(input_paths):
        if path.lower().endswith('.xyz'):
            a2, b2, c2, alpha2, beta2, gamma2 = a, b, c, alpha, beta, gamma
        elif path.lower().endswith('.pdb'):
            a2, b2, c2, alpha2, beta2, gamma2 = load_pdb_cell_params(path)
        else:
            raise Exception('Files must be a .pdb or .xyz file')
        if x_size == 0 or y_size == 0 or z_size == 0:
            # the reference falls back to the config's a, b, c here, even for .pdb
            x_size, y_size, z_size = a, b, c
            print('at least one slab size was not defined. Defaulting x,y,z slab dimensions to a, b, c')
        if any(val == 0 for val in [a2, b2, c2, alpha2, beta2, gamma2]):
            raise Exception('at least one unit cell parameter a, b, c, alpha, beta, gamma not defined')

        coords, elements = slabmaker_fitting(path, x_size, y_size, z_size, a2, b2, c2, alpha2, beta2,
                                             gamma2)
        iq, qx, qy, qz = voxelgridmaker_fitting(coords, elements, r_voxel_size, q_voxel_size, max_q,
                                                energy, num_cpus=None, fill_bkg=fill_bkg, smooth=smooth)
        det_sum, det_h, det_v = detectormaker_fitting(iq, qx, qy, qz, num_pixels, max_q, angle_init_vals,
                                                      angle_init_axs, psis, psi_weights_path, phis,
                                                      phi_weights_path, thetas, theta_weights_path,
                                                      mirror=mirror)
        sub = f'{save_folder}/{gen_names[i]}'
        os.makedirs(sub, exist_ok=True)
        if img_path:
            raise NotImplementedError("experimental comparison (img_path) needs fabio and the "
                                      "q-map rebinning, which giwaxsim_amd does not provide")
        if mirror:
            kh = np.where(det_h >= 0)[0]
            kv = np.where(det_v >= 0)[0]
            det_h = det_h[kh]
            det_v = det_v[kv]
            det_sum = det_sum[np.ix_(kv, kh)]
        np.save(f'{sub}/det_h.npy', det_h)
        np.save(f'{sub}/det_v.npy', det_v)
        np.save(f'{sub}/det_sum.npy', det_sum)
        outputs[gen_names[i]] = (det_sum, det_h, det_v)
        _plots(sub, det_sum, det_h, det_v)
        save_config_to_txt(config, f'{sub}/config.txt')
    return outputs


def _plots(sub, det_sum, det_h, det_v):
    try:
        import matplotlib
        matplotlib.use('Agg')
        import matplotlib.pyplot as plt
    except ImportError:
        return
    img = det_sum.copy()
    img[img <= 0] = 1e-6
    img[img != img] = 1e-6
    for name, norm in (('lin', matplotlib.colors.Normalize(vmin=np.percentile(img, 1),
                                                            vmax=np.percentile(img, 99.5))),
                       ('log', matplotlib.colors.LogNorm(vmin=np.percentile(img, 50),
                                                         vmax=np.percentile(img, 99.95)))):
        fig, ax = plt.subplots(figsize=(10, 5))
        cax = ax.imshow(img, norm=norm, extent=(np.min(det_v), np.max(det_v), np.min(det_h), np.max(det_h)),
                        cmap='turbo', origin='lower')
        fig.colorbar(cax, ax=ax, shrink=0.82, aspect=20, pad=0.02)
        ax.set_xlabel('$q_{xy}$ ($\\AA^{-1}$)')
        ax.set_ylabel('$q_z$ ($\\AA^{-1}$)')
        fig.tight_layout()
        fig.savefig(f'{sub}/det_sum_{name}.png', dpi=150)
        plt.close(fig)


def cli(argv=None):
    start = time.time()
    parser = argparse.ArgumentParser(description="Process a configuration file.")
    parser.add_argument('--config', type=str, required=True, help='Path to the configuration file')
    args = parser.parse_args(argv)
    main(parse_config_file(args.config))
    print(f'\nTotal Time: {time.time() - start}')


if __name__ == "__main__":
    cli()
