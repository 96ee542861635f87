"""Config-file entry point, mirroring the reference's ``simulate_GIWAXS.py``.

    python -m giwaxsim_amd.simulate_GIWAXS --config cfg.txt

Reads the same ``key=value`` config (keys and defaults of
simulate_GIWAXS.py:18-82), builds each slab, runs the GPU voxel-grid and
detector paths and saves ``det_h.npy``, ``det_v.npy``, ``det_sum.npy``
(+ PNGs when matplotlib is available) and ``config.txt`` per input file,
like simulate_GIWAXS.py:96-218.  With ``img_path`` set, the experimental-comparison
branch (simulate_GIWAXS.py:123-169) mirrors, trims and rebins the experimental map,
pads the simulated peaks, fits scale / offset when ``fit_scale_offset`` and saves the
trimmed axes, ``rebin_map.npy``, ``sim_comp_map.npy`` and the comparison figures; the
image is read from ``.npy`` directly, other formats through fabio when installed.
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

    mask_path = config.get('mask_path', None)
    img_path = config.get('img_path', None)
    qxy_path = config.get('qxy_path', None)
    qz_path = config.get('qz_path', None)
    fit_scale_offset = str_to_bool(config.get('fit_scale_offset', 'False'))
    pad_width = float(config.get('pad_width', 0))
    pad_range = (float(config.get('pad_range_min', 0)), float(config.get('pad_range_max', 0)))
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
            outputs[gen_names[i]] = _compare(sub, det_sum, det_h, det_v, img_path, qxy_path, qz_path, mask_path,
                                             pad_width, pad_range, fit_scale_offset, max_q)
            save_config_to_txt(config, f'{sub}/config.txt')
            continue
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


def _compare(sub, det_sum, det_h, det_v, img_path, qxy_path, qz_path, mask_path, pad_width, pad_range,
             fit_scale_offset, max_q):
    """simulate_GIWAXS.py:123-169 for one simulated image."""
    from .tools.comparison import (load_exp_image, mirror_qmap_positive_qxy_only, optimize_scale_offset,
                                   plot_fit, plot_linecuts, rebin_qmap_and_mask, shift_peak, trim_sim_data)
    exp_img = load_exp_image(img_path)
    exp_qxy = np.loadtxt(qxy_path)
    exp_qz = np.loadtxt(qz_path)
    mask = np.load(mask_path, allow_pickle=False)
    exp_img, exp_qxy, exp_qz = mirror_qmap_positive_qxy_only(exp_img, exp_qxy, exp_qz)
    sim_trim, det_h_trim, det_v_trim = trim_sim_data(det_sum, det_h, det_v, exp_qxy, exp_qz)
    rebin_map, rebin_mask = rebin_qmap_and_mask(det_h_trim, det_v_trim, exp_img, exp_qxy, exp_qz, mask)
    rebin_map[rebin_mask == 1] = 0
    sim_trim[rebin_mask == 1] = 0
    if pad_width > 0:
        sim_pad = shift_peak(sim_trim.copy(), det_h_trim, det_v_trim, pad_width, pad_range)
    else:
        sim_pad = sim_trim.copy()
    if fit_scale_offset:
        scale, offset = optimize_scale_offset(sim_pad, rebin_map, rebin_mask)
        sim_comp_map = scale * sim_pad + offset
        sim_comp_map[rebin_mask == 1] = 0
    else:
        sim_comp_map = sim_pad.copy()
    diff_map = np.where(rebin_map > 1e-10, sim_comp_map - rebin_map, 0)
    np.save(f'{sub}/det_h.npy', det_h_trim)
    np.save(f'{sub}/det_v.npy', det_v_trim)
    np.save(f'{sub}/rebin_map.npy', rebin_map)
    np.save(f'{sub}/sim_comp_map.npy', sim_comp_map)
    try:
        import matplotlib  # noqa: F401
    except ImportError:
        return rebin_map, sim_comp_map, diff_map
    plot_fit(f'{sub}/fit_result.png', rebin_map, sim_comp_map, diff_map, det_h_trim, det_v_trim, max_q, '')
    plot_linecuts(f'{sub}/fit_result_linecut.png', rebin_map, sim_comp_map, det_h_trim, det_v_trim, '')
    return rebin_map, sim_comp_map, diff_map


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
