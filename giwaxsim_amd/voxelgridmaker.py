"""First stage of the reference's legacy two-stage pipeline (old_modules/voxelgridmaker.py).

    python -m giwaxsim_amd.voxelgridmaker --config voxelgridmaker_config.txt

Reads the same config keys (input_folder / input_filepath, filetype, gen_name,
r_voxel_size, q_voxel_size, aff_num_qs, energy, max_q, output_dir, num_cpus,
scratch_folder, smooth, fill_bkg; old_modules/voxelgridmaker.py:10-25), builds each
structure's I(q) voxel grid on the GPU (``engine.run_voxelgrid``: the phi-frame path of
tools/voxelgrids.py generate_voxel_grid_low_mem 535-722, cropped to |q| < max_q + dq
as downselect_voxelgrid 16-48), averages the grids over the files, applies
(f0(q)/Z)^2 of the first file's most common element, and writes the stage handoff
``{output_dir}/{gen_name}_output_files/{gen_name}_iq|qx|qy|qz.npy`` plus
``{gen_name}_config.txt`` (voxelgridmaker.py:75-84), which ``giwaxsim_amd.detectormaker``
reads.

``aff_num_qs > 1`` follows the reference's q-shell branch (tools/voxelgrids.py:650-707)
as it computes: each shell's pass overwrites the grid, so the result is the last
shell's grid (f = f0(q_shell) + f1 + i f2 with its f0 evaluated at q, not q^2, as
tools/utilities.py:318-339 does) with the voxels of that shell doubled, and no final
f0 factor.  Only that last pass is run; the earlier ones do not reach the output.
"""

from __future__ import annotations

import argparse
import glob
import os
import time

import numpy as np

from . import engine
from .tools.ptable_dict import aff_dict
from .tools.utilities import (load_pdb, load_xyz, most_common_element, parse_config_file, save_config_to_txt,
                              str_to_bool)
from .tools.voxelgrids import add_f0_q_3d

__all__ = ['main', 'voxel_grid_cropped', 'shell_f0']


def shell_f0(q_val, element):
    """f0 of get_element_f0_dict (tools/utilities.py:318-339): the Cromer-Mann sum with
    q_val where the formula has q^2."""
    a = aff_dict[element]
    return (a[0] * np.exp(-a[1] * (q_val) / (16 * np.pi ** 2)) +
            a[2] * np.exp(-a[3] * (q_val) / (16 * np.pi ** 2)) +
            a[4] * np.exp(-a[5] * (q_val) / (16 * np.pi ** 2)) +
            a[6] * np.exp(-a[7] * (q_val) / (16 * np.pi ** 2)) +
            a[8])


def voxel_grid_cropped(input_path, r_voxel_size, q_voxel_size, max_q, aff_num_qs, energy, fill_bkg=False,
                       smooth=0):
    """(iq [qy, qx, qz] FP64 numpy, qx, qy, qz) of one structure file on the cropped
    grid, without the f0 factor (generate_voxel_grid_low_mem followed by
    downselect_voxelgrid)."""
    if input_path[-3:] == 'xyz':
        coords, elements = load_xyz(input_path)
    elif input_path[-3:] == 'pdb':
        coords, elements = load_pdb(input_path)
    else:
        raise Exception('files must be a .pdb or .xyz file')
    if aff_num_qs == 1:
        res = engine.run_voxelgrid(coords, elements, r_voxel_size, q_voxel_size, max_q, energy,
                                   fill_bkg=fill_bkg, smooth=int(smooth), f0_element=None)
        return res['iq64'].cpu().numpy(), res['qx'], res['qy'], res['qz']
    if aff_num_qs > 1:
        from .tools.f1f2 import f1_f2
        ids, _, _, uniq = engine.element_table(elements, energy)
        grid = engine.qgrid(float(q_voxel_size), float(max_q))
        step = grid.max_q_diag / int(aff_num_qs)
        q_val = 0.5 * step + (int(aff_num_qs) - 1) * step    # the last shell's centre
        upper_q, lower_q = q_val + step / 2, q_val - step / 2
        f_table = np.array([shell_f0(q_val, str(e)) + complex(*f1_f2(str(e), energy)) for e in uniq],
                           dtype=np.complex128)
        res = engine.run_voxelgrid(coords, None, r_voxel_size, q_voxel_size, max_q, energy, fill_bkg=fill_bkg,
                                   smooth=int(smooth), f_table=f_table, elem_ids=ids, f0_element=None)
        iq = res['iq64'].cpu().numpy()
        qx, qy, qz = res['qx'], res['qy'], res['qz']
        qxm, qym, qzm = np.meshgrid(qx, qy, qz)
        qr = np.sqrt(qxm ** 2 + qym ** 2 + qzm ** 2)
        iq += np.where((qr <= upper_q) & (qr > lower_q), iq, 0)
        return iq, qx, qy, qz
    raise Exception('Invalid aff_num_qs value. Must be non-negative integer')


def main(config):
    input_folder = config.get('input_folder', None)
    input_filepath = config.get('input_filepath', None)
    filetype = config.get('filetype', 'xyz')
    gen_name = config.get('gen_name')
    r_voxel_size = float(config.get('r_voxel_size', 0.3))
    q_voxel_size = float(config.get('q_voxel_size', 0.01))
    aff_num_qs = int(config.get('aff_num_qs', 1))
    energy = float(config.get('energy', 1))
    max_q = float(config.get('max_q', 2.5))
    output_dir = config.get('output_dir', os.getcwd())
    smooth = int(config.get('smooth', 0))
    fill_bkg = str_to_bool(config.get('fill_bkg', 'False'))
    if input_folder:
        input_paths = glob.glob(f'{input_folder}/*{filetype}')
    elif input_filepath:
        input_paths = [input_filepath]
    else:
        raise Exception('Either input_folder or input_path must be specified')
    iq_sum = None
    for input_path in input_paths:
        iq, qx, qy, qz = voxel_grid_cropped(input_path, r_voxel_size, q_voxel_size, max_q, aff_num_qs, energy,
                                            fill_bkg=fill_bkg, smooth=smooth)
        iq_sum = iq if iq_sum is None else iq_sum + iq
    iq = iq_sum / len(input_paths)
    if aff_num_qs == 1:
        iq = add_f0_q_3d(iq, qx, qy, qz, most_common_element(input_paths[0]))
    save_path = f'{output_dir}/{gen_name}_output_files'
    if not os.path.exists(save_path):
        os.mkdir(save_path)
    np.save(f'{save_path}/{gen_name}_iq.npy', iq)
    np.save(f'{save_path}/{gen_name}_qx.npy', qx)
    np.save(f'{save_path}/{gen_name}_qy.npy', qy)
    np.save(f'{save_path}/{gen_name}_qz.npy', qz)
    save_config_to_txt(config, f'{save_path}/{gen_name}_config.txt')
    return save_path


if __name__ == "__main__":
    start = time.time()
    parser = argparse.ArgumentParser(description="Process a configuration file.")
    parser.add_argument('--config', type=str, required=True, help='Path to the configuration file')
    args = parser.parse_args()
    main(parse_config_file(args.config))
    print(f'\nTotal Time: {str(time.time() - start)}')
