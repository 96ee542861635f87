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

from .tools.utilities import most_common_element, parse_config_file, save_config_to_txt, str_to_bool
from .tools.voxelgrids import add_f0_q_3d, shell_f0, structure_voxel_grid

__all__ = ['main', 'voxel_grid_cropped', 'shell_f0']


def voxel_grid_cropped(input_path, r_voxel_size, q_voxel_size, max_q, aff_num_qs, energy, fill_bkg=False,
                       smooth=0):
    """(iq [qy, qx, qz] FP64 numpy, qx, qy, qz) of one structure file on the cropped
    grid, without the f0 factor (generate_voxel_grid_low_mem followed by
    downselect_voxelgrid)."""
    return structure_voxel_grid(input_path, r_voxel_size, q_voxel_size, max_q, aff_num_qs, energy,
                                fill_bkg=fill_bkg, smooth=smooth, crop=True)


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
