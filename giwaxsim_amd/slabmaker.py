"""Standalone slab maker, mirroring the reference's ``slabmaker.py``.

    python -m giwaxsim_amd.slabmaker --config slab_config.txt

Reads the same ``key=value`` config (input_filepath, output_folder, gen_name,
x_size, y_size, z_size, a, b, c, alpha, beta, gamma; slabmaker.py:8-21), tiles the
unit cell and cuts the centred box on the GPU (``tools.comparison.SlabBuilder``:
the reference's replication order and arithmetic, so the atoms and their
coordinates are identical) and writes
``{output_folder}/{gen_name}_rect_cut{X}x{Y}x{Z}.xyz`` like slabmaker.py:110-111.
Unlike ``slabmaker_fitting``, this script writes the cut atoms without the second
shift to the origin (slabmaker.py:98-108).
"""

from __future__ import annotations

import argparse
import os

from .tools.comparison import SlabBuilder
from .tools.utilities import load_pdb, load_xyz, parse_config_file, write_xyz

__all__ = ['main', 'make_slab']


def make_slab(input_filepath, x_size, y_size, z_size, a, b, c, alpha, beta, gamma):
    """(coords (n, 3) FP64, element symbols) of the cut slab (slabmaker.py:27-108)."""
    import torch
    if input_filepath[-3:] == 'xyz':
        cell = load_xyz(input_filepath)
    elif input_filepath[-3:] == 'pdb':
        cell = load_pdb(input_filepath)
    else:
        raise Exception('files must be a .pdb or .xyz file')
    builder = SlabBuilder(input_filepath, a, b, c, alpha, beta, gamma, cell=cell)
    try:
        x, y, z, el = builder.build(x_size, y_size, z_size, shift_again=False)
    except AssertionError:
        raise AssertionError() from None   # slabmaker.py:87-89 asserts without a message
    coords = torch.stack((x, y, z), dim=1).cpu().numpy()
    return coords, builder.uniq[el.cpu().numpy()]


def main(config):
    input_filepath = config.get('input_filepath')
    output_folder = config.get('output_folder', os.path.dirname(input_filepath))
    gen_name = config.get('gen_name')
    x_size = float(config.get('x_size'))
    y_size = float(config.get('y_size'))
    z_size = float(config.get('z_size'))
    a = float(config.get('a'))
    b = float(config.get('b'))
    c = float(config.get('c'))
    alpha = float(config.get('alpha'))
    beta = float(config.get('beta'))
    gamma = float(config.get('gamma'))
    if not os.path.exists(output_folder):
        os.mkdir(output_folder)
    coords, elements = make_slab(input_filepath, x_size, y_size, z_size, a, b, c, alpha, beta, gamma)
    save_path = f'{output_folder}/{gen_name}_rect_cut{int(x_size)}x{int(y_size)}x{int(z_size)}.xyz'
    write_xyz(save_path, coords, elements)
    return save_path


if __name__ == "__main__":
    parser = argparse.ArgumentParser(description="Process a configuration file.")
    parser.add_argument('--config', type=str, required=True, help='Path to the configuration file')
    args = parser.parse_args()
    main(parse_config_file(args.config))
