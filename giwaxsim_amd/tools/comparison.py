"""Mirror of the reference's pipeline drivers (``tools/comparison.py``):
``voxelgridmaker_fitting``, ``detectormaker_fitting``,
``optimize_scale_offset`` and the slab builder ``slabmaker_fitting``.

The two drivers run every frame / orientation on the GPU with all
intermediate data device resident, and return numpy arrays in the
reference's layouts: iq [qy, qx, qz] and det_sum [v, h], both FP64.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import engine
from .._lib import GX_MOMENTS_WORK, call
from .utilities import calc_real_space_abc, load_pdb, load_xyz

__all__ = ['slabmaker_fitting', 'voxelgridmaker_fitting', 'detectormaker_fitting',
           'optimize_scale_offset', 'DeviceIQ']


class DeviceIQ(np.ndarray):
    """numpy FP64 iq that also keeps the device copy it was read back from,
    so ``detectormaker_fitting`` can gather from it without a re-upload."""

    def __array_finalize__(self, obj):
        self._gx_device = getattr(obj, '_gx_device', None) if obj is not None else None


def slabmaker_fitting(input_filepath, x_size, y_size, z_size, a, b, c, alpha, beta, gamma):
    """Tile the unit cell and crop a centred slab (tools/comparison.py:595-671).
    Structure preparation on the host (outside the forward hot path)."""
    if input_filepath.lower().endswith('.xyz'):
        coords, elements = load_xyz(input_filepath)
    elif input_filepath.lower().endswith('.pdb'):
        coords, elements = load_pdb(input_filepath)
    else:
        raise Exception('Files must be a .pdb or .xyz file')
    a_vect, b_vect, c_vect = calc_real_space_abc(a, b, c, alpha, beta, gamma)
    num_x = int(np.ceil(2 * x_size / a_vect[0]))
    num_y = int(np.ceil(2 * y_size / b_vect[1]))
    num_z = int(np.ceil(2 * z_size / c_vect[2]))
    # replicate along a, then b, then c: copies 0..num (the reference appends
    # num translated copies of the running set each time)
    for vec, num in ((np.array([a_vect[0], 0, 0]), num_x),
                     (np.array([b_vect[0], b_vect[1], 0]), num_y),
                     (np.array([c_vect[0], c_vect[1], c_vect[2]]), num_z)):
        base_c, base_e = coords.copy(), elements.copy()
        parts_c, parts_e = [coords], [elements]
        for i in range(num):
            parts_c.append(base_c + vec * (i + 1))
            parts_e.append(base_e)
        coords = np.concatenate(parts_c, axis=0)
        elements = np.concatenate(parts_e, axis=0)
    x_max = np.max(coords[:, 0]) - np.min(coords[:, 0])
    y_max = np.max(coords[:, 1]) - np.min(coords[:, 1])
    z_max = np.max(coords[:, 2]) - np.min(coords[:, 2])
    assert x_max > x_size, "x_max must be greater than x_size"
    assert y_max > y_size, "y_max must be greater than y_size"
    assert z_max > z_size, "z_max must be greater than z_size"
    x_buffer = (x_max - x_size) / 2
    y_buffer = (y_max - y_size) / 2
    z_buffer = (z_max - z_size) / 2
    coords[:, 0] -= np.min(coords[:, 0])
    coords[:, 1] -= np.min(coords[:, 1])
    coords[:, 2] -= np.min(coords[:, 2])
    mask = ((coords[:, 0] >= x_buffer) & (coords[:, 0] <= x_max - x_buffer) &
            (coords[:, 1] >= y_buffer) & (coords[:, 1] <= y_max - y_buffer) &
            (coords[:, 2] >= z_buffer) & (coords[:, 2] <= z_max - z_buffer))
    cf = coords[mask]
    ef = elements[mask]
    cf[:, 0] -= np.min(cf[:, 0])
    cf[:, 1] -= np.min(cf[:, 1])
    cf[:, 2] -= np.min(cf[:, 2])
    return cf, ef


def voxelgridmaker_fitting(coords, elements, r_voxel_size, q_voxel_size, max_q, energy,
                           num_cpus=None, fill_bkg=False, smooth=0):
    """3D I(q) voxel grid by the phi-frame method (tools/comparison.py:673-788).

    Returns (iq [qy, qx, qz] FP64, qx, qy, qz) of the cropped grid
    (|q| < max_q + dq per axis), f0_C(q)^2 applied.  ``num_cpus`` is accepted
    for signature compatibility; all frames run on the GPU.
    """
    res = engine.run_voxelgrid(coords, elements, r_voxel_size, q_voxel_size, max_q, energy,
                               fill_bkg=fill_bkg, smooth=int(smooth), want64=True)
    iq = res['iq64'].cpu().numpy().view(DeviceIQ)
    iq._gx_device = res['iq64']
    return iq, res['qx'], res['qy'], res['qz']


def _weights(path, angles):
    return np.load(path) if path else np.ones_like(angles) / len(angles)


def detectormaker_fitting(iq, qx, qy, qz, num_pixels, max_q, angle_init_vals, angle_init_axs, psis,
                          psi_weights_path, phis, phi_weights_path, thetas, theta_weights_path,
                          mirror=True):
    """Orientation-averaged detector image (tools/comparison.py:790-870).

    Orientations run in (psi outer, phi, theta inner) order and are summed in
    that order in FP64 on the GPU (deterministic; the reference's threaded
    ``+=`` is racy).  Returns (det_sum [v, h] FP64, det_h, det_v)."""
    import torch
    psi_w = _weights(psi_weights_path, psis)
    phi_w = _weights(phi_weights_path, phis)
    th_w = _weights(theta_weights_path, thetas)
    assert len(psis) == len(psi_w), 'psi weights length must equal psi_num'
    assert len(phis) == len(phi_w), 'phi weights length must equal phi_num'
    assert len(thetas) == len(th_w), 'theta weights length must equal theta_num'
    assert np.abs(1 - np.sum(psi_w)) < 0.01, 'psi weights must sum to 1'
    assert np.abs(1 - np.sum(phi_w)) < 0.01, 'phi weights must sum to 1'
    assert np.abs(1 - np.sum(th_w)) < 0.01, 'theta weights must sum to 1'
    plan = engine.detector_plan(num_pixels, max_q, angle_init_vals, angle_init_axs, psis, psi_w,
                                phis, phi_w, thetas, th_w)
    dev_iq = getattr(iq, '_gx_device', None)
    if dev_iq is None or tuple(dev_iq.shape) != tuple(np.shape(iq)):
        dev_iq = engine.to_dev(np.asarray(iq, dtype=np.float64), torch.float64)
    out, _ = engine.run_detector(dev_iq, qx, qy, qz, plan, mirror)
    return out.cpu().numpy(), plan.h, plan.v


def optimize_scale_offset(sim_map, rebin_map, rebin_mask):
    """Least-squares scale and offset over unmasked pixels
    (tools/comparison.py:873-882): centred FP64 moments reduced on the GPU,
    closed-form 2x2 solution on the host."""
    import torch
    s = engine.to_dev(np.asarray(sim_map, dtype=np.float64).ravel(), torch.float64)
    y = engine.to_dev(np.asarray(rebin_map, dtype=np.float64).ravel(), torch.float64)
    m = engine.to_dev((np.asarray(rebin_mask) != 0).ravel().astype(np.uint8), torch.uint8)
    work = torch.empty(GX_MOMENTS_WORK, dtype=torch.float64, device=s.device)
    res = (ctypes.c_double * 6)()
    call("gx_scale_offset_moments", engine.ptr(s), engine.ptr(y), engine.ptr(m), int(s.numel()),
         engine.ptr(work), res, engine.stream_handle())
    n, ms, my, sss, ssy = res[0], res[1], res[2], res[3], res[4]
    if n == 0:
        raise np.linalg.LinAlgError("no unmasked pixels")
    scale = ssy / sss if sss != 0 else 0.0
    offset = my - scale * ms
    return scale, offset
