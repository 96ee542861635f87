"""Mirror of the reference's ``tools/detector.py`` call surface on the GPU.

Rotation matrices are built on the host in FP64 from the detector's corner
pixels exactly like the reference; every per-pixel operation (rotation,
nearest-voxel intersection, weighting, accumulation, mirror) runs in the
gfx950 kernels of ``libgiwaxsim_amd.so``.  Rotations evaluate
``R @ (x, y, z)`` row by row, left to right, without FMA (the reference's
BLAS ``@`` rounding is machine dependent).
"""

from __future__ import annotations

import ctypes
from multiprocessing import shared_memory

import numpy as np

from .. import engine
from .._lib import GX_FIN_MIRROR, call
from ..engine import _axis_matrix

__all__ = ['make_detector', 'rotate_about_normal', 'rotate_about_horizontal',
           'rotate_about_vertical', 'intersect_detector', 'rotate_psi_phi_theta',
           'mirror_vertical_horizontal', 'generate_detector_ints']


def make_detector(h_max, num_pixels_h, v_max, num_pixels_v):
    """tools/detector.py:5-29: a detector in the qy-qz plane about the origin.
    Returns (det_x, det_y, det_z, h_axis_vals, v_axis_vals)."""
    h = np.linspace(-h_max, h_max, num_pixels_h)
    v = np.linspace(-v_max, v_max, num_pixels_v)
    det_y, det_z = np.meshgrid(h, v)
    det_x = np.zeros_like(det_y)
    return det_x, det_y, det_z, h, v


def _corners(x, y, z):
    return (np.array([x[0, 0], x[0, -1], x[-1, 0]]), np.array([y[0, 0], y[0, -1], y[-1, 0]]),
            np.array([z[0, 0], z[0, -1], z[-1, 0]]))


def _rotate(R, det_x, det_y, det_z):
    import torch
    shape = np.shape(det_x)
    x = engine.to_dev(np.ravel(det_x), torch.float64)
    y = engine.to_dev(np.ravel(det_y), torch.float64)
    z = engine.to_dev(np.ravel(det_z), torch.float64)
    out = torch.empty((3, x.numel()), dtype=torch.float64, device=x.device)
    call("gx_rotate_points", (ctypes.c_double * 9)(*np.asarray(R, dtype=np.float64).ravel()),
         engine.ptr(x), engine.ptr(y), engine.ptr(z), int(x.numel()), engine.ptr(out[0]),
         engine.ptr(out[1]), engine.ptr(out[2]), engine.stream_handle())
    o = out.cpu().numpy()
    return o[0].reshape(shape), o[1].reshape(shape), o[2].reshape(shape)


def rotate_about_normal(det_x_grid, det_y_grid, det_z_grid, psi):
    """Rotate about the detector normal by psi degrees (tools/detector.py:33-78)."""
    R = _axis_matrix('psi', _corners(det_x_grid, det_y_grid, det_z_grid), psi)
    return _rotate(R, det_x_grid, det_y_grid, det_z_grid)


def rotate_about_horizontal(det_x_grid, det_y_grid, det_z_grid, theta):
    """Rotate about the detector's horizontal axis (tools/detector.py:80-120)."""
    R = _axis_matrix('theta', _corners(det_x_grid, det_y_grid, det_z_grid), theta)
    return _rotate(R, det_x_grid, det_y_grid, det_z_grid)


def rotate_about_vertical(det_x_grid, det_y_grid, det_z_grid, phi):
    """Rotate about the detector's vertical axis (tools/detector.py:122-162)."""
    R = _axis_matrix('phi', _corners(det_x_grid, det_y_grid, det_z_grid), phi)
    return _rotate(R, det_x_grid, det_y_grid, det_z_grid)


def rotate_psi_phi_theta(det_x, det_y, det_z, psi, phi, theta):
    """tools/detector.py:234-244."""
    det_x, det_y, det_z = rotate_about_normal(det_x, det_y, det_z, psi)
    det_x, det_y, det_z = rotate_about_vertical(det_x, det_y, det_z, phi)
    det_x, det_y, det_z = rotate_about_horizontal(det_x, det_y, det_z, theta)
    return det_x, det_y, det_z


def _intersect_into(iq, qx, qy, qz, det_x, det_y, det_z, weight, out_dev):
    import torch
    iq_d = engine.to_dev(np.asarray(iq, dtype=np.float64), torch.float64)
    cy, cx, cz = iq_d.shape
    x = engine.to_dev(np.ravel(det_x), torch.float64)
    y = engine.to_dev(np.ravel(det_y), torch.float64)
    z = engine.to_dev(np.ravel(det_z), torch.float64)
    dq = np.diff(qz)[0]
    call("gx_intersect", engine.ptr(iq_d), cy, cx, cz, float(np.min(qx)), float(np.min(qy)),
         float(np.min(qz)), float(dq), engine.ptr(x), engine.ptr(y), engine.ptr(z), int(x.numel()),
         float(weight), 1 if out_dev is not None else 0,
         engine.ptr(out_dev), engine.stream_handle())


def intersect_detector(int_voxels, qx, qy, qz, det_x_grid, det_y_grid, det_z_grid):
    """Nearest-voxel intersection (the effective definition,
    tools/detector.py:194-232): idx = (det - min q) // dq, clipped to the grid."""
    import torch
    n = int(np.size(det_x_grid))
    out = torch.zeros(n, dtype=torch.float64, device=engine.device())
    _intersect_into(int_voxels, qx, qy, qz, det_x_grid, det_y_grid, det_z_grid, 1.0, out)
    return out.cpu().numpy().reshape(np.shape(det_x_grid))


def mirror_vertical_horizontal(qmap):
    """tools/detector.py:246-275 (odd centre row/column doubled, centre x4)."""
    import torch
    q = engine.to_dev(np.asarray(qmap, dtype=np.float64), torch.float64)
    nr, nc = q.shape
    out = torch.empty_like(q)
    call("gx_detector_finalize", engine.ptr(q), nr, nc, GX_FIN_MIRROR, engine.ptr(out),
         engine.stream_handle())
    return out.cpu().numpy()


def generate_detector_ints(args):
    """One orientation added into a named shared-memory image
    (tools/detector.py:278-300).  The addition happens on the GPU; the
    shared image is read and written back once."""
    import torch
    (iq, qx, qy, qz, det_x, det_y, det_z, psi, psi_weight, phi, phi_weight, theta, theta_weight,
     det_ints_shm_name) = args
    x2, y2, z2 = rotate_psi_phi_theta(det_x, det_y, det_z, psi, phi, theta)
    shm = shared_memory.SharedMemory(name=det_ints_shm_name)
    try:
        det_ints = np.ndarray(np.shape(det_x), dtype=np.float64, buffer=shm.buf)
        acc = engine.to_dev(det_ints.ravel(), torch.float64)
        _intersect_into(iq, qx, qy, qz, x2, y2, z2, psi_weight * phi_weight * theta_weight, acc)
        det_ints[...] = acc.cpu().numpy().reshape(det_ints.shape)
        del det_ints
    finally:
        shm.close()
