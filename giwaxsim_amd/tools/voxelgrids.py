"""Mirror of the reference's ``tools/voxelgrids.py`` call surface (the
phi-frame "low-mem" path) on the GPU.

The per-frame helpers keep the reference's signatures, including the
named shared-memory accumulators of ``rotate_project_fft_coords`` /
``process_file2``; the accumulation itself runs on the GPU and the shared
grids are read and written back once per call.  ``voxelgridmaker_fitting``
(giwaxsim_amd.tools.comparison) runs all frames device-resident instead.
"""

from __future__ import annotations

import ctypes
from multiprocessing import shared_memory

import numpy as np

from .. import engine
from .._lib import call, dbl_array
from .ptable_dict import aff_dict, ptable

__all__ = ['downselect_voxelgrid', 'rectangular_collapse_lengths', 'rotate_project_fft_coords',
           'process_file2', 'add_f0_q_3d']


def downselect_voxelgrid(grid, x_axis, y_axis, z_axis, max_val):
    """tools/voxelgrids.py:16-48: keep |axis| < max_val + voxel on each axis.
    Works on numpy arrays and device tensors (a view is returned)."""
    voxel_size = np.abs(x_axis[1] - x_axis[0])
    max_val = max_val + voxel_size
    xi = np.where(np.abs(x_axis) < max_val)[0]
    yi = np.where(np.abs(y_axis) < max_val)[0]
    zi = np.where(np.abs(z_axis) < max_val)[0]
    small = grid[yi[0]:yi[-1] + 1, xi[0]:xi[-1] + 1, zi[0]:zi[-1] + 1]
    return (small, x_axis[xi[0]:xi[-1] + 1], y_axis[yi[0]:yi[-1] + 1], z_axis[zi[0]:zi[-1] + 1])


def rectangular_collapse_lengths(x_vals, hor_length, ver_length, phi):
    """Path length through the phi-rotated rectangle at each x
    (tools/voxelgrids.py:238-309), evaluated on the GPU."""
    import torch
    from .._lib import GX_FP_STRIDE, FP_LEN_MODE, FP_CT
    rec = np.zeros(GX_FP_STRIDE)
    rec[FP_LEN_MODE:FP_CT + 1] = engine.collapse_record(phi, hor_length, ver_length)
    x = engine.to_dev(np.asarray(x_vals, dtype=np.float64).ravel(), torch.float64)
    out = torch.empty_like(x)
    call("gx_collapse_lengths", dbl_array(rec), engine.ptr(x), int(x.numel()), engine.ptr(out),
         engine.stream_handle())
    return out.cpu().numpy()


def add_f0_q_3d(iq, qx_axis, qy_axis, qz_axis, element):
    """iq * ((f0_element(|q|) / Z)^2) (tools/voxelgrids.py:828-857) on the GPU."""
    import torch
    Z = ptable[element]
    aff = aff_dict[element]
    g = engine.to_dev(np.asarray(iq, dtype=np.float64), torch.float64)
    ny, nx, nz = g.shape
    qx = engine.to_dev(np.asarray(qx_axis, dtype=np.float64), torch.float64)
    qy = engine.to_dev(np.asarray(qy_axis, dtype=np.float64), torch.float64)
    qz = engine.to_dev(np.asarray(qz_axis, dtype=np.float64), torch.float64)
    out = torch.empty_like(g)
    call("gx_apply_f0", engine.ptr(g), ny, nx, nz, engine.ptr(qx), engine.ptr(qy), engine.ptr(qz),
         dbl_array(aff), float(Z), engine.ptr(out), engine.stream_handle())
    return out.cpu().numpy()


def _same_axes(qx, qy, qz):
    if not (np.array_equal(qx, qy) and np.array_equal(qy, qz)):
        raise NotImplementedError("the GPU binning expects qx, qy and qz to be the same axis "
                                  "(as voxelgridmaker_fitting builds them)")


def _attach(name, shape):
    shm = shared_memory.SharedMemory(name=name)
    return shm, np.ndarray(shape, dtype=np.float64, buffer=shm.buf)


def process_file2(iq_2d, det_h_qx, det_h_qy, det_v_qz, qx, qy, qz, voxel_grid_shm_name,
                  voxel_grid_count_shm_name):
    """Bin one frame into the named (q_num,)^3 FP64 sum/count grids
    (tools/voxelgrids.py:464-506) on the GPU."""
    import torch
    q_num = len(qx)
    shape = (q_num, q_num, q_num)
    s_shm, vg = _attach(voxel_grid_shm_name, shape)
    c_shm, vc = _attach(voxel_grid_count_shm_name, shape)
    try:
        dev = engine.device()
        gs = engine.to_dev(vg.ravel(), torch.float64)
        gc = engine.to_dev(vc.ravel(), torch.float64)
        iq = engine.to_dev(np.asarray(iq_2d, dtype=np.float64), torch.float64)
        nv, nh = iq.shape
        hx = engine.to_dev(det_h_qx, torch.float64)
        hy = engine.to_dev(det_h_qy, torch.float64)
        vz = engine.to_dev(det_v_qz, torch.float64)
        qlim = dbl_array([np.min(qx), np.max(qx), np.min(qy), np.max(qy), np.min(qz), np.max(qz)])
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        call("gx_bin_frame_f64", engine.ptr(iq), nv, nh, engine.ptr(hx), engine.ptr(hy),
             engine.ptr(vz), qlim, float(np.diff(qz)[0]), q_num, engine.ptr(gs), engine.ptr(gc),
             engine.ptr(status), engine.stream_handle())
        if int(status.item()):
            raise IndexError(f"{int(status.item())} pixel(s) index outside the voxel grid")
        vg[...] = gs.cpu().numpy().reshape(shape)
        vc[...] = gc.cpu().numpy().reshape(shape)
        del vg, vc
    finally:
        s_shm.close()
        c_shm.close()


def rotate_project_fft_coords(args):
    """One phi-frame: rotate, project, scatter, fill/smooth, 2D FFT, |F|^2 and
    binning into the named shared grids (tools/voxelgrids.py:311-416), with
    every step on the GPU."""
    import torch
    (coords, f_values, phi, grid_size, r_voxel_size, avg_voxel_f, x_bound, y_bound, z_bound,
     fill_bkg, smooth, qx, qy, qz, voxel_grid_shm_name, voxel_grid_count_shm_name) = args
    _same_axes(qx, qy, qz)
    coords = np.asarray(coords, dtype=np.float64)
    if not engine._lib.load().gx_supported_grid(int(grid_size)):
        raise NotImplementedError(f"grid_size {grid_size} is outside the supported range 16..2048")
    g = engine.qgrid_from_axis(qz, [phi])
    plan = engine.build_frame_plan(int(grid_size), float(r_voxel_size), g, x_bound, y_bound,
                                   np.max(coords[:, 2]) - np.min(coords[:, 2]))
    f_table, elem_ids = np.unique(np.asarray(f_values, dtype=np.complex128), return_inverse=True)
    counts = np.bincount(elem_ids, minlength=len(f_table))
    B = engine.intensity_bound(float(np.sum(counts * np.abs(f_table))), plan.N, fill_bkg, smooth,
                               avg_voxel_f, plan.max_voxels, y_bound, x_bound, r_voxel_size)
    bins = engine.Bins(g, engine.bin_scale_for([(1, engine.pixels_per_bin(plan), B)]))
    slab = engine.Slab(coords, elem_ids.astype(np.int32), f_table, plan.N, r_voxel_size,
                       np.min(coords[:, 2]))
    runner = engine.FrameRunner(plan, slab, bins, fill_bkg, smooth, avg_voxel_f, batch=1)
    runner.run()
    runner.check()
    q_num = len(qz)
    shape = (q_num,) * 3
    s_shm, vg = _attach(voxel_grid_shm_name, shape)
    c_shm, vc = _attach(voxel_grid_count_shm_name, shape)
    try:
        gs = engine.to_dev(vg.ravel(), torch.float64)
        gc = engine.to_dev(vc.ravel(), torch.float64)
        call("gx_bins_accumulate_f64", engine.ptr(bins.sum), engine.ptr(bins.cnt), q_num ** 3,
             bins.scale, engine.ptr(gs), engine.ptr(gc), engine.stream_handle())
        vg[...] = gs.cpu().numpy().reshape(shape)
        vc[...] = gc.cpu().numpy().reshape(shape)
        del vg, vc
    finally:
        s_shm.close()
        c_shm.close()
