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
           'process_file2', 'process_file', 'frames_to_iq_parallel', 'add_f0_q_3d', 'generate_voxel_grid_low_mem',
           'structure_voxel_grid', 'shell_f0']


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


def _bin_frame_dev(iq_2d, det_h_qx, det_h_qy, det_v_qz, qx, qy, qz, gs, gc, status):
    """Enqueue the binning of one frame into the device sum / count grids gs, gc
    (flattened [q_num]^3); pixels outside the grid are counted in status."""
    import torch
    iq_2d = np.asarray(iq_2d, dtype=np.float64)
    if iq_2d.ndim != 2:
        raise ValueError(f"frame image must be 2-D, got shape {iq_2d.shape}")
    nv, nh = iq_2d.shape
    if not (np.size(det_h_qx) == nh and np.size(det_h_qy) == nh and np.size(det_v_qz) == nv):
        raise ValueError(f"frame axes ({np.size(det_h_qx)}, {np.size(det_h_qy)}, {np.size(det_v_qz)}) do not "
                         f"match the image's {nh} columns and {nv} rows")
    if len(qx) < 2 or gs.numel() != len(qx) ** 3 or gc.numel() != len(qx) ** 3:
        raise ValueError("sum / count grids must hold len(qx)^3 voxels")
    iq = engine.to_dev(iq_2d, torch.float64)
    hx = engine.to_dev(det_h_qx, torch.float64)
    hy = engine.to_dev(det_h_qy, torch.float64)
    vz = engine.to_dev(det_v_qz, torch.float64)
    qlim = dbl_array([np.min(qx), np.max(qx), np.min(qy), np.max(qy), np.min(qz), np.max(qz)])
    monotone = all(_sorted(a) for a in (det_h_qx, det_h_qy, det_v_qz))
    call("gx_bin_frame_f64", engine.ptr(iq), nv, nh, engine.ptr(hx), engine.ptr(hy),
         engine.ptr(vz), qlim, float(np.diff(qz)[0]), len(qx), engine.ptr(gs), engine.ptr(gc),
         engine.ptr(status), int(monotone), engine.stream_handle())


def _sorted(a):
    """True if ``a`` is non-decreasing or non-increasing (then every floor-divided bin
    index along it is too, and each bin's pixels form one rectangle)."""
    d = np.diff(np.asarray(a, dtype=np.float64).ravel())
    return bool(np.all(d >= 0) or np.all(d <= 0))


def _check_status(status):
    n = int(status.item())
    if n:
        raise IndexError(f"{n} pixel(s) index outside the voxel grid")


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
        gs = engine.to_dev(vg.ravel(), torch.float64)
        gc = engine.to_dev(vc.ravel(), torch.float64)
        status = torch.zeros(1, dtype=torch.int32, device=engine.device())
        _bin_frame_dev(iq_2d, det_h_qx, det_h_qy, det_v_qz, qx, qy, qz, gs, gc, status)
        _check_status(status)
        vg[...] = gs.cpu().numpy().reshape(shape)
        vc[...] = gc.cpu().numpy().reshape(shape)
        del vg, vc
    finally:
        s_shm.close()
        c_shm.close()


def _load_frame(filepath):
    """(iq, h_qx, h_qy, v_qz) of a frame saved as {filepath}_iq|h_qx|h_qy|v_qz.npy."""
    return tuple(np.load(f'{filepath}_{k}.npy', allow_pickle=False) for k in ('iq', 'h_qx', 'h_qy', 'v_qz'))


def process_file(filepath, q_num, qx, qy, qz, voxel_grid_shm_name, voxel_grid_count_shm_name):
    """Bin one saved frame ({filepath}_iq.npy, _h_qx.npy, _h_qy.npy, _v_qz.npy) into the
    named (q_num,)^3 shared grids, as process_file2 bins an in-memory frame
    (tools/voxelgrids.py:418-462)."""
    if int(q_num) != len(qx):
        raise ValueError(f"q_num {q_num} != len(qx) {len(qx)}")
    process_file2(*_load_frame(filepath), qx, qy, qz, voxel_grid_shm_name, voxel_grid_count_shm_name)


def frames_to_iq_parallel(filepaths, q_num, qx, qy, qz):
    """Mean I(q) [qy, qx, qz] over saved frames (tools/voxelgrids.py:508-533): every frame
    binned into one pair of device sum / count grids, then sum / count with 0 where no
    pixel landed.  Frames are binned one after another in the given order, each with
    np.add.at's order of additions (gx_bin_frame_f64 with sorted axes), so the result is
    deterministic; the reference's thread pool adds frames in completion order."""
    import torch
    _same_axes(qx, qy, qz)
    if int(q_num) != len(qx):
        raise ValueError(f"q_num {q_num} != len(qx) {len(qx)}")
    dev = engine.device()
    n = int(q_num) ** 3
    gs = torch.zeros(n, dtype=torch.float64, device=dev)
    gc = torch.zeros(n, dtype=torch.float64, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    for fp in filepaths:
        _bin_frame_dev(*_load_frame(fp), qx, qy, qz, gs, gc, status)
    _check_status(status)
    iq = torch.where(gc != 0, gs / torch.where(gc != 0, gc, 1.0), 0.0)
    return iq.cpu().numpy().reshape((int(q_num),) * 3)


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


# --------------------------------------------------------------------------------------
# whole-structure voxel grid (tools/voxelgrids.py:535-722)
# --------------------------------------------------------------------------------------
def shell_f0(q_val, element):
    """f0 of get_element_f0_dict (tools/utilities.py:318-339): the Cromer-Mann sum with
    q_val where the formula has q^2."""
    a = aff_dict[element]
    return (a[0] * np.exp(-a[1] * (q_val) / (16 * np.pi ** 2)) +
            a[2] * np.exp(-a[3] * (q_val) / (16 * np.pi ** 2)) +
            a[4] * np.exp(-a[5] * (q_val) / (16 * np.pi ** 2)) +
            a[6] * np.exp(-a[7] * (q_val) / (16 * np.pi ** 2)) +
            a[8])


def structure_voxel_grid(input_path, r_voxel_size, q_voxel_size, max_q, aff_num_qs, energy, fill_bkg=False,
                         smooth=0, crop=True):
    """(iq [qy, qx, qz] FP64 numpy, qx, qy, qz) of one .xyz / .pdb file, without the
    f0 factor: the full q_num^3 grid of generate_voxel_grid_low_mem (crop=False) or
    that grid cropped by downselect_voxelgrid (crop=True), from the device frame loop.

    ``aff_num_qs > 1`` follows the reference's q-shell branch (tools/voxelgrids.py:650-707)
    as it computes: each shell's pass overwrites the grid, so the result is the last
    shell's grid (f = f0(q_shell) + f1 + i f2, f0 evaluated at q, not q^2) with the voxels
    of that shell doubled.  Only that last pass is run."""
    from .utilities import load_pdb, load_xyz
    if input_path[-3:] == 'xyz':
        coords, elements = load_xyz(input_path)
    elif input_path[-3:] == 'pdb':
        coords, elements = load_pdb(input_path)
    else:
        raise Exception('files must be a .pdb or .xyz file')
    if aff_num_qs == 1:
        res = engine.run_voxelgrid(coords, elements, r_voxel_size, q_voxel_size, max_q, energy,
                                   fill_bkg=fill_bkg, smooth=int(smooth), f0_element=None, crop=crop)
        return res['iq64'].cpu().numpy(), res['qx'], res['qy'], res['qz']
    if aff_num_qs > 1:
        from .f1f2 import f1_f2
        ids, _, _, uniq = engine.element_table(elements, energy)
        grid = engine.qgrid(float(q_voxel_size), float(max_q))
        step = grid.max_q_diag / int(aff_num_qs)
        q_val = 0.5 * step + (int(aff_num_qs) - 1) * step    # the last shell's centre
        upper_q, lower_q = q_val + step / 2, q_val - step / 2
        f_table = np.array([shell_f0(q_val, str(e)) + complex(*f1_f2(str(e), energy)) for e in uniq],
                           dtype=np.complex128)
        res = engine.run_voxelgrid(coords, None, r_voxel_size, q_voxel_size, max_q, energy, fill_bkg=fill_bkg,
                                   smooth=int(smooth), f_table=f_table, elem_ids=ids, f0_element=None, crop=crop)
        iq = res['iq64'].cpu().numpy()
        qx, qy, qz = res['qx'], res['qy'], res['qz']
        qxm, qym, qzm = np.meshgrid(qx, qy, qz)
        qr = np.sqrt(qxm ** 2 + qym ** 2 + qzm ** 2)
        iq += np.where((qr <= upper_q) & (qr > lower_q), iq, 0)
        return iq, qx, qy, qz
    raise Exception('Invalid aff_num_qs value. Must be non-negative integer')


def generate_voxel_grid_low_mem(input_path, r_voxel_size, q_voxel_size, max_q, aff_num_qs, energy, gen_name,
                                output_dir=None, scratch_folder=None, num_cpus=None, fill_bkg=False, smooth=0):
    """I(q) voxel grid of one structure file on the full q_num^3 grid, all phi-frames on
    the device (tools/voxelgrids.py:535-722): returns (iq [qy, qx, qz], qx, qy, qz), or
    with ``output_dir`` writes ``{output_dir}/{gen_name}_output_files/{gen_name}_iq|qx|qy|qz.npy``
    and returns None.  Voxels no frame reaches are 0.  ``scratch_folder`` and ``num_cpus``
    are accepted and unused."""
    import os
    iq, qx, qy, qz = structure_voxel_grid(input_path, r_voxel_size, q_voxel_size, max_q, aff_num_qs, energy,
                                          fill_bkg=fill_bkg, smooth=smooth, crop=False)
    if not output_dir:
        return iq, qx, qy, qz
    save_path = f'{output_dir}/{gen_name}_output_files/'
    if not os.path.exists(save_path):
        os.mkdir(save_path)
    np.save(f'{save_path}{gen_name}_iq.npy', iq)
    np.save(f'{save_path}{gen_name}_qx.npy', qx)
    np.save(f'{save_path}{gen_name}_qy.npy', qy)
    np.save(f'{save_path}{gen_name}_qz.npy', qz)
    return None
