"""FP64 numpy restatement of the reference forward-scattering path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Reference files are cited relative to the reference repository root
(tchaney97/giwaxsim).  Pins for each function live in ``tests/test_oracle_*.py``
(closed forms, invariants, brute force, exact rational arithmetic, library
calls and the reference's own fixtures).  Functions with no independent pin
say "parity unpinned" in their docstring.
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor, as_completed

import numpy as np
from scipy.ndimage import gaussian_filter1d

__all__ = [
    "CROMER_MANN_C", "Z_C",
    "floor_div", "rotate_coords_z_y", "rotation_matrix", "rotate_canonical",
    "voxelgrid_plan", "rectangular_collapse_lengths", "frame_grid",
    "frame_axes", "frame_iq", "bin_frame", "downselect_voxelgrid",
    "add_f0_q_3d", "voxelgridmaker", "make_detector", "rotate_about_normal",
    "rotate_about_horizontal", "rotate_about_vertical", "intersect_detector",
    "rotate_psi_phi_theta", "mirror_vertical_horizontal", "detectormaker",
    "optimize_scale_offset", "avg_voxel_f_of", "default_threads",
]

# Cromer-Mann coefficients of carbon (a1,b1,a2,b2,a3,b3,a4,b4,c), International
# Tables for Crystallography Vol. C, Table 6.1.1.4; the reference hard-codes
# element 'C' for the f0(q) factor (tools/comparison.py:785-786,
# tools/ptable_dict.py:56).
CROMER_MANN_C = (2.31, 20.8439, 1.02, 10.2075, 1.5886, 0.5687, 0.865, 51.6512, 0.2156)
Z_C = 6


# --------------------------------------------------------------------------------------
# index arithmetic
# --------------------------------------------------------------------------------------
def floor_div(a, b):
    """numpy float ``//`` exactly as the reference uses it for every index
    (tools/voxelgrids.py:328-329, 497-499; tools/detector.py:217-219).

    numpy's float floor-divide is fmod based and returns the exact floor of the
    real quotient of the two doubles (``1.0 // 0.2 == 4.0`` although
    ``floor(1.0/0.2) == 5.0``); pinned against exact rational arithmetic in
    ``tests/test_oracle_index.py``.
    """
    return np.floor_divide(a, b)


def rotate_coords_z_y(coords, phi):
    """Rotated y of ``rotate_coords_z`` (tools/utilities.py:303-317).

    Reference: ``np.dot(coords, R.T)`` with R = [[c,-s,0],[s,c,0],[0,0,1]], so
    y' = x*s + y*c + z*0.  Canonical form: ``x*s + y*c`` evaluated elementwise
    (two roundings for the products, one for the sum, no FMA).  z is returned
    unchanged by the reference rotation (z' = 0*x + 0*y + 1*z == z exactly).
    """
    phi_rad = np.radians(phi)
    s = np.sin(phi_rad)
    c = np.cos(phi_rad)
    return coords[:, 0] * s + coords[:, 1] * c


def rotation_matrix(u, theta):
    """Rodrigues rotation matrix, tools/utilities.py:222-245 (same op order)."""
    ux, uy, uz = u[0], u[1], u[2]
    R = np.zeros((3, 3))
    ct = np.cos(theta)
    st = np.sin(theta)
    R[0, 0] = ct + ux ** 2 * (1 - ct)
    R[0, 1] = ux * uy * (1 - ct) - uz * st
    R[0, 2] = ux * uz * (1 - ct) + uy * st
    R[1, 0] = uy * ux * (1 - ct) + uz * st
    R[1, 1] = ct + uy ** 2 * (1 - ct)
    R[1, 2] = uy * uz * (1 - ct) - ux * st
    R[2, 0] = uz * ux * (1 - ct) - uy * st
    R[2, 1] = uz * uy * (1 - ct) + ux * st
    R[2, 2] = ct + uz ** 2 * (1 - ct)
    return R


def rotate_canonical(R, x, y, z):
    """``R @ vstack([x, y, z])`` (tools/detector.py:71,113,155) in canonical
    order: row i = (R[i,0]*x + R[i,1]*y) + R[i,2]*z, elementwise, no FMA."""
    xr = R[0, 0] * x + R[0, 1] * y + R[0, 2] * z
    yr = R[1, 0] * x + R[1, 1] * y + R[1, 2] * z
    zr = R[2, 0] * x + R[2, 1] * y + R[2, 2] * z
    return xr, yr, zr


# --------------------------------------------------------------------------------------
# voxel-grid driver planning (tools/comparison.py:705-744)
# --------------------------------------------------------------------------------------
def voxelgrid_plan(coords, r_voxel_size, q_voxel_size, max_q):
    """Sizes and axes of ``voxelgridmaker_fitting`` (tools/comparison.py:705-731).

    Returns dict with grid_size N, q_num, q axis, phis (degrees), bounds.
    """
    max_q_diag = np.sqrt(2) * max_q
    if max_q_diag > 2 * np.pi / r_voxel_size:
        raise Exception('Max_q is non-physical for given voxel size')
    grid_size = int(np.ceil(2 * np.pi / (q_voxel_size * r_voxel_size)))
    x_bound = np.max(coords[:, 0]) - np.min(coords[:, 0])
    y_bound = np.max(coords[:, 1]) - np.min(coords[:, 1])
    z_bound = np.max(coords[:, 2]) - np.min(coords[:, 2])
    min_bound = np.min([x_bound, y_bound, z_bound])
    if grid_size * r_voxel_size < min_bound:
        raise Exception('Calculated real-space bounds smaller than simulation. '
                        'Please lower delta_q value')
    max_q_diag = max_q_diag + max_q_diag % q_voxel_size
    q_num = ((2 * max_q_diag / q_voxel_size) + 1).astype(int)
    if q_num % 2 == 0:
        q_num += 1
    q_axis = np.linspace(-max_q_diag, max_q_diag, q_num)
    delta_phi_rad = np.arctan(q_voxel_size / max_q_diag)
    phi_num = np.ceil(2 * np.pi / delta_phi_rad).astype(int)
    last_phi = 180 - (180 / phi_num)
    phis = np.linspace(0, last_phi, num=phi_num)
    return dict(grid_size=grid_size, q_num=int(q_num), q_axis=q_axis, phis=phis,
                max_q_diag=max_q_diag, x_bound=x_bound, y_bound=y_bound,
                z_bound=z_bound)


def avg_voxel_f_of(f_values, x_bound, y_bound, z_bound, r_voxel_size):
    """tools/comparison.py:741-744."""
    sum_f_values = np.sum(f_values)
    volume = x_bound * y_bound * z_bound
    return (sum_f_values / volume) * r_voxel_size ** 3


# --------------------------------------------------------------------------------------
# per-frame projection (tools/voxelgrids.py:238-416)
# --------------------------------------------------------------------------------------
def rectangular_collapse_lengths(x_vals, hor_length, ver_length, phi):
    """Path length of a vertical line through the phi-rotated rectangle,
    tools/voxelgrids.py:238-309 (branch structure and op order kept; the
    per-x Python loop is evaluated as masked numpy expressions with the same
    elementwise operations)."""
    x_vals = np.asarray(x_vals, dtype=np.float64)
    if phi > 90:
        phi = phi - 90
        hor_length, ver_length = ver_length, hor_length
    theta = 90 - phi
    theta_rad = np.deg2rad(theta)
    phi_rad = np.deg2rad(phi)
    if phi == 0:
        return np.where(x_vals < hor_length, ver_length, 0).astype(np.float64)
    if phi == 90:
        return np.where(x_vals < ver_length, hor_length, 0).astype(np.float64)
    stop1 = ver_length * np.cos(theta_rad)
    stop2 = hor_length * np.cos(phi_rad)
    if stop1 < stop2:
        mid_length = ver_length / np.sin(theta_rad)
    else:
        stop1, stop2 = stop2, stop1
        mid_length = hor_length / np.sin(phi_rad)
    A = np.sqrt(ver_length ** 2 - (ver_length * np.cos(theta_rad)) ** 2)
    tp = np.tan(phi_rad)
    vc = ver_length * np.cos(theta_rad)
    tt = np.tan(theta_rad)
    cp = np.cos(phi_rad)
    ct = np.cos(theta_rad)
    out = np.zeros_like(x_vals)
    x = x_vals
    with np.errstate(all='ignore'):
        r1 = (A + x * tp) - ((vc - x) * tt)
        r3 = (hor_length - ((x - vc) / cp)) / ct
    s12 = stop1 + stop2
    c0 = x == 0
    c1 = ~c0 & (x < stop1)
    c2 = ~c0 & ~c1 & (x <= stop2)
    c3 = ~c0 & ~c1 & ~c2 & (x < s12)
    out[c1] = r1[c1]
    out[c2] = mid_length
    out[c3] = r3[c3]
    return out


def frame_grid(coords, f_values, phi, grid_size, r_voxel_size, avg_voxel_f,
               x_bound, y_bound, fill_bkg, smooth, return_stats=False):
    """Real-space projected grid of one phi-frame, tools/voxelgrids.py:319-379.

    Canonical rotation (see ``rotate_coords_z_y``); everything else in the
    reference's order.  Returns complex128 (N, N) indexed [z, y].
    """
    N = grid_size
    yr = rotate_coords_z_y(coords, phi)
    yr = yr - np.min(yr)
    zr = coords[:, 2] - np.min(coords[:, 2])
    y_indices = floor_div(yr, r_voxel_size).astype(int)
    z_indices = floor_div(zr, r_voxel_size).astype(int)
    valid_mask = (y_indices >= 0) & (y_indices < N) & (z_indices >= 0) & (z_indices < N)
    y_indices = y_indices[valid_mask]
    z_indices = z_indices[valid_mask]
    valid_fs = f_values[valid_mask]
    grid = np.zeros((N, N), dtype=complex)
    np.add.at(grid, (z_indices, y_indices), valid_fs)

    max_voxels = np.sqrt(x_bound ** 2 + y_bound ** 2) // r_voxel_size
    y_min = np.min(y_indices)
    z_min = np.min(z_indices)
    y_max = np.max(y_indices)
    z_max = np.max(z_indices)
    if fill_bkg:
        x_vals = np.arange(N) * r_voxel_size
        lengths = rectangular_collapse_lengths(x_vals, y_bound, x_bound, phi)
        num_missing_voxels = (max_voxels - floor_div(lengths, r_voxel_size)).astype(int)
        amorphous = np.tile(num_missing_voxels * avg_voxel_f, (N, 1))
        grid = grid + amorphous
        grid[:z_min, :] = avg_voxel_f * max_voxels
        grid[z_max:, :] = avg_voxel_f * max_voxels
        grid[z_min:z_max + 1, :y_min + 1] = avg_voxel_f * max_voxels
        grid[z_min:z_max + 1, y_max:] = avg_voxel_f * max_voxels
    if smooth:
        avg_val = avg_voxel_f * max_voxels
        sigma = smooth
        mask_y = np.zeros(grid.shape[1])
        mask_z = np.zeros(grid.shape[0])
        mask_y[y_min + sigma:y_max - sigma] = 1
        mask_z[z_min + sigma:z_max - sigma] = 1
        sm_y = gaussian_filter1d(mask_y, sigma=sigma, mode='wrap')
        sm_z = gaussian_filter1d(mask_z, sigma=sigma, mode='wrap')
        m = np.outer(sm_z, sm_y)
        grid = grid * m + avg_val * (1 - m)
    if return_stats:
        return grid, dict(y_min=int(y_min), y_max=int(y_max), z_min=int(z_min),
                          z_max=int(z_max), max_voxels=max_voxels,
                          y_idx=y_indices, z_idx=z_indices)
    return grid


def frame_axes(grid_size, r_voxel_size, phi):
    """Per-frame q coordinates, tools/voxelgrids.py:381-402.

    Returns (det_h_qx, det_h_qy, det_v_qz, q_h_shifted)."""
    q_h = np.fft.fftfreq(grid_size, d=r_voxel_size) * 2 * np.pi
    q_v = np.fft.fftfreq(grid_size, d=r_voxel_size) * 2 * np.pi
    q_h_shifted = np.fft.fftshift(q_h)
    q_v_shifted = np.fft.fftshift(q_v)
    right_qy = np.max(q_h_shifted) * np.cos(np.deg2rad(-phi))
    left_qy = -right_qy
    right_qx = -np.max(q_h_shifted) * np.sin(np.deg2rad(-phi))
    left_qx = -right_qx
    det_h_qy = np.linspace(left_qy, right_qy, num=len(q_h_shifted))
    det_h_qx = np.linspace(left_qx, right_qx, num=len(q_h_shifted))
    return det_h_qx, det_h_qy, q_v_shifted, q_h_shifted


def frame_iq(grid):
    """|fftshift(fftn(grid))|^2, tools/voxelgrids.py:387-392."""
    ft = np.fft.fftshift(np.fft.fftn(grid))
    return np.abs(ft) ** 2


def bin_frame(iq_2d, det_h_qx, det_h_qy, det_v_qz, qx, qy, qz, voxel_grid, voxel_grid_count):
    """Accumulate one frame into the 3D bins, tools/voxelgrids.py:474-503.

    ``voxel_grid`` / ``voxel_grid_count`` are (q_num,)*3 float64 arrays
    indexed [qy, qx, qz], updated in place with ``np.add.at``."""
    det_h_mask = ((det_h_qx <= np.max(qx)) & (det_h_qx >= np.min(qx)) &
                  (det_h_qy <= np.max(qy)) & (det_h_qy >= np.min(qy)))
    det_v_mask = (det_v_qz <= np.max(qz)) & (det_v_qz >= np.min(qz))
    det_h_qx = det_h_qx[det_h_mask]
    det_h_qy = det_h_qy[det_h_mask]
    det_v_qz = det_v_qz[det_v_mask]
    iq_2d = iq_2d[det_v_mask, :][:, det_h_mask]
    qx_mesh, qz_mesh = np.meshgrid(det_h_qx, det_v_qz)
    qy_mesh, qz_mesh = np.meshgrid(det_h_qy, det_v_qz)
    qx_vals = np.ravel(qx_mesh)
    qy_vals = np.ravel(qy_mesh)
    qz_vals = np.ravel(qz_mesh)
    iq_vals = np.ravel(iq_2d)
    actual_q_voxel = np.diff(qz)[0]
    qx_i = floor_div(qx_vals - np.min(qx), actual_q_voxel).astype(int)
    qy_i = floor_div(qy_vals - np.min(qy), actual_q_voxel).astype(int)
    qz_i = floor_div(qz_vals - np.min(qz), actual_q_voxel).astype(int)
    np.add.at(voxel_grid, (qy_i, qx_i, qz_i), iq_vals)
    np.add.at(voxel_grid_count, (qy_i, qx_i, qz_i), np.ones_like(iq_vals))
    return qy_i, qx_i, qz_i


def downselect_voxelgrid(grid, x_axis, y_axis, z_axis, max_val):
    """tools/voxelgrids.py:16-48."""
    voxel_size = np.abs(x_axis[1] - x_axis[0])
    max_val = max_val + voxel_size
    x_idxs = np.where(np.abs(x_axis) < max_val)[0]
    y_idxs = np.where(np.abs(y_axis) < max_val)[0]
    z_idxs = np.where(np.abs(z_axis) < max_val)[0]
    small = grid[y_idxs[0]:y_idxs[-1] + 1, x_idxs[0]:x_idxs[-1] + 1, z_idxs[0]:z_idxs[-1] + 1]
    return (small, x_axis[x_idxs[0]:x_idxs[-1] + 1], y_axis[y_idxs[0]:y_idxs[-1] + 1],
            z_axis[z_idxs[0]:z_idxs[-1] + 1])


def add_f0_q_3d(iq, qx_axis, qy_axis, qz_axis, aff=CROMER_MANN_C, Z=Z_C):
    """tools/voxelgrids.py:828-857 (element 'C' as called from
    tools/comparison.py:785-786)."""
    iq_new = np.copy(iq)
    qx, qy, qz = np.meshgrid(qx_axis, qy_axis, qz_axis)
    q2 = qx ** 2 + qy ** 2 + qz ** 2
    fq_norm = ((aff[0] * np.exp(-aff[1] * q2 / (16 * np.pi ** 2)) +
                aff[2] * np.exp(-aff[3] * q2 / (16 * np.pi ** 2)) +
                aff[4] * np.exp(-aff[5] * q2 / (16 * np.pi ** 2)) +
                aff[6] * np.exp(-aff[7] * q2 / (16 * np.pi ** 2)) +
                aff[8]) / Z) ** 2
    iq_new *= fq_norm
    return iq_new


def _one_frame(coords, f_values, phi, N, r, avg, xb, yb, fill_bkg, smooth, qx, qy, qz,
               vg, vc):
    grid = frame_grid(coords, f_values, phi, N, r, avg, xb, yb, fill_bkg, smooth)
    iq2 = frame_iq(grid)
    hx, hy, vz, _ = frame_axes(N, r, phi)
    bin_frame(iq2, hx, hy, vz, qx, qy, qz, vg, vc)


def voxelgridmaker(coords, f_values, r_voxel_size, q_voxel_size, max_q,
                   fill_bkg=False, smooth=0, phis=None, threads=0):
    """``voxelgridmaker_fitting`` (tools/comparison.py:673-788).

    ``f_values`` is the complex per-atom f = Z + f1 + i f2 (the reference
    builds it with xraydb, tools/comparison.py:735-739; that lookup is parity
    unpinned and is an input here).  ``threads=0`` runs frames sequentially in
    phi order (canonical); ``threads=None`` uses ``ThreadPoolExecutor()`` with
    its default worker count exactly as the reference (CPU baseline mode);
    ``phis`` restricts the run to a subset of frames (bounded samples).

    Returns dict(iq, qx, qy, qz, sum, count, plan).
    """
    plan = voxelgrid_plan(coords, r_voxel_size, q_voxel_size, max_q)
    N = plan['grid_size']
    q_num = plan['q_num']
    qx = qy = qz = plan['q_axis']
    avg = avg_voxel_f_of(f_values, plan['x_bound'], plan['y_bound'], plan['z_bound'],
                         r_voxel_size)
    vg = np.zeros((q_num,) * 3)
    vc = np.zeros((q_num,) * 3)
    run_phis = plan['phis'] if phis is None else np.asarray(phis)
    args = [(coords, f_values, phi, N, r_voxel_size, avg, plan['x_bound'],
             plan['y_bound'], fill_bkg, smooth, qx, qy, qz, vg, vc) for phi in run_phis]
    if threads == 0:
        for a in args:
            _one_frame(*a)
    else:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            futs = [ex.submit(_one_frame, *a) for a in args]
            for fu in as_completed(futs):
                fu.result()
    mean = np.divide(vg, vc, out=np.zeros_like(vg), where=vc != 0)
    small, qxs, qys, qzs = downselect_voxelgrid(mean, qx, qy, qz, max_q)
    iq = add_f0_q_3d(small, qxs, qys, qzs)
    return dict(iq=iq, qx=qxs, qy=qys, qz=qzs, sum=vg, count=vc, plan=plan, avg=avg)


# --------------------------------------------------------------------------------------
# detector (tools/detector.py)
# --------------------------------------------------------------------------------------
def make_detector(h_max, num_pixels_h, v_max, num_pixels_v):
    """tools/detector.py:5-29."""
    h = np.linspace(-h_max, h_max, num_pixels_h)
    v = np.linspace(-v_max, v_max, num_pixels_v)
    det_y, det_z = np.meshgrid(h, v)
    det_x = np.zeros_like(det_y)
    return det_x, det_y, det_z, h, v


def _corners(x, y, z):
    p1 = np.array([x[0, 0], y[0, 0], z[0, 0]])
    p2 = np.array([x[0, -1], y[0, -1], z[0, -1]])
    p3 = np.array([x[-1, 0], y[-1, 0], z[-1, 0]])
    return p1, p2, p3


def _apply(R, x, y, z):
    xr, yr, zr = rotate_canonical(R, x.ravel(), y.ravel(), z.ravel())
    return xr.reshape(x.shape), yr.reshape(y.shape), zr.reshape(z.shape)


def normal_matrix(x, y, z, psi):
    """Axis + matrix of tools/detector.py:49-68."""
    p1, p2, p3 = _corners(x, y, z)
    n = np.cross(p2 - p1, p3 - p1)
    n /= np.linalg.norm(n)
    return rotation_matrix(n, np.radians(psi))


def horizontal_matrix(x, y, z, theta):
    """tools/detector.py:96-110."""
    p1, p2, _ = _corners(x, y, z)
    h = p2 - p1
    h /= np.linalg.norm(h)
    return rotation_matrix(h, np.radians(theta))


def vertical_matrix(x, y, z, phi):
    """tools/detector.py:138-152 (axis = column 0 end - start)."""
    p1, _, p3 = _corners(x, y, z)
    v = p3 - p1
    v /= np.linalg.norm(v)
    return rotation_matrix(v, np.radians(phi))


def rotate_about_normal(x, y, z, psi):
    """tools/detector.py:33-78 (canonical matmul)."""
    return _apply(normal_matrix(x, y, z, psi), x, y, z)


def rotate_about_horizontal(x, y, z, theta):
    """tools/detector.py:80-120 (canonical matmul)."""
    return _apply(horizontal_matrix(x, y, z, theta), x, y, z)


def rotate_about_vertical(x, y, z, phi):
    """tools/detector.py:122-162 (canonical matmul)."""
    return _apply(vertical_matrix(x, y, z, phi), x, y, z)


def intersect_detector(iq, qx, qy, qz, det_x, det_y, det_z):
    """The effective (second) definition, tools/detector.py:194-232."""
    dq = np.diff(qz)[0]
    xi = floor_div(det_x.ravel() - np.min(qx), dq).astype(int)
    yi = floor_div(det_y.ravel() - np.min(qy), dq).astype(int)
    zi = floor_div(det_z.ravel() - np.min(qz), dq).astype(int)
    xi = np.clip(xi, 0, iq.shape[1] - 1)
    yi = np.clip(yi, 0, iq.shape[0] - 1)
    zi = np.clip(zi, 0, iq.shape[2] - 1)
    return iq[yi, xi, zi].reshape(det_x.shape)


def rotate_psi_phi_theta(x, y, z, psi, phi, theta):
    """tools/detector.py:234-244."""
    x, y, z = rotate_about_normal(x, y, z, psi)
    x, y, z = rotate_about_vertical(x, y, z, phi)
    x, y, z = rotate_about_horizontal(x, y, z, theta)
    return x, y, z


def mirror_vertical_horizontal(qmap):
    """tools/detector.py:246-275."""
    lr = np.fliplr(qmap)
    ud = np.flipud(qmap)
    udlr = np.fliplr(ud)
    s = qmap + lr + ud + udlr
    if qmap.shape[0] % 2 != 0:
        c = qmap.shape[0] // 2
        s[c, :] = (qmap[c, :] + qmap[c, :][::-1]) * 2
    if qmap.shape[1] % 2 != 0:
        c = qmap.shape[1] // 2
        s[:, c] = (qmap[:, c] + qmap[:, c][::-1]) * 2
    if qmap.shape[1] % 2 != 0 and qmap.shape[0] % 2 != 0:
        s[qmap.shape[0] // 2, qmap.shape[1] // 2] = qmap[qmap.shape[0] // 2, qmap.shape[1] // 2] * 4
    return s


def _init_rotations(det_x, det_y, det_z, angle_init_vals, angle_init_axs):
    """tools/comparison.py:799-818 (each slot may apply psi, phi or theta)."""
    for val, ax in zip(angle_init_vals, angle_init_axs):
        if ax == 'psi':
            det_x, det_y, det_z = rotate_about_normal(det_x, det_y, det_z, val)
        if ax == 'phi':
            det_x, det_y, det_z = rotate_about_vertical(det_x, det_y, det_z, val)
        if ax == 'theta':
            det_x, det_y, det_z = rotate_about_horizontal(det_x, det_y, det_z, val)
    return det_x, det_y, det_z


def detectormaker(iq, qx, qy, qz, num_pixels, max_q, angle_init_vals, angle_init_axs,
                  psis, psi_weights, phis, phi_weights, thetas, theta_weights,
                  mirror=True, threads=0):
    """``detectormaker_fitting`` (tools/comparison.py:790-870).

    Weights are arrays (the reference loads them from ``*_weights_path`` or
    uses uniform 1/len, 821-823).  ``threads=0`` sums orientations
    sequentially in (psi outer, phi, theta inner) order; ``threads=None``
    reproduces the reference's ``ThreadPoolExecutor()`` fan-out (with a lock
    so the sum is not racy) for CPU-baseline timing.
    """
    det_x, det_y, det_z, det_h, det_v = make_detector(max_q, num_pixels, max_q, num_pixels)
    det_x, det_y, det_z = _init_rotations(det_x, det_y, det_z, angle_init_vals, angle_init_axs)
    psi_w = np.ones_like(psis) / len(psis) if psi_weights is None else np.asarray(psi_weights)
    phi_w = np.ones_like(phis) / len(phis) if phi_weights is None else np.asarray(phi_weights)
    th_w = np.ones_like(thetas) / len(thetas) if theta_weights is None else np.asarray(theta_weights)
    assert len(psis) == len(psi_w), 'psi weights length must equal psi_num'
    assert len(phis) == len(phi_w), 'phi weights length must equal phi_num'
    assert len(thetas) == len(th_w), 'theta weights length must equal theta_num'
    assert np.abs(1 - np.sum(psi_w)) < 0.01, 'psi weights must sum to 1'
    assert np.abs(1 - np.sum(phi_w)) < 0.01, 'phi weights must sum to 1'
    assert np.abs(1 - np.sum(th_w)) < 0.01, 'theta weights must sum to 1'
    det_sum = np.zeros((num_pixels, num_pixels))
    orients = [(ps, pw, ph, phw, th, tw)
               for ps, pw in zip(psis, psi_w)
               for ph, phw in zip(phis, phi_w)
               for th, tw in zip(thetas, th_w)]

    def one(o):
        ps, pw, ph, phw, th, tw = o
        x2, y2, z2 = rotate_psi_phi_theta(det_x, det_y, det_z, ps, ph, th)
        d = intersect_detector(iq, qx, qy, qz, x2, y2, z2)
        d *= pw * phw * tw
        return d

    if threads == 0:
        for o in orients:
            det_sum += one(o)
    else:
        import threading
        lock = threading.Lock()

        def work(o):
            d = one(o)
            with lock:
                det_sum[...] += d
        with ThreadPoolExecutor(max_workers=threads) as ex:
            futs = [ex.submit(work, o) for o in orients]
            for fu in as_completed(futs):
                fu.result()
    if mirror:
        det_sum = mirror_vertical_horizontal(det_sum)
    det_sum[det_sum != det_sum] = 1e-6
    det_sum[det_sum <= 0] = 1e-6
    det_sum *= 1e-6
    return det_sum, det_h, det_v


def optimize_scale_offset(sim_map, rebin_map, rebin_mask):
    """tools/comparison.py:873-882."""
    s = sim_map[rebin_mask == 0].ravel()
    y = rebin_map[rebin_mask == 0].ravel()
    A = np.vstack([s, np.ones_like(s)]).T
    coeffs, _, _, _ = np.linalg.lstsq(A, y, rcond=None)
    return coeffs[0], coeffs[1]


def default_threads():
    """``ThreadPoolExecutor()`` default worker count (Python 3.8+)."""
    return min(32, (os.cpu_count() or 1) + 4)
