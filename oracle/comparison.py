"""FP64 CPU oracle of the experimental-comparison step of a fit evaluation.

TEST INFRASTRUCTURE ONLY: ``tests/`` and the ``cpu_baseline`` leg of ``bench.py``
use it as the checker (or the timed CPU baseline); ``giwaxsim_amd`` never imports it.

A numpy / scipy restatement of the reference's functions, with the same library
calls the reference makes (scipy.ndimage.map_coordinates, scipy.interpolate.griddata):

* ``trim_sim_data``    tools/comparison.py:161-191
* ``rebin_qmap_and_mask`` tools/comparison.py:406-445
* ``linear_polar``     tools/comparison.py:469-497
* ``polar_linear``     tools/comparison.py:499-532
* ``add_pad``          tools/comparison.py:534-557
* ``shift_peak``       tools/comparison.py:559-592
* ``optimize_scale_offset`` tools/comparison.py:873-882
* ``fit_residual``     the residual of fullfit_model, fit_slabsize.py:19-47
  (after the detector image: trim, shift_peak, scale / offset, masked residual)
* ``slabmaker``        tools/comparison.py:595-671 after the file load and the lattice
  vectors: replicate, bounding box, shift, centred crop, shift (without the last shift:
  the standalone slabmaker.py:33-116)
"""

from __future__ import annotations

import numpy as np
from scipy.interpolate import griddata
from scipy.ndimage import map_coordinates


def trim_sim_data(sim_det_ints, det_h, det_v, exp_qxy, exp_qz):
    """tools/comparison.py:161-191."""
    qxy_min, qxy_max = np.min(exp_qxy), np.max(exp_qxy)
    qz_min, qz_max = np.min(exp_qz), np.max(exp_qz)
    h_mask = (det_h >= qxy_min) & (det_h <= qxy_max)
    v_mask = (det_v >= qz_min) & (det_v <= qz_max)
    return sim_det_ints[v_mask, :][:, h_mask], det_h[h_mask], det_v[v_mask]


def rebin_qmap_and_mask(qxy1, qz1, qmap2, qxy2, qz2, qmap2_mask):
    """tools/comparison.py:406-445: griddata 'linear' of the map and of the mask,
    mask thresholded at > 0."""
    grid1_x, grid1_z = np.meshgrid(qxy1, qz1)
    grid2_x, grid2_z = np.meshgrid(qxy2, qz2)
    points2 = np.vstack((grid2_x.ravel(), grid2_z.ravel())).T
    rebinned = griddata(points2, qmap2.ravel(), (grid1_x, grid1_z), method='linear')
    mask_r = griddata(points2, qmap2_mask.ravel(), (grid1_x, grid1_z), method='linear')
    return rebinned, np.where(mask_r > 0, 1, 0)


def linear_polar(img, o=None, r=None, output=None, order=1, cont=0):
    """tools/comparison.py:469-497 (output shape (round(2 pi r), round(r)))."""
    if o is None:
        o = np.array(img.shape[:2]) / 2 - 0.5
    else:
        o = np.array(o)
    if r is None:
        r = np.sqrt((np.array(img.shape[:2]) ** 2).sum()) / 2
    if output is None:
        shp = int(round(r * 2 * np.pi)), int(round(r))
        output = np.zeros(shp, dtype=img.dtype)
    elif isinstance(output, tuple):
        output = np.zeros(output, dtype=img.dtype)
    out_h, out_w = output.shape
    rs = np.linspace(0, r, out_w)
    ts = np.linspace(0, 2 * np.pi, out_h)
    xs = rs * np.cos(ts[:, None]) + o[1]
    ys = rs * np.sin(ts[:, None]) + o[0]
    map_coordinates(img, (ys, xs), order=order, output=output, cval=cont)
    return output


def polar_linear(img, o=None, r=None, output=None, order=1, cont=0):
    """tools/comparison.py:499-532."""
    if r is None:
        r = img.shape[1]
    if output is None:
        output = np.zeros((int(r * 2), int(r * 2)), dtype=img.dtype)
    elif isinstance(output, tuple):
        output = np.zeros(output, dtype=img.dtype)
    if o is None:
        o = np.array(output.shape) / 2 - 0.5
    else:
        o = np.array(o)
    out_h, out_w = output.shape
    ys, xs = np.mgrid[:out_h, :out_w] - o[:, None, None]
    rs = np.sqrt(ys ** 2 + xs ** 2)
    ts = np.arctan2(ys, xs)
    ts[ts < 0] += 2 * np.pi
    rs = (rs / r) * (img.shape[1] - 1)
    ts = (ts / (2 * np.pi)) * (img.shape[0] - 1)
    map_coordinates(img, (ts, rs), order=order, output=output, cval=cont)
    return output


def add_pad(polar_image, r_axis, pad_width, pad_range):
    """tools/comparison.py:534-557."""
    pad_range_min, pad_range_max = pad_range
    widths = np.diff(r_axis)
    widths_trim = widths[widths > 0]
    pixel_width = np.mean(widths_trim)
    pad_width_pixels = int(np.round(pad_width / pixel_width))
    if pad_width_pixels == 0:
        return polar_image
    pad_start_idx = int(np.argmin(np.abs(r_axis - pad_range_min)))
    pad_end_idx = int(np.argmin(np.abs(r_axis - pad_range_max)))
    pad_range_pixels = pad_end_idx - pad_start_idx
    spacing = int((pad_range_pixels - 1) / pad_width_pixels - 1)
    assert spacing >= 0, 'pad_range is too small for desired pad_width'
    for i in range(pad_width_pixels):
        pad_idx = pad_start_idx + 2 * i + spacing * i
        pad_values = polar_image[:, pad_idx]
        polar_image = np.insert(polar_image, pad_idx + 1, pad_values, axis=1)
    return polar_image[:, 0:-pad_width_pixels]


def shift_peak(image, det_h, det_v, pad_width_qspace, pad_range_qspace):
    """tools/comparison.py:559-592."""
    xx, yy = np.meshgrid(det_h, det_v)
    r, chi = np.hypot(xx, yy), np.arctan2(yy, xx)
    row, col = image.shape
    radius = float(np.sqrt(row ** 2 + col ** 2))
    center_row = int(np.argmin(np.abs(det_v)))
    center_col = int(np.argmin(np.abs(det_h)))
    polar_image = linear_polar(image, o=(center_row, center_col), r=radius, output=None, order=1, cont=0)
    polar_r = linear_polar(r, o=(center_row, center_col), r=radius, output=None, order=1, cont=0)
    r_axis = polar_r[0, :]
    polar_image_pad = add_pad(polar_image.copy(), r_axis, pad_width_qspace, pad_range_qspace)
    polar_image_pad[polar_image == 0] = 0
    return polar_linear(polar_image_pad, o=(center_row, center_col), r=None, output=image.shape)


def optimize_scale_offset(sim_map, rebin_map, rebin_mask):
    """tools/comparison.py:873-882: least squares of rebin = s sim + b over mask == 0."""
    keep = rebin_mask == 0
    A = np.vstack([sim_map[keep], np.ones(int(keep.sum()))]).T
    (s, b), *_ = np.linalg.lstsq(A, rebin_map[keep], rcond=None)
    return s, b


def fit_residual(det_sum, det_h, det_v, rebin_map, rebin_mask, exp_qxy, exp_qz, pad_width, pad_range):
    """fullfit_model after the detector image (fit_slabsize.py:33-46)."""
    sim_trim, h_trim, v_trim = trim_sim_data(det_sum, det_h, det_v, exp_qxy, exp_qz)
    sim_pad = shift_peak(sim_trim.copy(), h_trim, v_trim, pad_width, pad_range)
    s, b = optimize_scale_offset(sim_pad, rebin_map, rebin_mask)
    return ((rebin_map - (s * sim_pad + b)) * (rebin_mask == 0)).ravel(), (s, b)


def slabmaker(coords, elements, x_size, y_size, z_size, a_vect, b_vect, c_vect, shift_again=True):
    """tools/comparison.py:608-671: copies 1..num of the running set appended along a,
    then b, then c (each copy = set + (i + 1) * vector), the bounding box, the shift to
    the origin, the centred box of the requested size and the final shift.  With
    ``shift_again=False``, the standalone slabmaker.py:33-116 (the same replication,
    whose appended element blocks equal the running set's, and no final shift)."""
    num_x = int(np.ceil(2 * x_size / a_vect[0]))
    num_y = int(np.ceil(2 * y_size / b_vect[1]))
    num_z = int(np.ceil(2 * z_size / c_vect[2]))
    for num, vec in ((num_x, (a_vect[0], 0, 0)), (num_y, (b_vect[0], b_vect[1], 0)),
                     (num_z, (c_vect[0], c_vect[1], c_vect[2]))):
        base_c, base_e = coords.copy(), elements.copy()
        for i in range(num):
            step = np.array([vec[0] * (i + 1), vec[1] * (i + 1), vec[2] * (i + 1)])
            coords = np.concatenate((coords, base_c + step), axis=0)
            elements = np.concatenate((elements, base_e), axis=0)
    x_max = np.max(coords[:, 0]) - np.min(coords[:, 0])
    y_max = np.max(coords[:, 1]) - np.min(coords[:, 1])
    z_max = np.max(coords[:, 2]) - np.min(coords[:, 2])
    assert x_max > x_size, "x_max must be greater than x_size"
    assert y_max > y_size, "y_max must be greater than y_size"
    assert z_max > z_size, "z_max must be greater than z_size"
    xb, yb, zb = (x_max - x_size) / 2, (y_max - y_size) / 2, (z_max - z_size) / 2
    for k in range(3):
        coords[:, k] -= np.min(coords[:, k])
    keep = ((coords[:, 0] >= xb) & (coords[:, 0] <= x_max - xb) & (coords[:, 1] >= yb) &
            (coords[:, 1] <= y_max - yb) & (coords[:, 2] >= zb) & (coords[:, 2] <= z_max - zb))
    out, out_e = coords[keep], elements[keep]
    if shift_again:
        for k in range(3):
            out[:, k] -= np.min(out[:, k])
    return out, out_e
