"""Mirror of the reference's pipeline drivers (``tools/comparison.py``):
``voxelgridmaker_fitting``, ``detectormaker_fitting``,
``optimize_scale_offset``, the slab builder ``slabmaker_fitting``, and the
experimental-comparison step of a fit evaluation (``trim_sim_data``,
``rebin_qmap_and_mask``, ``linear_polar`` / ``polar_linear`` / ``add_pad`` /
``shift_peak``, and ``FitComparison`` for fit_slabsize.py's residual).

The two drivers run every frame / orientation on the GPU with all
intermediate data device resident, and return numpy arrays in the
reference's layouts: iq [qy, qx, qz] and det_sum [v, h], both FP64.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import engine
from .._lib import GX_MOMENTS_WORK, GxSlab, call
from .ptable_dict import aff_dict, ptable
from .utilities import calc_real_space_abc, load_pdb, load_xyz

__all__ = ['slabmaker_fitting', 'voxelgridmaker_fitting', 'detectormaker_fitting',
           'optimize_scale_offset', 'trim_sim_data', 'rebin_qmap_and_mask',
           'linear_polar', 'polar_linear', 'add_pad', 'shift_peak', 'pad_columns', 'ShiftPeakPlan',
           'RebinWeights', 'rebin_weights', 'FitComparison', 'SlabBuilder', 'mirror_qmap_positive_qxy_only', 'evaluate_fit',
           'make_linecut', 'plot_fit', 'plot_linecuts', 'load_exp_image', 'mirror_qmap_fix_detgap',
           'mirror_qmap_positive_qxy_only_old', 'find_nearest', 'match_qmap_index', 'normalize_qmap',
           'normalize_qmap_position', 'mask_forbidden_pixels', 'create_wedge_mask', 'rebin_and_combine_qmaps',
           'scale_offset_mask_qmap', 'add_f0_q_dependence']


def _load_cell(input_filepath):
    if input_filepath.lower().endswith('.xyz'):
        return load_xyz(input_filepath)
    if input_filepath.lower().endswith('.pdb'):
        return load_pdb(input_filepath)
    raise Exception('Files must be a .pdb or .xyz file')


class SlabBuilder:
    """slabmaker_fitting (tools/comparison.py:595-671) for one unit-cell file and cell,
    on the device: the cell is loaded and uploaded once, ``build(x, y, z)`` replicates
    it (gx_slab_tile: the reference's concatenation order and addition order, so the
    coordinates are bit-identical), takes the bounding box (gx_slab_bounds, exact),
    keeps the centred box in the original order (gx_slab_crop) and shifts to the
    origin (gx_shift_min).  The host only sees the six extremes (for the reference's
    asserts and buffers) and the kept count."""

    def __init__(self, input_filepath, a, b, c, alpha, beta, gamma, cell=None):
        """``cell=(coords, elements)`` gives the unit cell directly (``input_filepath``
        is then not read)."""
        import torch
        coords, elements = _load_cell(input_filepath) if cell is None else cell
        self.uniq, inv = np.unique(np.asarray(elements), return_inverse=True)
        self.a_vect, self.b_vect, self.c_vect = calc_real_space_abc(a, b, c, alpha, beta, gamma)
        dev = engine.device()
        self.n_cell = len(coords)
        self.cell = torch.as_tensor(np.ascontiguousarray(coords, dtype=np.float64)).to(dev)
        self.cell_elem = torch.as_tensor(inv.astype(np.int32)).to(dev)
        self.vec = (ctypes.c_double * 6)(self.a_vect[0], self.b_vect[0], self.b_vect[1], self.c_vect[0],
                                         self.c_vect[1], self.c_vect[2])

    def _bounds(self, x, y, z, el):
        import torch
        out = torch.empty(6 + len(self.uniq), dtype=torch.float64, device=x.device)
        d = GxSlab(n_atoms=int(x.numel()), n_elem=len(self.uniq), x=engine.ptr(x).value, y=engine.ptr(y).value,
                   z=engine.ptr(z).value, elem=engine.ptr(el).value)
        call("gx_slab_bounds", ctypes.byref(d), engine.ptr(out), engine.ptr(out[6:].view(torch.int64)),
             engine.stream_handle())
        return out   # xmin, xmax, ymin, ymax, zmin, zmax | per-element counts (int64 words)

    def build(self, x_size, y_size, z_size, shift_again=True):
        """Device (x, y, z, elem_id) columns of the slab; elem ids index ``self.uniq``.
        ``shift_again=False`` skips the final shift of the kept atoms to the origin
        (the standalone slabmaker.py:83-116 writes them as cut)."""
        import torch
        num_x = int(np.ceil(2 * x_size / self.a_vect[0]))
        num_y = int(np.ceil(2 * y_size / self.b_vect[1]))
        num_z = int(np.ceil(2 * z_size / self.c_vect[2]))
        n = self.n_cell * (num_x + 1) * (num_y + 1) * (num_z + 1)
        dev = self.cell.device
        x, y, z = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3))
        el = torch.empty(n, dtype=torch.int32, device=dev)
        call("gx_slab_tile", engine.ptr(self.cell), engine.ptr(self.cell_elem), self.n_cell, num_x, num_y, num_z,
             self.vec, engine.ptr(x), engine.ptr(y), engine.ptr(z), engine.ptr(el), engine.stream_handle())
        mm = self._bounds(x, y, z, el)[:6].cpu().numpy()
        x_max, y_max, z_max = mm[1] - mm[0], mm[3] - mm[2], mm[5] - mm[4]
        assert x_max > x_size, "x_max must be greater than x_size"
        assert y_max > y_size, "y_max must be greater than y_size"
        assert z_max > z_size, "z_max must be greater than z_size"
        x_buffer = (x_max - x_size) / 2
        y_buffer = (y_max - y_size) / 2
        z_buffer = (z_max - z_size) / 2
        lim = (ctypes.c_double * 9)(mm[0], mm[2], mm[4], x_buffer, x_max - x_buffer, y_buffer, y_max - y_buffer,
                                    z_buffer, z_max - z_buffer)
        ox, oy, oz = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3))
        oel = torch.empty(n, dtype=torch.int32, device=dev)
        work = torch.empty((n + 2047) // 2048 + 1, dtype=torch.int64, device=dev)
        kept = ctypes.c_int64(0)
        call("gx_slab_crop", engine.ptr(x), engine.ptr(y), engine.ptr(z), engine.ptr(el), n, lim, engine.ptr(ox),
             engine.ptr(oy), engine.ptr(oz), engine.ptr(oel), engine.ptr(work), ctypes.byref(kept),
             engine.stream_handle())
        m = int(kept.value)
        del x, y, z, el
        if m == 0 and shift_again:   # the reference's np.min of the empty selection raises
            raise ValueError("zero-size array to reduction operation minimum which has no identity")
        ox, oy, oz, oel = ox[:m], oy[:m], oz[:m], oel[:m]
        if shift_again:
            mm2 = self._bounds(ox, oy, oz, oel)
            call("gx_shift_min", engine.ptr(ox), engine.ptr(oy), engine.ptr(oz), m, engine.ptr(mm2),
                 engine.stream_handle())
        return ox, oy, oz, oel


def slabmaker_fitting(input_filepath, x_size, y_size, z_size, a, b, c, alpha, beta, gamma):
    """Tile the unit cell and crop a centred slab (tools/comparison.py:595-671): the
    replication, bounding boxes, crop and shifts on the GPU (``SlabBuilder``), the
    coordinates and the order of the atoms identical to the reference's."""
    import torch
    sb = SlabBuilder(input_filepath, a, b, c, alpha, beta, gamma)
    x, y, z, el = sb.build(x_size, y_size, z_size)
    coords = torch.stack((x, y, z), dim=1).cpu().numpy()
    return coords, sb.uniq[el.cpu().numpy()]


def voxelgridmaker_fitting(coords, elements, r_voxel_size, q_voxel_size, max_q, energy,
                           num_cpus=None, fill_bkg=False, smooth=0):
    """3D I(q) voxel grid by the phi-frame method (tools/comparison.py:673-788).

    Returns (iq [qy, qx, qz] FP64, qx, qy, qz) of the cropped grid
    (|q| < max_q + dq per axis), f0_C(q)^2 applied.  ``num_cpus`` is accepted
    for signature compatibility; all frames run on the GPU.
    """
    res = engine.run_voxelgrid(coords, elements, r_voxel_size, q_voxel_size, max_q, energy,
                               fill_bkg=fill_bkg, smooth=int(smooth), want64=True)
    return res['iq64'].cpu().numpy(), res['qx'], res['qy'], res['qz']


def _weights(path, angles):
    return np.load(path) if path else np.ones_like(angles) / len(angles)


def detectormaker_fitting(iq, qx, qy, qz, num_pixels, max_q, angle_init_vals, angle_init_axs, psis,
                          psi_weights_path, phis, phi_weights_path, thetas, theta_weights_path,
                          mirror=True):
    """Orientation-averaged detector image (tools/comparison.py:790-870).

    Orientations run in (psi outer, phi, theta inner) order and are summed in
    that order in FP64 on the GPU (deterministic; the reference's threaded
    ``+=`` is racy).  Returns (det_sum [v, h] FP64, det_h, det_v)."""
    out, h, v = detector_image(iq, qx, qy, qz, num_pixels, max_q, angle_init_vals, angle_init_axs, psis,
                               psi_weights_path, phis, phi_weights_path, thetas, theta_weights_path, mirror)
    return out.cpu().numpy(), h, v


def detector_image(iq, qx, qy, qz, num_pixels, max_q, angle_init_vals, angle_init_axs, psis, psi_weights_path,
                   phis, phi_weights_path, thetas, theta_weights_path, mirror, floor_scale=True):
    """The device image of detectormaker_fitting (``floor_scale``: NaN / <= 0 -> 1e-6,
    then * 1e-6) or, without ``floor_scale``, the mirrored sum the legacy
    detectormaker floors on its own.  ``iq`` is a host array (uploaded on every call) or
    a device tensor (used as it is).  Returns (device image [v, h], det_h, det_v)."""
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
    # a host iq is uploaded on every call: the caller may have edited it or derived it
    # from an earlier result, so no device copy of an earlier array is reused
    dev_iq = engine.to_dev(iq if torch.is_tensor(iq) else np.asarray(iq, dtype=np.float64), torch.float64)
    out, _ = engine.run_detector(dev_iq, qx, qy, qz, plan, mirror, floor_scale=floor_scale)
    return out, plan.h, plan.v


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


# --------------------------------------------------------------------------------------
# experimental comparison of a fit evaluation (tools/comparison.py:161-191, 406-445,
# 469-592).  Sample positions, pad column maps and rebin weights depend on geometry
# only: they are planned on the host with the reference's own formulas and cached;
# the device applies them to every new image (gx_map_linear, gx_polar_pad,
# gx_csr_apply, gx_fit_residual).
# --------------------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _is_dev(a):
    torch = _torch()
    return torch.is_tensor(a)


def _as_dev(a):
    torch = _torch()
    if torch.is_tensor(a):
        return a.to(device=engine.device(), dtype=torch.float64).contiguous()
    return engine.to_dev(np.asarray(a, dtype=np.float64), torch.float64)


def _map_linear_dev(img_dev, ys_dev, xs_dev, cval=0.0):
    """scipy.ndimage.map_coordinates(img, (ys, xs), order=1, mode='constant', cval)."""
    torch = _torch()
    h, w = img_dev.shape
    out = torch.empty(ys_dev.shape, dtype=torch.float64, device=img_dev.device)
    call("gx_map_linear", engine.ptr(img_dev), h, w, engine.ptr(ys_dev), engine.ptr(xs_dev),
         int(ys_dev.numel()), float(cval), engine.ptr(out), engine.stream_handle())
    return out


def _check_order(order):
    if order != 1:
        raise NotImplementedError("only order=1 (linear) interpolation, the reference's setting")


def _linear_polar_coords(shape, o, r, out_shape):
    """Sample positions of linear_polar (tools/comparison.py:475-494)."""
    if o is None:
        o = np.array(shape[:2]) / 2 - 0.5
    else:
        o = np.array(o)
    if r is None:
        r = np.sqrt((np.array(shape[:2]) ** 2).sum()) / 2
    if out_shape is None:
        out_shape = (int(round(r * 2 * np.pi)), int(round(r)))
    out_h, out_w = out_shape
    rs = np.linspace(0, r, out_w)
    ts = np.linspace(0, 2 * np.pi, out_h)
    xs = rs * np.cos(ts[:, None]) + o[1]
    ys = rs * np.sin(ts[:, None]) + o[0]
    return ys, xs


def _polar_linear_coords(polar_shape, o, r, out_shape):
    """Sample positions (theta row, radius column) of polar_linear
    (tools/comparison.py:505-530)."""
    if r is None:
        r = polar_shape[1]
    if out_shape is None:
        out_shape = (int(r * 2), int(r * 2))
    if o is None:
        o = np.array(out_shape) / 2 - 0.5
    else:
        o = np.array(o)
    out_h, out_w = out_shape
    ys, xs = np.mgrid[:out_h, :out_w] - o[:, None, None]
    rs = np.sqrt(ys ** 2 + xs ** 2)
    ts = np.arctan2(ys, xs)
    ts[ts < 0] += 2 * np.pi
    rs = (rs / r) * (polar_shape[1] - 1)
    ts = (ts / (2 * np.pi)) * (polar_shape[0] - 1)
    return ts, rs


def _out_shape(output):
    if output is None:
        return None
    return tuple(output) if isinstance(output, tuple) else tuple(np.shape(output))


def _deliver(out_dev, like, output):
    """Device result -> the caller's kind: a device tensor for device input, else
    numpy (written into ``output`` when that is an array, as map_coordinates does)."""
    if _is_dev(like):
        return out_dev
    res = out_dev.cpu().numpy()
    if isinstance(output, np.ndarray):
        output[...] = res
        return output
    return res


def linear_polar(img, o=None, r=None, output=None, order=1, cont=0):
    """Cartesian image -> (theta, r) polar image (tools/comparison.py:469-497),
    bilinear samples on the GPU."""
    _check_order(order)
    ys, xs = _linear_polar_coords(np.shape(img), o, r, _out_shape(output))
    out = _map_linear_dev(_as_dev(img), _as_dev(ys), _as_dev(xs), cont)
    return _deliver(out, img, output)


def polar_linear(img, o=None, r=None, output=None, order=1, cont=0):
    """(theta, r) polar image -> Cartesian image (tools/comparison.py:499-532),
    bilinear samples on the GPU."""
    _check_order(order)
    ts, rs = _polar_linear_coords(np.shape(img), o, r, _out_shape(output))
    out = _map_linear_dev(_as_dev(img), _as_dev(ts), _as_dev(rs), cont)
    return _deliver(out, img, output)


def pad_columns(width, r_axis, pad_width, pad_range):
    """Column map of add_pad (tools/comparison.py:534-557): padded[:, c] =
    polar[:, src[c]], or None when nothing is padded.  Same arithmetic and the
    same errors as the reference's repeated np.insert."""
    pad_range_min, pad_range_max = pad_range
    widths = np.diff(r_axis)
    widths_trim = widths[widths > 0]
    pixel_width = np.mean(widths_trim)
    pad_width_pixels = int(np.round(pad_width / pixel_width))
    if pad_width_pixels == 0:
        return None
    pad_start_idx = int(np.argmin(np.abs(r_axis - pad_range_min)))
    pad_end_idx = int(np.argmin(np.abs(r_axis - pad_range_max)))
    pad_range_pixels = pad_end_idx - pad_start_idx
    spacing = int((pad_range_pixels - 1) / pad_width_pixels - 1)
    assert spacing >= 0, 'pad_range is too small for desired pad_width'
    cols = list(range(int(width)))
    for i in range(pad_width_pixels):
        pad_idx = pad_start_idx + 2 * i + spacing * i
        if pad_idx >= len(cols):
            raise IndexError(f"index {pad_idx} is out of bounds for axis 1 with size {len(cols)}")
        cols.insert(pad_idx + 1, cols[pad_idx])
    return np.asarray(cols[:len(cols) - pad_width_pixels], dtype=np.int32)


def _polar_pad_dev(polar_dev, src, mask_zero):
    torch = _torch()
    h, w = polar_dev.shape
    if src is None:
        src = np.arange(w, dtype=np.int32)
    src_dev = torch.as_tensor(src, dtype=torch.int32).to(polar_dev.device)
    out = torch.empty_like(polar_dev)
    call("gx_polar_pad", engine.ptr(polar_dev), h, w, engine.ptr(src_dev), int(bool(mask_zero)),
         engine.ptr(out), engine.stream_handle())
    return out


def add_pad(polar_image, r_axis, pad_width, pad_range):
    """Duplicate columns of a polar image to stretch a radial band
    (tools/comparison.py:534-557); the column copy runs on the GPU."""
    src = pad_columns(np.shape(polar_image)[1], np.asarray(r_axis), pad_width, pad_range)
    if src is None:
        return polar_image
    return _deliver(_polar_pad_dev(_as_dev(polar_image), src, False), polar_image, None)


class ShiftPeakPlan:
    """shift_peak (tools/comparison.py:559-592) for one trimmed detector geometry and
    pad setting: the polar and inverse sample positions and the pad column map,
    planned once; ``__call__`` applies them to a device image [v, h]."""

    def __init__(self, shape, det_h, det_v, pad_width_qspace, pad_range_qspace):
        det_h = np.asarray(det_h, dtype=np.float64)
        det_v = np.asarray(det_v, dtype=np.float64)
        row, col = shape
        self.shape = (int(row), int(col))
        radius = float(np.sqrt(row ** 2 + col ** 2))
        o = (int(np.argmin(np.abs(det_v))), int(np.argmin(np.abs(det_h))))
        ys, xs = _linear_polar_coords(self.shape, o, radius, None)
        self.polar_shape = ys.shape
        self.ys, self.xs = _as_dev(ys), _as_dev(xs)
        # r_axis = row 0 of linear_polar(hypot(det_h, det_v)), sampled like the image
        xx, yy = np.meshgrid(det_h, det_v)
        r_dev = _as_dev(np.hypot(xx, yy))
        r_axis = _map_linear_dev(r_dev, self.ys[0].contiguous(), self.xs[0].contiguous(), 0.0).cpu().numpy()
        self.src = pad_columns(self.polar_shape[1], r_axis, pad_width_qspace, pad_range_qspace)
        ts, rs = _polar_linear_coords(self.polar_shape, o, None, self.shape)
        self.ts, self.rs = _as_dev(ts), _as_dev(rs)

    def __call__(self, image_dev):
        polar = _map_linear_dev(image_dev, self.ys, self.xs, 0.0)
        padded = _polar_pad_dev(polar, self.src, True)
        return _map_linear_dev(padded, self.ts, self.rs, 0.0)


_SHIFT_PLANS = {}


def _shift_plan(shape, det_h, det_v, pad_width, pad_range):
    torch = _torch()
    key = (tuple(shape), np.asarray(det_h, np.float64).tobytes(), np.asarray(det_v, np.float64).tobytes(),
           float(pad_width), tuple(float(p) for p in pad_range), torch.cuda.current_device())
    plan = _SHIFT_PLANS.get(key)
    if plan is None:
        if len(_SHIFT_PLANS) >= 8:
            _SHIFT_PLANS.pop(next(iter(_SHIFT_PLANS)))
        plan = _SHIFT_PLANS[key] = ShiftPeakPlan(shape, det_h, det_v, pad_width, pad_range)
    return plan


def shift_peak(image, det_h, det_v, pad_width_qspace, pad_range_qspace):
    """Stretch the pi-pi band of a detector image radially by pad_width_qspace inside
    pad_range_qspace (tools/comparison.py:559-592), on the GPU."""
    plan = _shift_plan(np.shape(image), det_h, det_v, pad_width_qspace, pad_range_qspace)
    return _deliver(plan(_as_dev(image)), image, None)


def trim_sim_data(sim_det_ints, det_h, det_v, exp_qxy, exp_qz):
    """Trim a detector image and its axes to the experimental q range
    (tools/comparison.py:161-191).  A device image stays on the device."""
    qxy_min, qxy_max = np.min(exp_qxy), np.max(exp_qxy)
    qz_min, qz_max = np.min(exp_qz), np.max(exp_qz)
    det_h = np.asarray(det_h)
    det_v = np.asarray(det_v)
    h_mask = (det_h >= qxy_min) & (det_h <= qxy_max)
    v_mask = (det_v >= qz_min) & (det_v <= qz_max)
    if _is_dev(sim_det_ints):
        torch = _torch()
        iv = torch.as_tensor(np.nonzero(v_mask)[0]).to(sim_det_ints.device)
        ih = torch.as_tensor(np.nonzero(h_mask)[0]).to(sim_det_ints.device)
        trimmed = sim_det_ints.index_select(0, iv).index_select(1, ih).contiguous()
    else:
        trimmed = sim_det_ints[v_mask, :][:, h_mask]
    return trimmed, det_h[h_mask], det_v[v_mask]


def rebin_weights(qxy1, qz1, qxy2, qz2):
    """griddata(points2, values, (grid1_x, grid1_z), method='linear') as sparse
    weights (host): (row_ptr int64 [n+1], col int32, val FP64, target shape).  The
    triangle of each target point comes from the Delaunay triangulation of the source
    grid and its point location (scipy.spatial.Delaunay / find_simplex, which griddata
    uses), its barycentric coordinates in griddata's arithmetic:
    c_i = T_i0 (x0 - r0) + T_i1 (x1 - r1), c_2 = (1 - c_0) - c_1, so that applying
    the weights in the stored order (0 + c_0 v_0 + c_1 v_1 + c_2 v_2) reproduces
    griddata bit for bit; rows outside the hull are empty (NaN there, griddata's fill)."""
    from scipy.spatial import Delaunay
    grid1_x, grid1_z = np.meshgrid(qxy1, qz1)
    grid2_x, grid2_z = np.meshgrid(qxy2, qz2)
    points2 = np.vstack((grid2_x.ravel(), grid2_z.ravel())).T
    xi = np.vstack((grid1_x.ravel(), grid1_z.ravel())).T
    tri = Delaunay(points2)
    simplex = tri.find_simplex(xi)
    inside = simplex >= 0
    T = tri.transform[simplex[inside]]                      # [m, 3, 2]
    d0 = xi[inside, 0] - T[:, 2, 0]
    d1 = xi[inside, 1] - T[:, 2, 1]
    c0 = T[:, 0, 0] * d0 + T[:, 0, 1] * d1
    c1 = T[:, 1, 0] * d0 + T[:, 1, 1] * d1
    bary = np.stack([c0, c1, (1.0 - c0) - c1], axis=1)
    verts = tri.simplices[simplex[inside]]
    counts = np.where(inside, 3, 0)
    row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return row_ptr, verts.astype(np.int32).ravel(), bary.ravel(), grid1_x.shape


class RebinWeights:
    """``rebin_weights`` resident on the device, applied by gx_csr_apply."""

    def __init__(self, qxy1, qz1, qxy2, qz2):
        torch = _torch()
        row_ptr, col, val, self.shape = rebin_weights(qxy1, qz1, qxy2, qz2)
        self.n_rows = len(row_ptr) - 1
        dev = engine.device()
        self.row_ptr = torch.as_tensor(row_ptr).to(dev)
        self.col = torch.as_tensor(col).to(dev)
        self.val = torch.as_tensor(val).to(dev)

    def apply(self, values_dev, empty=np.nan):
        torch = _torch()
        y = torch.empty(self.n_rows, dtype=torch.float64, device=values_dev.device)
        call("gx_csr_apply", engine.ptr(self.row_ptr), engine.ptr(self.col), engine.ptr(self.val), self.n_rows,
             engine.ptr(values_dev), float(empty), engine.ptr(y), engine.stream_handle())
        return y.view(self.shape)


def rebin_qmap_and_mask(qxy1, qz1, qmap2, qxy2, qz2, qmap2_mask):
    """Rebin an experimental map and its mask onto the simulated (qxy1, qz1) grid
    (tools/comparison.py:406-445): griddata 'linear' as precomputed barycentric
    weights applied on the GPU; the mask is 1 wherever its interpolant is > 0."""
    W = RebinWeights(qxy1, qz1, qxy2, qz2)
    rebinned = W.apply(_as_dev(np.asarray(qmap2, dtype=np.float64).ravel()))
    mask_r = W.apply(_as_dev(np.asarray(qmap2_mask, dtype=np.float64).ravel()))
    mask = (mask_r > 0).to(dtype=_torch().int64)
    return rebinned.cpu().numpy(), mask.cpu().numpy()


def _scale_offset_dev(sim_dev, ref_dev, mask_u8_dev):
    """optimize_scale_offset (tools/comparison.py:873-882) on device arrays."""
    torch = _torch()
    work = torch.empty(GX_MOMENTS_WORK, dtype=torch.float64, device=sim_dev.device)
    res = (ctypes.c_double * 6)()
    call("gx_scale_offset_moments", engine.ptr(sim_dev), engine.ptr(ref_dev), engine.ptr(mask_u8_dev),
         int(sim_dev.numel()), engine.ptr(work), res, engine.stream_handle())
    n, ms, my, sss, ssy = res[0], res[1], res[2], res[3], res[4]
    if n == 0:
        raise np.linalg.LinAlgError("no unmasked pixels")
    scale = ssy / sss if sss != 0 else 0.0
    return scale, my - scale * ms


class FitComparison:
    """The experimental side of fit_slabsize.py's fullfit_model (19-47) for a fixed
    experimental map: trim the simulated detector image to the experimental range,
    shift_peak, least-squares scale / offset over unmasked pixels, masked residual
    -- all on the device, from the device detector image."""

    def __init__(self, det_h, det_v, rebin_map, rebin_mask, exp_qxy, exp_qz, pad_width, pad_range):
        torch = _torch()
        qxy_min, qxy_max = np.min(exp_qxy), np.max(exp_qxy)
        qz_min, qz_max = np.min(exp_qz), np.max(exp_qz)
        det_h, det_v = np.asarray(det_h), np.asarray(det_v)
        h_mask = (det_h >= qxy_min) & (det_h <= qxy_max)
        v_mask = (det_v >= qz_min) & (det_v <= qz_max)
        dev = engine.device()
        self.iv = torch.as_tensor(np.nonzero(v_mask)[0]).to(dev)
        self.ih = torch.as_tensor(np.nonzero(h_mask)[0]).to(dev)
        self.h_trim, self.v_trim = det_h[h_mask], det_v[v_mask]
        shape = (len(self.v_trim), len(self.h_trim))
        self.shift = _shift_plan(shape, self.h_trim, self.v_trim, pad_width, pad_range)
        self.ref = _as_dev(np.asarray(rebin_map, dtype=np.float64))
        self.mask = torch.as_tensor((np.asarray(rebin_mask) != 0).astype(np.uint8)).to(dev)
        assert tuple(self.ref.shape) == shape, "rebin_map must match the trimmed detector grid"
        self.n = int(np.prod(shape))

    def padded(self, det_dev):
        """Trimmed, shift_peak-ed device image [v_trim, h_trim] of a device image [v, h]."""
        sim = det_dev.index_select(0, self.iv).index_select(1, self.ih).contiguous()
        return self.shift(sim)

    def residual(self, det_dev):
        """(residual device tensor [n], scale, offset) for a device image [v, h]."""
        torch = _torch()
        sim = self.padded(det_dev)
        scale, offset = _scale_offset_dev(sim, self.ref, self.mask)
        out = torch.empty(self.n, dtype=torch.float64, device=sim.device)
        call("gx_fit_residual", engine.ptr(sim), engine.ptr(self.ref), engine.ptr(self.mask), self.n,
             float(scale), float(offset), engine.ptr(out), engine.stream_handle())
        return out, scale, offset


# --------------------------------------------------------------------------------------
# experimental inputs and result figures of the comparison / fit scripts (host side)
# --------------------------------------------------------------------------------------
def load_exp_image(img_path):
    """The experimental q-map image.  ``.npy`` is read with numpy (no pickles); other
    formats go through fabio, as the reference's ``fabio.open(img_path).data``
    (simulate_GIWAXS.py:124, fit_slabsize.py:139)."""
    if str(img_path).lower().endswith('.npy'):
        return np.load(img_path, allow_pickle=False)
    try:
        import fabio
    except ImportError as e:
        raise ImportError(f"reading {img_path} needs fabio (not installed); save the image as .npy") from e
    return fabio.open(img_path).data


def _positive_side_filled(qmap, qxy):
    """(qxy > 0 columns of qmap, their qxy) with each pixel replaced by its mirrored
    qxy < 0 partner where that one is brighter than 1.1x (the mirrored side
    zero-padded to the positive side's width): the detector-gap fill shared by
    tools/comparison.py:268-319 and 322-378."""
    keep = qxy > 0
    q_pos = qxy[keep]
    m_pos = qmap[:, keep]
    m_neg = qmap[:, qxy < 0][:, ::-1]
    if m_neg.shape[1] < m_pos.shape[1]:
        m_neg = np.pad(m_neg, ((0, 0), (0, m_pos.shape[1] - m_neg.shape[1])), mode='constant', constant_values=0)
    brighter = m_pos * 1.1 < m_neg
    m_pos[brighter] = m_neg[brighter]
    return m_pos, q_pos


def mirror_qmap_positive_qxy_only(qmap, qxy, qz):
    """tools/comparison.py:268-319: keep the qxy > 0 columns; where the mirrored
    qxy < 0 column (zero-padded to the positive side's width) is brighter than 1.1x
    the positive pixel, take it (detector gaps); put the qxy == 0 column first."""
    qmap = np.asarray(qmap)
    qxy = np.asarray(qxy)
    m_pos, q_pos = _positive_side_filled(qmap, qxy)
    if np.any(qxy == 0):
        zero_col = qmap[:, np.flatnonzero(qxy == 0)].reshape(-1, 1)
        m_pos = np.concatenate((zero_col, m_pos), axis=1)
        q_pos = np.concatenate(([0], q_pos))
    return m_pos, q_pos, qz


def mirror_qmap_fix_detgap(qmap, qxy, qz):
    """tools/comparison.py:322-378: the gap-filled qxy > 0 side (as
    mirror_qmap_positive_qxy_only) reflected across qxy = 0, [-q reversed, q], with the
    qxy == 0 column, when there is one, appended at the END (qxy 0 last)."""
    qmap = np.asarray(qmap)
    qxy = np.asarray(qxy)
    m_pos, q_pos = _positive_side_filled(qmap, qxy)
    q_out = np.concatenate((-q_pos[::-1], q_pos))
    m_out = np.concatenate((m_pos[:, ::-1], m_pos), axis=1)
    if np.any(qxy == 0):
        m_out = np.concatenate((m_out, qmap[:, np.flatnonzero(qxy == 0)].reshape(-1, 1)), axis=1)
        q_out = np.append(q_out, 0)
    return m_out, q_out, qz


def mirror_qmap_positive_qxy_only_old(qmap, qxy, qz):
    """tools/comparison.py:239-266: the qxy >= 0 columns reflected across qxy = 0
    ([-q reversed, q]; a qxy == 0 column appears twice), no gap fill."""
    qmap = np.asarray(qmap)
    qxy = np.asarray(qxy)
    keep = qxy >= 0
    q_pos, m_pos = qxy[keep], qmap[:, keep]
    return np.concatenate((m_pos[:, ::-1], m_pos), axis=1), np.concatenate((-q_pos[::-1], q_pos)), qz


def find_nearest(array, value):
    """Index of the element of ``array`` closest to ``value`` (first on ties),
    tools/comparison.py:18-21."""
    return np.abs(np.asarray(array) - value).argmin()


def match_qmap_index(qxy_val, qz_val, qxy, qz):
    """(qz index, qxy index) of the map pixel closest to (qxy_val, qz_val),
    tools/comparison.py:23-45."""
    return np.abs(qz - qz_val).argmin(), np.abs(qxy - qxy_val).argmin()


def normalize_qmap(qmap1, qmap2):
    """qmap1 scaled so its NaN-ignoring sum equals qmap2's, tools/comparison.py:47-55."""
    return qmap1 * (np.nansum(qmap2) / np.nansum(qmap1))


def normalize_qmap_position(qmap1, qmap2, qxy_axis, qz_axis, qxy_pos, qz_pos):
    """qmap1 scaled so its pixel nearest (qxy_pos, qz_pos) equals qmap2's there (both
    maps on the same axes), tools/comparison.py:57-71; prints the pixel and values."""
    ix = np.argmin(np.abs(qxy_axis - qxy_pos))
    iz = np.argmin(np.abs(qz_axis - qz_pos))
    v1, v2 = qmap1[iz, ix], qmap2[iz, ix]
    print(f'qxy_idx={ix}, qz_idx={iz}, qmap1_int={v1}, qmap2_int={v2}')
    return qmap1 * (v2 / v1)


def _wedge(qxy, qz, alpha_i_deg, energy):
    """Boolean [qz, qxy] map of the missing wedge: |qxy| below the smallest qxy
    reachable at each qz for incidence angle alpha_i and wavelength 12400 / energy A
    (exit angle theta = asin(qz wl / 2 pi - sin alpha_i) + alpha_i, qxy_min =
    |2 pi / wl (cos(theta - alpha_i) - cos alpha_i)|), tools/comparison.py:106-114.
    A qz the geometry cannot reach (NaN theta) masks nothing."""
    wl = 12400 / energy
    a = np.deg2rad(alpha_i_deg)
    qz = np.asarray(qz, dtype=np.float64)
    with np.errstate(invalid='ignore'):
        theta = np.arcsin(qz * wl / (2 * np.pi) - np.sin(a)) + a
        qxy_min = np.abs((2 * np.pi / wl) * (np.cos(theta - a) - np.cos(a)))
        return np.abs(np.asarray(qxy))[None, :] < qxy_min[:, None]


def mask_forbidden_pixels(image, qxy, qz, alpha_i_deg, energy):
    """Copy of ``image`` [qz, qxy] with the missing-wedge pixels set to NaN,
    tools/comparison.py:73-116."""
    out = np.array(image, copy=True)
    out[_wedge(qxy, qz, alpha_i_deg, energy)] = np.nan
    return out


def create_wedge_mask(image, qxy, qz, alpha_i_deg, energy):
    """Integer [qz, qxy] mask, 1 on the missing-wedge pixels and 0 elsewhere,
    tools/comparison.py:118-159 (``image`` is not used)."""
    return _wedge(qxy, qz, alpha_i_deg, energy).astype(np.int64)


def rebin_and_combine_qmaps(qmap1, qxy1, qz1, qmap2, qxy2, qz2, pos=0):
    """qmap1 with its columns from qxy = 0 on replaced by qmap2 linearly interpolated
    onto qmap1's grid (NaN outside qmap2's hull) and scaled: to qmap1 at pos = (qxy, qz)
    when given, else by qmap1's 15th over the interpolated map's 75th NaN-percentile,
    tools/comparison.py:193-237.  The interpolation weights are rebin_weights',
    bit-identical to scipy's griddata 'linear'."""
    qmap2 = np.asarray(qmap2, dtype=np.float64)
    row_ptr, col, val, shape = rebin_weights(np.asarray(qxy1), np.asarray(qz1), np.asarray(qxy2), np.asarray(qz2))
    rebinned = _csr_apply_host(row_ptr, col, val, qmap2.ravel()).reshape(shape)
    if pos:
        rebinned = normalize_qmap_position(rebinned, qmap1, qxy1, qz1, pos[0], pos[1])
    else:
        rebinned = rebinned * (np.nanpercentile(qmap1, 15) / np.nanpercentile(rebinned, 75))
    zero = np.abs(qxy1).argmin()
    out = np.array(qmap1, copy=True)
    out[:, zero:] = rebinned[:, zero:]
    return out


def _csr_apply_host(row_ptr, col, val, x):
    """out[r] = 0 + sum over the row's entries in stored order of val * x[col]; NaN for
    rows without entries (gx_csr_apply's arithmetic on the host)."""
    n = len(row_ptr) - 1
    counts = np.diff(row_ptr)
    out = np.full(n, np.nan)
    rows = np.flatnonzero(counts)
    if len(rows) == 0:
        return out
    width = int(counts.max())
    acc = np.zeros(len(rows))
    for k in range(width):
        has = counts[rows] > k
        idx = row_ptr[rows[has]] + k
        acc[has] = acc[has] + val[idx] * x[col[idx]]
    out[rows] = acc
    return out


def scale_offset_mask_qmap(qmap, qmap_mask, scale=1, offset=0):
    """qmap * scale + offset with the pixels where qmap_mask == 1 set to 0,
    tools/comparison.py:447-467."""
    return np.where(qmap_mask == 1, 0, qmap * scale + offset)


def add_f0_q_dependence(det_img, det_h, det_v, element):
    """det_img [v, h] multiplied IN PLACE by f0(|q|) / Z of ``element`` (Cromer-Mann
    a0 e^(-a1 q^2 / 16 pi^2) + ... + a6 e^(-a7 q^2 / 16 pi^2) + a8, terms added in that
    order, q^2 = qxy^2 + qz^2 at each pixel) and returned, tools/comparison.py:380-404."""
    a = aff_dict[element]
    qxy, qz = np.meshgrid(np.asarray(det_h, dtype=np.float64), np.asarray(det_v, dtype=np.float64))
    q2 = np.sqrt(qxy ** 2 + qz ** 2) ** 2
    c = 16 * np.pi ** 2
    fq = a[0] * np.exp(-a[1] * q2 / c)
    fq = fq + a[2] * np.exp(-a[3] * q2 / c)
    fq = fq + a[4] * np.exp(-a[5] * q2 / c)
    fq = fq + a[6] * np.exp(-a[7] * q2 / c)
    fq = fq + a[8]
    det_img *= fq / ptable[element]
    return det_img


def evaluate_fit(best_params, fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params,
                 fixed_exp_params):
    """(rebin_map, sim_comp_map, diff_map) of the slab of size ``best_params``
    (tools/comparison.py:884-912): the fit model's image, trimmed and shift_peak-ed on
    the device, scaled by the least-squares scale / offset, masked pixels zeroed, and
    the difference where the experimental map exceeds 1e-10."""
    from ..fit_slabsize import model_for
    model = model_for(fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params, fixed_exp_params)
    rebin_map, rebin_mask = fixed_exp_params[0], fixed_exp_params[1]
    sim_pad, scale, offset = model.padded(*best_params)
    sim_comp_map = scale * sim_pad.cpu().numpy() + offset
    sim_comp_map[rebin_mask == 1] = 0
    diff_map = np.where(rebin_map > 1e-10, sim_comp_map - rebin_map, 0)
    return rebin_map, sim_comp_map, diff_map


def _pyplot():
    import matplotlib
    matplotlib.use('Agg')
    import matplotlib.pyplot as plt
    return matplotlib, plt


def _q_axes(ax, max_q=None, fontsize=12):
    from matplotlib import ticker
    ax.set_xlabel('$\\mathregular{q_{xy}}$ ($\\AA^{-1}$)', fontsize=fontsize)
    ax.set_ylabel('$\\mathregular{q_z}$ ($\\AA^{-1}$)', fontsize=fontsize)
    ax.yaxis.set_major_locator(ticker.MultipleLocator(0.5))
    ax.xaxis.set_major_locator(ticker.MultipleLocator(0.5))
    if max_q is not None:
        ax.set_xlim(0, max_q)
        ax.set_ylim(0, max_q)


def plot_fit(savepath, rebin_map, sim_comp_map, diff_map, det_h_trim, det_v_trim, max_q, suptitle):
    """Simulated, experimental and difference maps side by side
    (tools/comparison.py:914-975): turbo colour scale on the experimental 1st-99.9th
    percentiles for both maps, symmetric bwr scale for the difference."""
    matplotlib, plt = _pyplot()
    extent = (np.min(det_h_trim), np.max(det_h_trim), np.min(det_v_trim), np.max(det_v_trim))
    norm = matplotlib.colors.Normalize(vmin=np.percentile(rebin_map, 1), vmax=np.percentile(rebin_map, 99.9))
    bound = max(abs(np.percentile(diff_map, 0.1)), abs(np.percentile(diff_map, 99.9)))
    fig, (ax_sim, ax_exp, ax_diff) = plt.subplots(1, 3, figsize=(12, 6))
    panels = ((ax_exp, rebin_map, norm, 'turbo', 'Experimental GIWAXS'),
              (ax_sim, sim_comp_map, norm, 'turbo', 'Simulated GIWAXS'),
              (ax_diff, diff_map, matplotlib.colors.Normalize(vmin=-bound, vmax=bound), 'bwr', 'Sim. - Exp.'))
    for ax, img, nrm, cmap, title in panels:
        im = ax.imshow(img, norm=nrm, extent=extent, cmap=cmap, origin='lower')
        cb = fig.colorbar(im, ax=ax, shrink=0.45, aspect=25, pad=0.05)
        if ax is ax_diff:
            cb.set_label('(Percent)', fontsize=12)
        _q_axes(ax, max_q)
        ax.set_title(title, fontsize=12)
    plt.suptitle(suptitle)
    plt.tight_layout()
    plt.savefig(savepath, dpi=300)
    plt.close(fig)


def make_linecut(polar_image, chi_axis, r_axis, lower_chi, upper_chi):
    """tools/comparison.py:977-989: mean over the polar rows with lower < chi < upper
    (pixels below 0.1 ignored), dropping q = 0 and empty columns."""
    cut = polar_image[np.logical_and(chi_axis > lower_chi, chi_axis < upper_chi), :]
    cut[cut < 0.1] = np.nan
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter('ignore', RuntimeWarning)
        ints = np.nanmean(cut, axis=0)
    qs = r_axis.copy()
    assert len(ints) == len(qs), f'intensity array shape {np.shape(ints)} does not match q array shape {np.shape(qs)}'
    keep = (qs != 0) & (ints == ints)
    return qs[keep], ints[keep]


def plot_linecuts(savepath, rebin_map, sim_comp_map, det_h_trim, det_v_trim, suptitle):
    """In-plane (chi 0-30), out-of-plane (60-90) and full (0-90) radial line cuts of
    the simulated and experimental maps (tools/comparison.py:991-1060), from their
    polar warps (``linear_polar`` on the device) about the q = 0 pixel."""
    from matplotlib import ticker
    _, plt = _pyplot()
    xx, yy = np.meshgrid(det_h_trim, det_v_trim)
    r, chi = np.hypot(xx, yy), np.arctan2(yy, xx)
    rows, cols = sim_comp_map.shape
    radius = float(np.sqrt(rows ** 2 + cols ** 2))
    o = (int(np.argmin(np.abs(det_v_trim))), int(np.argmin(np.abs(det_h_trim))))
    warp = lambda img: linear_polar(np.ascontiguousarray(img, dtype=np.float64), o=o, r=radius)
    polar_sim, polar_exp, polar_r, polar_chi = (warp(m) for m in (sim_comp_map, rebin_map, r, chi))
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter('ignore', RuntimeWarning)
        chi_axis = np.nanmean(np.where(polar_chi == 0, np.nan, polar_chi), axis=1) * (180 / np.pi)
        r_axis = np.nanmean(np.where(polar_r == 0, np.nan, polar_r), axis=0)
    fig, axes = plt.subplots(1, 3, figsize=(12, 4))
    for ax, (lo, hi), title in zip(axes, ((0, 30), (60, 90), (0, 90)),
                                   ('IP comparions', 'OOP comparions', 'Full comparions')):
        ax.plot(*make_linecut(polar_exp, chi_axis, r_axis, lo, hi), label='Experiment')
        ax.plot(*make_linecut(polar_sim, chi_axis, r_axis, lo, hi), label='Simulation')
        ax.set_xlabel('$\\mathregular{q_{r}}$ ($\\AA^{-1}$)', fontsize=12)
        ax.set_ylabel('Intensity (a.u.)', fontsize=12)
        ax.xaxis.set_major_locator(ticker.MultipleLocator(0.5))
        ax.set_title(title, fontsize=12)
        ax.legend(prop={'size': 12})
    plt.suptitle(suptitle)
    plt.tight_layout()
    plt.savefig(savepath, dpi=300)
    plt.close(fig)
