"""Differentiable variant of the phi-frame path, with a hand-written adjoint.

The reference has no gradients; nothing here replaces a reference function.
Its nearest-grid-point deposit (tools/voxelgrids.py:328-341) is piecewise
constant in the atom positions, so this variant deposits every atom with
cloud-in-cell (bilinear) weights on the y-z grid; the FFT, |F|^2, binning with
counts, mean, crop, f0 weighting and the detector gather / mirror / floor are
the forward path's (tools/voxelgrids.py:381-503, tools/comparison.py:765-870).

``DifferentiableSlab`` maps (atom coordinates, per-element complex f) to the
finalized detector image through ``torch.autograd.Function``; its backward is
the HIP adjoint (include/giwaxsim_amd.h, gx_adj_* / gx_bins_adjoint /
gx_detector_adjoint), which recomputes the phi-frames' row transforms instead of
storing them all; the batch the forward left in its frame buffer (all frames, when
they fit in one batch) is reused rather than recomputed.  Gradients: dL/d coords [n, 3], dL/d f_re and dL/d f_im per
element.  A scale / flat background on the image is ordinary torch algebra
(``image_loss``).

fill_bkg and smooth act on the deposited grid as in the forward path: almost
everywhere in the positions the frame grid is M S + avg_voxel_f p with fixed
masks M and a fixed pattern p (the filled region and smoothed edges follow the
integer index bounds, which are piecewise constant), and avg_voxel_f = sum f r^3
/ (x_b y_b z_b) carries gradient to f and to the atoms that fix the bounding box.

Grid edges that are powers of two (256..2048) use the in-wave transforms; other edges
in [16, 2048] (the reference's fit geometry has N = 1048) the forward path's zoomed
chirp-z band transforms and their adjoints (band values -> grid points through the same
block transforms).  Every atom at least one voxel inside the grid edge.
"""

from __future__ import annotations

import ctypes
import hashlib
import math

import numpy as np
import torch

from . import _lib
from ._lib import GxAdjoint, GxDetector, GxFrames, GxSlab, call
from .engine import (_empty, bin_scale_for, chirpz_adjoint_tables, chirpz_band_tables, chirpz_bands,
                     chirpz_tables, collapse_records, detector_plan, device, element_table,
                     fft_precision, frame_plan, gauss_prefix, intensity_bound, is_pow2_grid,
                     pixels_per_bin, ptr, stream_handle, twiddles)
from .tools.ptable_dict import aff_dict, ptable


def _uniform(a):
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    return a, np.ones_like(a) / len(a)



# transpose maps of the detector gather, by geometry (see DifferentiableSlab)
MAP_BYTES_MAX = 32e9
_DET_MAPS: dict = {}

class _DetectorImage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, f_re, f_im, model):
        out = model._forward(coords, f_re, f_im)
        ctx.model = model
        ctx.generation = model._generation
        ctx.coords_dtype = coords.dtype
        return out

    @staticmethod
    def backward(ctx, g):
        model = ctx.model
        if model._generation != ctx.generation:
            raise RuntimeError("DifferentiableSlab ran another forward before this backward; "
                               "keep one forward/backward pair in flight per model")
        gc, gf = model._backward(g)
        return gc.to(ctx.coords_dtype), gf[:, 0].clone(), gf[:, 1].clone(), None


class DifferentiableSlab:
    """Detector image of one slab as a differentiable function of its atoms.

    ``elements``: element symbol per atom (fixed); ``coords0``: (n, 3)
    coordinates that fix the plan (grid edge N = ceil(2 pi / (q_voxel r)), q
    axis, crop and phi frames exactly as voxelgridmaker_fitting,
    tools/comparison.py:705-731); ``detector``: detectormaker_fitting's
    arguments as a dict (num_pixels, max_q, psis, phis, thetas, optional
    *_weights, angle_init_vals, angle_init_axs, mirror).  ``frames=(lo, hi)``
    restricts the phi-frames.  ``fill_bkg`` / ``smooth`` as in
    voxelgridmaker_fitting.  ``precision``: 'auto' (FP32 transforms for filled
    or smoothed grids, FP64 for sparse ones, as the forward path), 'fp32' or
    'fp64'.

    ``model(coords, f_re, f_im)`` -> finalized image [n_pix, n_pix] (FP64
    device tensor).  f_re / f_im (per element, order of ``self.elements``)
    default to the tabulated f = Z + f1 + i f2 at ``energy``.
    """

    def __init__(self, elements, coords0, r_voxel_size, q_voxel_size, max_q, detector,
                 energy=12700.0, frames=None, precision='auto', batch=None, fill_bkg=False, smooth=0):
        dev = device()
        ids, ftab, _, uniq = element_table(elements, energy)
        self.elements = [str(u) for u in uniq]
        self.f_re0 = torch.as_tensor(ftab.real.copy()).to(dev)
        self.f_im0 = torch.as_tensor(ftab.imag.copy()).to(dev)
        self.n_elem = len(ftab)
        c0 = np.asarray(coords0, dtype=np.float64)
        self.n = n = int(c0.shape[0])
        xb, yb, zb = (float(np.ptp(c0[:, i])) for i in range(3))
        self.r = r = float(r_voxel_size)
        plan = frame_plan(r, float(q_voxel_size), float(max_q), xb, yb, zb)
        if not is_pow2_grid(plan.N) and not 16 <= plan.N <= 2048:
            raise NotImplementedError(f"the differentiable path needs a grid edge in [16, 2048], got N={plan.N}")
        self.plan = plan
        N, W = plan.N, plan.W
        F = plan.n_frames
        self.frames = (0, F) if frames is None else (int(frames[0]), int(frames[1]))
        self.fill, self.smooth = bool(fill_bkg), int(smooth)
        self.dense = self.fill or self.smooth > 0
        self.f64 = fft_precision(self.fill, self.smooth, precision)
        self._batch_arg = batch
        self._pixels_per_bin = pixels_per_bin(plan)

        # atoms and the regrouped slab
        self.x, self.y, self.z = (_empty(n, torch.float64) for _ in range(3))
        self.elem = torch.as_tensor(ids).to(dev).to(torch.int32).contiguous()
        self.ftab = _empty(2 * self.n_elem, torch.float64)
        self.sx, self.sy = _empty(n, torch.float64), _empty(n, torch.float64)
        self.sq = _empty(2 * n, torch.int32)
        self.szrow = _empty(n, torch.int32)
        self.row_off = _empty(N + 2, torch.int32)
        self.row_exp = _empty(N, torch.int32)
        self.work = _empty(N * self.n_elem + 2 * (N + 1), torch.int32)
        self.szf = _empty(n, torch.float64)
        self.sperm = _empty(n, torch.int32)
        self.selem = _empty(n, torch.int32)
        self.stable_work = _empty(-(-n // _lib.GX_STABLE_CHUNK) * (N + 1), torch.int32)
        self.minmax = _empty(6, torch.float64)
        self.ecount = _empty(self.n_elem, torch.int64)

        # frames, bins
        g = plan.grid
        self.fparams = torch.as_tensor(plan.fparams).to(dev).contiguous()
        self.tw = torch.as_tensor(twiddles(N, self.f64)).to(dev).contiguous()
        self.tw64 = torch.as_tensor(twiddles(N, True)).to(dev).contiguous()
        # edges that are not powers of two: the forward path's zoomed chirp-z bands and
        # the kernels of their adjoints (band values -> grid points)
        self.bs = None
        if not is_pow2_grid(N):
            t = lambda a: torch.as_tensor(a).to(dev).contiguous()
            _, ch, _, _ = chirpz_tables(N, self.f64)
            (rlo, rk), (clo, ck) = chirpz_bands(plan)
            # the row band sized to the rows the cloud-in-cell deposit reaches (nodes up
            # to max_voxels + 1, one row of rounding margin); the kernels check it per frame
            M, Kr, twr = chirpz_band_tables(N, rlo, rk, self.f64, support=int(plan.max_voxels) + 3)
            # the column band sized to the rows that hold atoms (with the batch buffers' 16
            # rows of slack) unless smoothing without fill spreads the background below N
            zs = min(N, self._rows_for(zb) + 16) if (self.fill or not self.smooth) else None
            Mc, Kc, twc = chirpz_band_tables(N, clo, ck, self.f64, support=zs)
            self.bs = dict(M=M, Mc=Mc, rlo=rlo, rk=rk, clo=clo, ck=ck, chirp=t(ch), K=t(Kr), tw=t(twr),
                           Kc=t(Kc), twc=t(twc), Kadj=t(chirpz_adjoint_tables(N, rlo, rk, M, self.f64)),
                           Kadj_c=t(chirpz_adjoint_tables(N, clo, ck, Mc, self.f64)))
        cdt = torch.float64 if self.f64 else torch.float32
        # per-frame statistics of every frame (the backward reuses the forward's)
        self.ymin_key = _empty(3 * F, torch.int64)
        self.idx_bounds = _empty(4 * F, torch.int32)
        self._size_batch(self._rows_for(zb))
        B = self.batch
        self.status = torch.zeros(4, dtype=torch.int32, device=dev)
        nb = g.c_n ** 3
        self.bin_sum = _empty(nb, torch.int64)
        self.bin_cnt = _empty(nb, torch.int32)
        self.iq64 = _empty(nb, torch.float64)

        # adjoint
        self.adj_row_exp = _empty(N, torch.int32)
        self.ainv = _empty(n, torch.int32)
        self.g_bins = _empty(nb, torch.float64)
        self.red = _empty(max(B * (N + 1), 2 * (1 + self.n_elem) * _lib.GX_ADJ_RED_BLOCKS), torch.float64)
        self.amin = _empty(B + 7, torch.int32)
        # fill_bkg / smooth tables (per batch) and the background-level adjoint
        self.col_nm = _empty(B * N, torch.int32) if self.fill else None
        self.col_sy = _empty(B * N, torch.float64) if self.smooth else None
        self.col_sz = _empty(B * N, torch.float64) if self.smooth else None
        self.col_spec = _empty(B * W * 2, cdt) if self.dense else None
        if self.smooth:
            cw, self.smooth_radius = gauss_prefix(self.smooth)
            self.gauss = torch.as_tensor(cw).to(dev).contiguous()
        else:
            self.gauss, self.smooth_radius = None, 0
        self.avg_spec = _empty(B * (N + W) * 2, torch.float64) if self.dense else None
        self.avg_part = _empty(B * W * 2, torch.float64) if self.dense else None
        self.g_avg = _empty(2, torch.float64)
        self._records_for = (xb, yb)
        self.g_atom = _empty(5 * n, torch.float64)
        self.e_sum = _empty(2 * self.n_elem, torch.float64)
        self.g_coords = _empty(3 * n, torch.float64)
        self.g_f = _empty(2 * self.n_elem, torch.float64)

        # detector
        d = dict(detector)
        psis, pw = _uniform(d['psis'])
        phis, phw = _uniform(d.get('phis', [0.0]))
        ths, thw = _uniform(d.get('thetas', [0.0]))
        pw = pw if d.get('psi_weights') is None else np.asarray(d['psi_weights'], dtype=np.float64)
        phw = phw if d.get('phi_weights') is None else np.asarray(d['phi_weights'], dtype=np.float64)
        thw = thw if d.get('theta_weights') is None else np.asarray(d['theta_weights'], dtype=np.float64)
        dp = detector_plan(int(d['num_pixels']), float(d['max_q']), d.get('angle_init_vals', ()),
                           d.get('angle_init_axs', ()), psis, pw, phis, phw, ths, thw)
        self.n_pix = npx = dp.n_pix
        self.mirror = bool(d.get('mirror', True))
        self.det_h = torch.as_tensor(dp.h).to(dev)
        self.det_v = torch.as_tensor(dp.v).to(dev)
        self.det_iR = torch.as_tensor(dp.init_R).to(dev).contiguous()
        self.det_oR = torch.as_tensor(dp.orient_R).to(dev).contiguous()
        self.det_ow = torch.as_tensor(dp.orient_w).to(dev).contiguous()
        self.det_raw = _empty(npx * npx, torch.float64)
        self.det_out = _empty(npx * npx, torch.float64)
        self.det_work = _empty(3 * npx * npx + (npx * npx + 255) // 256, torch.float64)
        self.fx_scale = _empty(1, torch.float64)
        ax = g.crop_axis
        self.detector_desc = GxDetector(
            n_pix=npx, n_init=len(dp.init_R), n_orient=len(dp.orient_w),
            h_axis=ptr(self.det_h).value, v_axis=ptr(self.det_v).value,
            init_R=ptr(self.det_iR).value, orient_R=ptr(self.det_oR).value,
            orient_w=ptr(self.det_ow).value, iq=ptr(self.iq64).value, iq_f64=1,
            cy=g.c_n, cx=g.c_n, cz=g.c_n, qx_min=float(np.min(ax)), qy_min=float(np.min(ax)),
            qz_min=float(np.min(ax)), dq=float(np.diff(ax)[0]), det=ptr(self.det_raw).value)
        # the detector adjoint goes through a transpose map of the orientation gather
        # when it fits (<= 32 GB of runs of pixels, 9 bytes a run, of the 288 GB of HBM),
        # else scatters into copies of its accumulator (<= 256 MB)
        pbits = int(np.ceil(np.log2(max(npx * npx, 1))))
        self.det_map = None
        if pbits < 32 and len(dp.orient_w) <= 2 ** (32 - pbits) and nb < 2 ** 31:
            # the map depends on the detector geometry and the q grid only: models of
            # the same geometry (a batch of slabs) share the last one built
            key = hashlib.sha1(b''.join(np.ascontiguousarray(a, dtype=np.float64).tobytes() for a in
                                        (dp.h, dp.v, dp.init_R, dp.orient_R, np.array(
                                            [g.c_n, np.min(ax), np.diff(ax)[0], dev.index or 0])))).hexdigest()
            if key not in _DET_MAPS:
                _DET_MAPS.clear()
                _DET_MAPS[key] = self._build_map(nb)
            self.det_map = _DET_MAPS[key]
        self.mapped = self.det_map is not None
        self.n_rep = 1 if self.mapped else int(max(1, min(8, 256e6 // (8 * nb))))
        self.g_fx = _empty(self.n_rep * nb, torch.int64)
        self._generation = 0
        self._prepared = None

    # ----------------------------------------------------------------------------
    def _build_map(self, nb):
        """The detector's transpose map as runs of pixels (gx_detector_map_count, then
        gx_detector_map_fill into arrays of exactly the counted size): (run_voxel, runs,
        run_len, n_runs), or None when the runs would take more than MAP_BYTES_MAX."""
        st = stream_handle()
        cnt = _empty(nb, torch.int32)
        scan = _empty(-(-nb // 2048) + 1, torch.int64)
        off = _empty(nb + 1, torch.int64)
        call("gx_detector_map_count", ctypes.byref(self.detector_desc), ptr(cnt), ptr(scan), ptr(off), st)
        n_runs = int(off[nb].item())
        if n_runs < 1 or n_runs * 9 > MAP_BYTES_MAX:
            return None
        runs = _empty(n_runs, torch.int32)
        run_voxel = _empty(n_runs, torch.int32)
        run_len = _empty(n_runs, torch.uint8)
        call("gx_detector_map_fill", ctypes.byref(self.detector_desc), ptr(off), ptr(cnt), ptr(runs),
             ptr(run_voxel), ptr(run_len), n_runs, st)
        return run_voxel, runs, run_len, n_runs

    # ----------------------------------------------------------------------------
    def _rows_for(self, zb):
        """Row range [0, Z) of a slab of z extent zb: rows v0 and v0 + 1 of every atom."""
        N, r = self.plan.N, self.r
        return int(min(N // 8 * 8, max(8, -(-(int(np.floor_divide(zb, r)) + 2) // 8) * 8)))

    def _size_batch(self, Z):
        """(Re)size the per-batch buffers for row range Z (with 16 rows of slack)."""
        N, W, n = self.plan.N, self.plan.W, self.n
        Zc = int(min(N, Z + 16))
        # kept rows of the column transforms: the forward stores them for the backward of
        # its last batch (power-of-two grids)
        K = int(self.plan.i_hi - self.plan.i_lo + 1) if self.bs is None else 0
        nf = self.frames[1] - self.frames[0]
        esz = 16 if self.f64 else 8
        B = self._batch_arg
        if B is None:
            # frame buffer + per-atom partial gradients of a batch; up to a quarter of
            # the free device memory (at most 64 GB, at least 2.5 GB), so that a slab's
            # frames fit in as few batches as possible (the backward reuses the
            # forward's last batch instead of recomputing it)
            per_frame = W * Zc * esz + 4 * (esz // 2) * max(n, 1) + W * K * esz
            free, _ = torch.cuda.mem_get_info()
            budget = max(2.5e9, min(0.25 * free, 64e9))
            B = max(1, min(nf, int(budget // per_frame), 8192))
        if getattr(self, 'batch', None) not in (None, B):
            raise RuntimeError("internal: batch size changed")
        self.batch = int(B)
        self.Z_cap = Zc
        cdt = torch.float64 if self.f64 else torch.float32
        self.inter = _empty(self.batch * W * Zc * 2, cdt)
        self.f_band = _empty(self.batch * W * K * 2, cdt) if K else None
        self.g_part = _empty(self.batch * 4 * n, torch.float64 if self.f64 else torch.float32)
        # per-element dL/df of each frame's row blocks (at most Zc blocks per frame)
        self.e_part = (_empty(self.batch * Zc * self.n_elem * 2, torch.float64)
                       if self.n_elem <= _lib.GX_ADJ_ELEM_REG else None)

    def batches(self, fr):
        """(f0, nf, descriptor) per batch: the descriptor's statistics pointers
        address frame f0's slots of the all-frame arrays."""
        lo, hi = self.frames
        out = []
        for f0 in range(lo, hi, self.batch):
            d = GxFrames.from_buffer_copy(fr)
            d.ymin_key = self.ymin_key.data_ptr() + 8 * 3 * f0
            d.idx_bounds = self.idx_bounds.data_ptr() + 4 * 4 * f0
            out.append((f0, min(self.batch, hi - f0), d))
        return out

    def __call__(self, coords, f_re=None, f_im=None):
        f_re = self.f_re0 if f_re is None else f_re
        f_im = self.f_im0 if f_im is None else f_im
        return _DetectorImage.apply(coords, f_re, f_im, self)

    def descriptors(self, coords, f_re, f_im):
        """Upload atoms and f, regroup the slab and return (gx_frames, gx_adjoint)."""
        dev = device()
        c = torch.as_tensor(coords).detach().to(device=dev, dtype=torch.float64)
        if c.shape != (self.n, 3):
            raise ValueError(f"coords must be ({self.n}, 3), got {tuple(c.shape)}")
        self.x.copy_(c[:, 0])
        self.y.copy_(c[:, 1])
        self.z.copy_(c[:, 2])
        ft = self.ftab.view(self.n_elem, 2)
        ft[:, 0].copy_(torch.as_tensor(f_re).detach().to(dev, torch.float64))
        ft[:, 1].copy_(torch.as_tensor(f_im).detach().to(dev, torch.float64))
        st = stream_handle()
        sb = GxSlab(n_atoms=self.n, n_elem=self.n_elem, x=ptr(self.x).value, y=ptr(self.y).value,
                    z=ptr(self.z).value, elem=ptr(self.elem).value)
        call("gx_slab_bounds", ctypes.byref(sb), ptr(self.minmax), ptr(self.ecount), st)
        mm = self.minmax.cpu().numpy()
        counts = self.ecount.cpu().numpy()
        xb, yb, zb, zmin = mm[1] - mm[0], mm[3] - mm[2], mm[5] - mm[4], float(mm[4])
        N, r = self.plan.N, self.r
        if math.hypot(xb, yb) / r + 2.0 >= N or zb / r + 2.0 >= N:
            raise ValueError(f"slab extent ({xb:.3f}, {yb:.3f}, {zb:.3f}) A does not fit the "
                             f"{N} x {N} grid of {r} A voxels with a one-voxel margin")
        Z = self._rows_for(zb)
        if Z > self.Z_cap:
            self._batch_arg = self.batch
            self._size_batch(Z)
        slab = GxSlab(n_atoms=self.n, n_elem=self.n_elem, N=N, r=r, zmin=zmin,
                      x=ptr(self.x).value, y=ptr(self.y).value, z=ptr(self.z).value,
                      elem=ptr(self.elem).value, f_table=ptr(self.ftab).value,
                      sx=ptr(self.sx).value, sy=ptr(self.sy).value, sq=ptr(self.sq).value,
                      szrow=ptr(self.szrow).value, row_off=ptr(self.row_off).value,
                      row_exp=ptr(self.row_exp).value, work=ptr(self.work).value,
                      szf=ptr(self.szf).value, sperm=ptr(self.sperm).value,
                      selem=ptr(self.selem).value, stable_work=ptr(self.stable_work).value)
        call("gx_slab_sort_rows", ctypes.byref(slab), st)
        ftab = ft.cpu().numpy()
        fc = ftab[:, 0] + 1j * ftab[:, 1]
        sum_abs = float(np.sum(counts * np.abs(fc)))
        nf = self.frames[1] - self.frames[0]
        # background level and grid constants of the current bounding box
        # (tools/comparison.py:741-744, tools/voxelgrids.py:343)
        vol = xb * yb * zb
        self.avg = (np.sum(counts * fc) / vol) * r ** 3 if self.dense else 0.0
        max_voxels = np.sqrt(np.float64(xb) ** 2 + np.float64(yb) ** 2) // r
        if self.fill and self._records_for != (xb, yb):
            fpr = self.plan.fparams.copy()
            fpr[:, _lib.FP_LEN_MODE:_lib.FP_CT + 1] = collapse_records(self.plan.grid.phis, yb, xb)
            self.fparams.copy_(torch.as_tensor(fpr))
            self._records_for = (xb, yb)
        B = intensity_bound(sum_abs, N, self.fill, self.smooth, self.avg, max_voxels, yb, xb, r)
        self.bin_scale = bin_scale_for([(nf, self._pixels_per_bin, B)])
        p, g = self.plan, self.plan.grid
        fr = GxFrames(
            N=N, r=r, n_atoms=self.n, ax=ptr(self.sx).value, ay=ptr(self.sy).value,
            aq=ptr(self.sq).value, azrow=ptr(self.szrow).value, row_off=ptr(self.row_off).value,
            row_exp=ptr(self.row_exp).value, fill_bkg=int(self.fill), smooth=self.smooth,
            max_voxels=float(max_voxels), avg_re=float(np.real(self.avg)),
            avg_im=float(np.imag(self.avg)), gauss_cw=ptr(self.gauss).value,
            smooth_radius=self.smooth_radius, col_nm=ptr(self.col_nm).value,
            col_sy=ptr(self.col_sy).value, col_sz=ptr(self.col_sz).value,
            col_spec=ptr(self.col_spec).value,
            zrow_max=int(np.floor_divide(zb, r)), qmin=g.qmin, dq=g.dq, c_lo=g.c_lo, c_n=g.c_n,
            q_val=p.q_val, i_lo=p.i_lo, i_hi=p.i_hi, W=p.W, Z=Z, bin_scale=self.bin_scale,
            fft_f64=int(self.f64), twiddle=ptr(self.tw).value, twiddle64=ptr(self.tw64).value,
            bs_M=0, n_frames=p.n_frames, fparams=ptr(self.fparams).value, batch=self.batch,
            ymin_key=ptr(self.ymin_key).value, idx_bounds=ptr(self.idx_bounds).value,
            inter=ptr(self.inter).value, bin_sum=ptr(self.bin_sum).value,
            bin_cnt=ptr(self.bin_cnt).value, status=ptr(self.status).value,
            f_band=ptr(self.f_band).value if self.f_band is not None else None)
        if self.bs is not None:
            b = self.bs
            fr.bs_M, fr.bs_Mc = b['M'], b['Mc']
            fr.bs_rlo, fr.bs_rk, fr.bs_clo, fr.bs_ck = b['rlo'], b['rk'], b['clo'], b['ck']
            fr.bs_chirp, fr.bs_kernel, fr.bs_twiddle = ptr(b['chirp']).value, ptr(b['K']).value, ptr(b['tw']).value
            fr.bs_kernel_c, fr.bs_twiddle_c = ptr(b['Kc']).value, ptr(b['twc']).value
        adj = GxAdjoint(
            n_elem=self.n_elem, f_table=ptr(self.ftab).value, afz=ptr(self.szf).value,
            aperm=ptr(self.sperm).value, aelem=ptr(self.selem).value, ecnt=ptr(self.work).value,
            z=ptr(self.z).value, zmin=zmin, row_exp=ptr(self.adj_row_exp).value,
            ainv=ptr(self.ainv).value, g_bins=ptr(self.g_bins).value,
            g_part=ptr(self.g_part).value, red=ptr(self.red).value, amin=ptr(self.amin).value,
            g_atom=ptr(self.g_atom).value, x=ptr(self.x).value, y=ptr(self.y).value,
            ext=ptr(self.minmax).value, ecount=ptr(self.ecount).value, vol_r3=r ** 3 / vol,
            avg_spec=ptr(self.avg_spec).value, avg_part=ptr(self.avg_part).value,
            g_avg=ptr(self.g_avg).value, e_part=ptr(self.e_part).value, e_sum=ptr(self.e_sum).value,
            bs_kernel_adj=ptr(self.bs['Kadj']).value if self.bs else None,
            bs_kernel_adj_c=ptr(self.bs['Kadj_c']).value if self.bs else None)
        return fr, adj

    def _check_status(self):
        s = self.status.cpu().numpy()
        if s[1]:
            raise RuntimeError(f"differentiable path: {int(s[1])} frame(s) outside the planned bands (a "
                               "frame's kept columns, or its atoms' y rows beyond the chirp-z row band "
                               "sized to the slab the model was built for): rebuild the model for the "
                               f"new slab extent (status {s.tolist()})")
        if s[2]:
            raise RuntimeError(f"internal: differentiable path status {s.tolist()}")

    def _forward(self, coords, f_re, f_im):
        fr, adj = self.descriptors(coords, f_re, f_im)
        self.status.zero_()
        st = stream_handle()
        call("gx_adj_prepare", ctypes.byref(fr), ctypes.byref(adj), st)
        g = self.plan.grid
        call("gx_bins_clear", ptr(self.bin_sum), ptr(self.bin_cnt), g.c_n ** 3, st)
        for f0, nf, d in self.batches(fr):
            call("gx_adj_forward", ctypes.byref(d), ctypes.byref(adj), f0, nf, st)
        self.finalize_image(st)
        self._check_status()
        self._prepared = (fr, adj)
        self._inter_fr = fr
        self._generation += 1
        return self.det_out.view(self.n_pix, self.n_pix).clone()

    def finalize_image(self, st):
        g = self.plan.grid
        start, step, stop = g.linspace_params
        aff = _lib.dbl_array(aff_dict['C'])
        call("gx_bins_finalize", ptr(self.bin_sum), ptr(self.bin_cnt), g.c_n, self.bin_scale, g.c_lo,
             g.q_num, start, step, stop, aff, float(ptable['C']), ptr(self.iq64), ptr(None), st)
        call("gx_detector_accumulate", ctypes.byref(self.detector_desc), st)
        call("gx_detector_finalize", ptr(self.det_raw), self.n_pix, self.n_pix, self._flags(),
             ptr(self.det_out), st)

    def _flags(self):
        return _lib.GX_FIN_FLOOR_SCALE | (_lib.GX_FIN_MIRROR if self.mirror else 0)

    def _backward(self, g_out):
        fr, adj = self._prepared
        st = stream_handle()
        go = g_out.detach().to(device=device(), dtype=torch.float64).contiguous()
        self.backward_enqueue(fr, adj, go, st)
        self._check_status()
        return (self.g_coords.view(self.n, 3).clone(), self.g_f.view(self.n_elem, 2).clone())

    def backward_enqueue(self, fr, adj, g_out, st):
        """Enqueue the whole adjoint for the image cotangent ``g_out`` [n_pix^2]."""
        g = self.plan.grid
        if self.mapped:
            call("gx_detector_adjoint_mapped", ctypes.byref(self.detector_desc), ptr(g_out), self._flags(),
                 ptr(self.det_work), ptr(self.det_map[0]), ptr(self.det_map[1]), ptr(self.det_map[2]),
                 self.det_map[3], ptr(self.g_fx), ptr(self.fx_scale), st)
        else:
            call("gx_detector_adjoint", ctypes.byref(self.detector_desc), ptr(g_out), self._flags(),
                 ptr(self.det_work), ptr(self.g_fx), self.n_rep, ptr(self.fx_scale), st)
        start, step, stop = g.linspace_params
        aff = _lib.dbl_array(aff_dict['C'])
        call("gx_bins_adjoint", ptr(self.g_fx), self.n_rep, ptr(self.fx_scale), ptr(self.bin_cnt), g.c_n, g.c_lo,
             g.q_num, start, step, stop, aff, float(ptable['C']), ptr(self.g_bins), st)
        call("gx_adj_prepare", ctypes.byref(fr), ctypes.byref(adj), st)
        # last batch first: the frame buffer still holds its row pass when the forward
        # of these descriptors was the last thing enqueued on this model
        stored = getattr(self, '_inter_fr', None) is fr
        self._inter_fr = None
        for i, (f0, nf, d) in enumerate(reversed(self.batches(fr))):
            name = "gx_adj_backward_stored" if (stored and i == 0) else "gx_adj_backward"
            call(name, ctypes.byref(d), ctypes.byref(adj), f0, nf, st)
        call("gx_adj_finish", ctypes.byref(fr), ctypes.byref(adj), ptr(self.g_coords), ptr(self.g_f), st)

    def forward_enqueue(self, fr, adj, st):
        """Enqueue the forward (bins cleared, every frame, image) for prepared descriptors."""
        g = self.plan.grid
        call("gx_adj_prepare", ctypes.byref(fr), ctypes.byref(adj), st)
        call("gx_bins_clear", ptr(self.bin_sum), ptr(self.bin_cnt), g.c_n ** 3, st)
        for f0, nf, d in self.batches(fr):
            call("gx_adj_forward", ctypes.byref(d), ctypes.byref(adj), f0, nf, st)
        self.finalize_image(st)
        self._inter_fr = fr


def image_loss(image, target, scale, background):
    """sum (scale * image + background - target)^2 (torch; differentiable in
    every argument)."""
    return torch.sum((scale * image + background - target) ** 2)


def fit_scale_background(image, target):
    """Least-squares scale and flat background of ``image`` onto ``target``
    (optimize_scale_offset, tools/comparison.py:873-882, without a mask)."""
    dev = device()
    s = image.detach().to(device=dev, dtype=torch.float64).contiguous().reshape(-1)
    y = target.detach().to(device=dev, dtype=torch.float64).contiguous().reshape(-1)
    m = torch.zeros(s.numel(), dtype=torch.uint8, device=dev)
    work = torch.empty(_lib.GX_MOMENTS_WORK, dtype=torch.float64, device=dev)
    res = (ctypes.c_double * 6)()
    call("gx_scale_offset_moments", ptr(s), ptr(y), ptr(m), int(s.numel()), ptr(work), res,
         stream_handle())
    ms, my, sss, ssy = res[1], res[2], res[3], res[4]
    scale = ssy / sss if sss != 0 else 0.0
    return scale, my - scale * ms
