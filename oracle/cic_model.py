"""FP64 autograd reference of the differentiable (cloud-in-cell) variant.

TEST INFRASTRUCTURE ONLY: used by ``tests/`` as the checker of the hand-written
HIP adjoint, never by the product.

The reference (tchaney97/giwaxsim) has no gradients.  This model is the
oracle's forward path — ``frame_grid`` (tools/voxelgrids.py:319-341),
``frame_iq`` (387-392), ``frame_axes`` + ``bin_frame`` (381-402, 464-503),
the mean / crop / f0 of ``voxelgridmaker`` (tools/comparison.py:765-788) and
``detectormaker`` (tools/comparison.py:790-870) — with one change: the
nearest-grid-point deposit ``np.add.at(grid, (z // r, y' // r), f)`` is
replaced by cloud-in-cell weights,

    u = (y' - min y') / r,  v = (z - min z) / r,  u0 = u // 1,  tu = u - u0 (same for v)
    node (v0 + a, u0 + b) += f * [(1 - tv), tv][a] * [(1 - tu), tu][b],

written in torch (CPU, float64) so torch.autograd differentiates it.  Every
index (atom nodes, frame pixel -> bin, crop, detector voxel) comes from the
numpy oracle, exactly; only the values flow through torch.  With
``deposit='ngp'`` the weights collapse to the reference's deposit and the
model reproduces ``voxelgridmaker`` / ``detectormaker``, which is how the
model's index maps are pinned.
"""

from __future__ import annotations

import numpy as np
import torch

from . import giwaxs_oracle as O


def _frame_bin_map(N, r, phi, q_axis, c_lo, c_n):
    """(pixel, bin) pairs of one frame: pixel = flat index of the unshifted FFT
    output [kz, ky]; bin = flat index of the cropped [qy, qx, qz] grid.  The
    masks and floor divisions are ``bin_frame``'s (tools/voxelgrids.py:474-499)."""
    hx, hy, vz, _ = O.frame_axes(N, r, phi)
    qmin, qmax = np.min(q_axis), np.max(q_axis)
    hmask = (hx <= qmax) & (hx >= qmin) & (hy <= qmax) & (hy >= qmin)
    vmask = (vz <= qmax) & (vz >= qmin)
    js = np.nonzero(hmask)[0]
    iz = np.nonzero(vmask)[0]
    dq = np.diff(q_axis)[0]
    bx = O.floor_div(hx[js] - qmin, dq).astype(np.int64)
    by = O.floor_div(hy[js] - qmin, dq).astype(np.int64)
    bz = O.floor_div(vz[iz] - qmin, dq).astype(np.int64)
    I, J = np.meshgrid(np.arange(len(iz)), np.arange(len(js)), indexing='ij')
    BX, BY, BZ = bx[J], by[J], bz[I]
    keep = ((BX >= c_lo) & (BX < c_lo + c_n) & (BY >= c_lo) & (BY < c_lo + c_n) &
            (BZ >= c_lo) & (BZ < c_lo + c_n))
    # shifted row/column -> unshifted FFT index (fftshift of an even length)
    kz = (iz[I[keep]] + N // 2) % N
    ky = (js[J[keep]] + N // 2) % N
    pix = kz * N + ky
    b = ((BY[keep] - c_lo) * c_n + (BX[keep] - c_lo)) * c_n + (BZ[keep] - c_lo)
    return pix.astype(np.int64), b.astype(np.int64)


def _detector_maps(num_pixels, max_q, angle_init_vals, angle_init_axs, psis, psi_w, phis, phi_w,
                   thetas, theta_w, qx, qy, qz):
    """Per orientation: flat voxel index of every pixel into the cropped grid and
    the orientation weight (detectormaker + intersect_detector)."""
    det_x, det_y, det_z, _, _ = O.make_detector(max_q, num_pixels, max_q, num_pixels)
    det_x, det_y, det_z = O._init_rotations(det_x, det_y, det_z, angle_init_vals, angle_init_axs)
    dq = np.diff(qz)[0]
    idx, w = [], []
    for ps, pw in zip(psis, psi_w):
        for ph, phw in zip(phis, phi_w):
            for th, tw in zip(thetas, theta_w):
                x2, y2, z2 = O.rotate_psi_phi_theta(det_x, det_y, det_z, ps, ph, th)
                xi = np.clip(O.floor_div(x2.ravel() - np.min(qx), dq).astype(int), 0, len(qx) - 1)
                yi = np.clip(O.floor_div(y2.ravel() - np.min(qy), dq).astype(int), 0, len(qy) - 1)
                zi = np.clip(O.floor_div(z2.ravel() - np.min(qz), dq).astype(int), 0, len(qz) - 1)
                idx.append((yi * len(qx) + xi) * len(qz) + zi)
                w.append(pw * phw * tw)
    return np.array(idx, dtype=np.int64), np.array(w, dtype=np.float64)


def mirror_torch(q):
    """mirror_vertical_horizontal (tools/detector.py:246-275) on a torch image."""
    s = q + torch.flip(q, [1]) + torch.flip(q, [0]) + torch.flip(q, [0, 1])
    nr, nc = q.shape
    if nr % 2:
        c = nr // 2
        s = s.clone()
        s[c, :] = (q[c, :] + torch.flip(q[c, :], [0])) * 2
    if nc % 2:
        c = nc // 2
        s = s.clone()
        s[:, c] = (q[:, c] + torch.flip(q[:, c], [0])) * 2
    if nr % 2 and nc % 2:
        s[nr // 2, nc // 2] = q[nr // 2, nc // 2] * 4
    return s


class CicModel:
    """The differentiable variant of one slab, end to end, in torch FP64 (CPU).

    ``coords0`` fixes the plan (grid edge, q axis, crop, frames) the way
    voxelgrid_plan does; ``frames=(lo, hi)`` restricts the phi-frames;
    ``detector`` is a dict of detectormaker's arguments (num_pixels, max_q,
    angle_init_vals, angle_init_axs, psis, psi_weights, phis, phi_weights,
    thetas, theta_weights, mirror)."""

    def __init__(self, coords0, elem_ids, r_voxel_size, q_voxel_size, max_q, detector,
                 frames=None, deposit='cic', fill_bkg=False, smooth=0):
        plan = O.voxelgrid_plan(np.asarray(coords0), r_voxel_size, q_voxel_size, max_q)
        self.N = N = plan['grid_size']
        self.r = float(r_voxel_size)
        self.deposit = deposit
        self.fill, self.smooth = bool(fill_bkg), int(smooth)
        q_axis = plan['q_axis']
        voxel = np.abs(q_axis[1] - q_axis[0])
        keep = np.where(np.abs(q_axis) < max_q + voxel)[0]
        self.c_lo, self.c_n = int(keep[0]), int(len(keep))
        self.q_crop = q_axis[keep[0]:keep[-1] + 1]
        phis = plan['phis']
        if frames is not None:
            phis = phis[frames[0]:frames[1]]
        self.phis = phis
        self.elem = torch.as_tensor(np.asarray(elem_ids, dtype=np.int64))
        self.maps = []
        count = np.zeros(self.c_n ** 3)
        for phi in phis:
            rad = np.radians(phi)
            pix, b = _frame_bin_map(N, r_voxel_size, phi, q_axis, self.c_lo, self.c_n)
            np.add.at(count, b, 1.0)
            self.maps.append((phi, np.sin(rad), np.cos(rad), torch.as_tensor(pix), torch.as_tensor(b)))
        self.count = torch.as_tensor(count)
        ones = np.ones((self.c_n,) * 3)
        qc = self.q_crop
        self.f0w = torch.as_tensor(O.add_f0_q_3d(ones, qc, qc, qc).ravel())
        d = dict(detector)
        n = int(d['num_pixels'])
        psis, phis_d, thetas = (np.atleast_1d(np.asarray(d[k], dtype=np.float64))
                                for k in ('psis', 'phis', 'thetas'))
        pw = d.get('psi_weights')
        phw = d.get('phi_weights')
        tw = d.get('theta_weights')
        pw = np.ones_like(psis) / len(psis) if pw is None else np.asarray(pw)
        phw = np.ones_like(phis_d) / len(phis_d) if phw is None else np.asarray(phw)
        tw = np.ones_like(thetas) / len(thetas) if tw is None else np.asarray(tw)
        idx, w = _detector_maps(n, d['max_q'], d.get('angle_init_vals', ()),
                                d.get('angle_init_axs', ()), psis, pw, phis_d, phw, thetas, tw,
                                qc, qc, qc)
        self.n_pix = n
        self.det_idx = torch.as_tensor(idx)
        self.det_w = torch.as_tensor(w)
        self.mirror = bool(d.get('mirror', True))

    # ----------------------------------------------------------------------------
    def grid(self, coords, f_re, f_im, s, c, phi=0.0):
        """One frame's grid (real, imag), each [N*N] indexed z*N + y: the deposit,
        then fill_bkg and smoothing as frame_grid (tools/voxelgrids.py:343-379)."""
        N, r = self.N, self.r
        x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
        yp = x * s + y * c
        yy = yp - torch.min(yp)
        zz = z - torch.min(z)
        u0 = torch.as_tensor(O.floor_div(yy.detach().numpy(), r).astype(np.int64))
        v0 = torch.as_tensor(O.floor_div(zz.detach().numpy(), r).astype(np.int64))
        if self.deposit == 'ngp':
            tu = torch.zeros_like(yy)
            tv = torch.zeros_like(zz)
        else:
            tu = torch.clamp(yy * (1.0 / r) - u0.to(yy.dtype), 0.0, 1.0)
            tv = torch.clamp(zz * (1.0 / r) - v0.to(zz.dtype), 0.0, 1.0)
        if int(u0.max()) + 1 >= N or int(v0.max()) + 1 >= N:
            raise ValueError("atoms reach the last grid row/column")
        fr = f_re[self.elem]
        fi = f_im[self.elem]
        gr = torch.zeros(N * N, dtype=torch.float64)
        gi = torch.zeros(N * N, dtype=torch.float64)
        for a, wv in ((0, 1.0 - tv), (1, tv)):
            for b, wu in ((0, 1.0 - tu), (1, tu)):
                node = (v0 + a) * N + (u0 + b)
                w = wv * wu
                gr = gr.index_add(0, node, fr * w)
                gi = gi.index_add(0, node, fi * w)
        if not (self.fill or self.smooth):
            return gr, gi
        # avg_voxel_f over the bounding box (tools/comparison.py:741-744), differentiable
        xb = torch.max(x) - torch.min(x)
        yb = torch.max(y) - torch.min(y)
        zb = torch.max(z) - torch.min(z)
        vol = xb * yb * zb
        avg = torch.complex(torch.sum(fr) / vol * r ** 3, torch.sum(fi) / vol * r ** 3)
        xbn, ybn = float(xb.detach()), float(yb.detach())
        mv = np.sqrt(np.float64(xbn) ** 2 + np.float64(ybn) ** 2) // r
        y_min, y_max = int(u0.min()), int(u0.max())
        z_min, z_max = int(v0.min()), int(v0.max())
        G = torch.complex(gr, gi).reshape(N, N)
        A = avg * mv
        if self.fill:
            L = O.rectangular_collapse_lengths(np.arange(N) * r, ybn, xbn, phi)
            nm = torch.as_tensor((mv - O.floor_div(L, r)).astype(int).astype(np.float64))
            in_z = np.zeros(N)
            in_z[z_min:z_max] = 1
            in_y = np.zeros(N)
            in_y[y_min + 1:y_max] = 1
            inside = torch.as_tensor(np.outer(in_z, in_y))
            G = inside * (G + nm[None, :] * avg) + (1 - inside) * A
        if self.smooth:
            from scipy.ndimage import gaussian_filter1d
            sg = self.smooth
            mask_y = np.zeros(N)
            mask_z = np.zeros(N)
            mask_y[y_min + sg:y_max - sg] = 1
            mask_z[z_min + sg:z_max - sg] = 1
            m = torch.as_tensor(np.outer(gaussian_filter1d(mask_z, sigma=sg, mode='wrap'),
                                         gaussian_filter1d(mask_y, sigma=sg, mode='wrap')))
            G = G * m + A * (1 - m)
        G = G.reshape(-1)
        return G.real, G.imag

    def frame_intensity(self, coords, f_re, f_im, s, c, phi=0.0):
        """|FFT2|^2 of one frame, [N*N] unshifted [kz, ky]."""
        gr, gi = self.grid(coords, f_re, f_im, s, c, phi)
        F = torch.fft.fft2(torch.complex(gr, gi).reshape(self.N, self.N))
        return (F.real ** 2 + F.imag ** 2).reshape(-1)

    def iq(self, coords, f_re, f_im):
        """Cropped, f0-weighted mean I(q), flat [c^3] ([qy, qx, qz])."""
        total = torch.zeros(self.c_n ** 3, dtype=torch.float64)
        for phi, s, c, pix, b in self.maps:
            I = self.frame_intensity(coords, f_re, f_im, s, c, phi)
            total = total.index_add(0, b, I[pix])
        cnt = self.count
        mean = torch.where(cnt > 0, total / torch.clamp(cnt, min=1.0), torch.zeros_like(total))
        return mean * self.f0w

    def detector(self, iq):
        raw = (iq[self.det_idx] * self.det_w[:, None]).sum(0).reshape(self.n_pix, self.n_pix)
        v = mirror_torch(raw) if self.mirror else raw
        v = torch.where(v > 0, v, torch.full_like(v, 1e-6))
        return v * 1e-6

    def __call__(self, coords, f_re, f_im):
        return self.detector(self.iq(coords, f_re, f_im))
