"""gx_workspace_size against the buffers the package itself allocates and runs with:
the frame runner, slab and bins of a forward run, and every buffer of a differentiable
slab (power-of-two and chirp-z grid edges) after a forward and a backward, byte for byte.
The per-frame statistics are kept for all frames (sized with batch = n_frames), the other
per-frame buffers for one batch."""

import sys

import numpy as np
import pytest
import torch

from giwaxsim_amd import _lib, engine
from giwaxsim_amd.adjoint import DifferentiableSlab
from giwaxsim_amd.synth import graphite_slab

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _nbytes(t):
    return t.numel() * t.element_size()


def _query(desc, n_elem, **extra):
    return dict(N=desc.N, n_atoms=desc.n_atoms, n_elem=n_elem, n_frames=desc.n_frames, W=desc.W, Z=desc.Z,
                c_n=desc.c_n, i_lo=desc.i_lo, i_hi=desc.i_hi, fill_bkg=desc.fill_bkg, smooth=desc.smooth,
                fft_f64=desc.fft_f64, bs_M=desc.bs_M, bs_Mc=desc.bs_Mc, **extra)


@pytest.mark.parametrize("fill,smooth", [(False, 0), (True, 3)])
def test_forward_buffers(fill, smooth):
    _say(f"[gpu-workspace] forward graphite 16x16x12 N=256 fill={fill} smooth={smooth}")
    coords, el = graphite_slab(size=(16.0, 16.0, 12.0), seed=21)
    r = 0.25
    qv = 2 * np.pi / (r * 256) * (1 + 1e-9)
    ids, ftab, _, _ = engine.element_table(el, 12700.0)
    res = engine.run_voxelgrid(coords, el, r, qv, 2.5, 12700.0, fill_bkg=fill, smooth=smooth, f_table=ftab,
                               elem_ids=ids)
    run, slab, bins = res['runner'], res['slab'], res['bins']
    d = run.desc
    q = _query(d, len(ftab))
    per_batch = _lib.workspace_size(**q, batch=d.batch)
    all_frames = _lib.workspace_size(**q, batch=d.n_frames)
    have = dict(sx=slab.sx, sy=slab.sy, sq=slab.sq, szrow=slab.szrow, row_off=slab.row_off,
                row_exp=slab.row_exp, work=slab.work, fparams=run.fparams, twiddle=run.tw, twiddle64=run.tw64,
                inter=run.inter, bin_sum=bins.sum, bin_cnt=bins.cnt, status=run.status, avg_q=run.avg_q,
                iq=res['iq64'])
    stats = dict(ymin_key=run.ymin_key, idx_bounds=run.idx_bounds)
    if fill:
        stats['col_nm'] = run.col_nm
    if smooth:
        stats.update(col_sy=run.col_sy, col_sz=run.col_sz)
        have['gauss_cw'] = run.gauss
    if fill or smooth:
        stats['col_spec'] = run.col_spec
    for k, t in have.items():
        assert _nbytes(t) == per_batch[k], (k, _nbytes(t), per_batch[k])
    for k, t in stats.items():
        assert _nbytes(t) == all_frames[k], (k, _nbytes(t), all_frames[k])
    for k in ('col_nm', 'col_sy', 'col_sz', 'col_spec', 'gauss_cw'):
        if k not in have and k not in stats:
            assert per_batch[k] == 0, k


@pytest.mark.parametrize("N", [256, 300])
def test_differentiable_buffers(N):
    _say(f"[gpu-workspace] differentiable graphite 16x16x12 N={N} fill smooth 3, 48^2 detector")
    coords, el = graphite_slab(size=(16.0, 16.0, 12.0), seed=21)
    r = 0.25
    qv = 2 * np.pi / (r * N) * (1 + 1e-9)
    det = dict(num_pixels=48, max_q=2.5, psis=list(np.arange(0.0, 91.0, 10.0)), phis=[0.0], thetas=[0.0],
               angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    m = DifferentiableSlab(el, coords, r, qv, 2.5, det, frames=(0, 10), fill_bkg=True, smooth=3)
    c = torch.tensor(coords, device='cuda', requires_grad=True)
    m(c).sum().backward()
    torch.cuda.synchronize()
    fr, _ = m.descriptors(c.detach(), m.f_re0, m.f_im0)
    q = _query(fr, m.n_elem, differentiable=1, store_band=int(m.f_band is not None), n_pix=m.n_pix,
               n_init=int(m.detector_desc.n_init), n_orient=int(m.detector_desc.n_orient))
    q['Z'] = m.Z_cap   # the model sizes its batch buffers for its row capacity
    s = _lib.workspace_size(**q, batch=m.batch)
    all_frames = _lib.workspace_size(**q, batch=fr.n_frames)
    have = dict(sx=m.sx, sy=m.sy, sq=m.sq, szrow=m.szrow, row_off=m.row_off, row_exp=m.row_exp, work=m.work,
                szf=m.szf, sperm=m.sperm, selem=m.selem, stable_work=m.stable_work, twiddle=m.tw,
                twiddle64=m.tw64, gauss_cw=m.gauss, col_nm=m.col_nm, col_sy=m.col_sy, col_sz=m.col_sz,
                col_spec=m.col_spec, inter=m.inter, bin_sum=m.bin_sum, bin_cnt=m.bin_cnt, status=m.status,
                iq=m.iq64, adj_row_exp=m.adj_row_exp, ainv=m.ainv, g_bins=m.g_bins, g_part=m.g_part, red=m.red,
                amin=m.amin, g_atom=m.g_atom, avg_spec=m.avg_spec, avg_part=m.avg_part, g_avg=m.g_avg,
                e_part=m.e_part, e_sum=m.e_sum, g_coords=m.g_coords, g_f=m.g_f, det=m.det_raw,
                det_out=m.det_out, det_work=m.det_work, fx_scale=m.fx_scale, g_fx=m.g_fx,
                h_axis=m.det_h, v_axis=m.det_v, init_R=m.det_iR, orient_R=m.det_oR, orient_w=m.det_ow)
    if m.f_band is not None:
        have['f_band'] = m.f_band
    if m.bs is not None:
        b = m.bs
        have.update(bs_chirp=b['chirp'], bs_kernel=b['K'], bs_twiddle=b['tw'], bs_kernel_c=b['Kc'],
                    bs_twiddle_c=b['twc'], bs_kernel_adj=b['Kadj'], bs_kernel_adj_c=b['Kadj_c'])
    assert m.mapped and m.n_rep == 1
    for k, t in have.items():
        assert _nbytes(t) == s[k], (k, _nbytes(t), s[k])
    for k in ('ymin_key', 'idx_bounds'):
        assert _nbytes(getattr(m, k)) == all_frames[k], k
    run_voxel, runs, run_len, n_runs = m.det_map
    assert _nbytes(runs) + _nbytes(run_voxel) + _nbytes(run_len) == n_runs * s['map_run_bytes']
    assert n_runs <= s['map_runs_max']
