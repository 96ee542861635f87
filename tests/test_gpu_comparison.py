"""GPU parity of the experimental-comparison step (giwaxsim_amd.tools.comparison)
against the oracle (oracle/comparison.py: the reference's functions with its scipy
calls), on the same seeded inputs:

* gx_map_linear vs scipy.ndimage.map_coordinates(order=1, mode='constant') at random
  positions including exact edges and points just outside: max |diff| <= 1e-14 max|img|;
* linear_polar, polar_linear, add_pad, shift_peak (drop-in, numpy in and out) on a
  detector-like image with exact zeros: <= 1e-12 of max|image|; add_pad exact;
* rebin_qmap_and_mask vs griddata: same NaN positions, values and masks bit-exact
  (griddata's barycentric arithmetic, FMA-free kernel);
* FitComparison.residual (trim, shift_peak, scale / offset, masked residual, all on
  the device) vs the oracle's fit_residual: <= 1e-10 of max|residual|, scale and
  offset to 1e-10;
* fit_slabsize.fullfit_model end to end (slab builder, GPU voxel grid and detector,
  device comparison) vs the oracle pipeline on the graphite test cell: <= 1e-4 of the
  experimental map's scale (the image itself carries the FP32-transform error).
"""

import sys

import numpy as np
import pytest
import torch
from scipy.ndimage import map_coordinates

from giwaxsim_amd import engine
from giwaxsim_amd.tools import comparison as G
from oracle import comparison as C

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _image(shape, seed):
    """Smooth positive field with a zero border and zero holes (like a mirrored,
    floored detector image after the 1e-6 scale)."""
    rng = np.random.default_rng(seed)
    h, w = shape
    y, x = np.mgrid[:h, :w] / np.array([h, w])[:, None, None]
    img = 1e-6 * (1.5 + np.sin(7 * x + 3 * y) + 0.3 * rng.random(shape))
    img[:2] = 0.0
    img[:, -3:] = 0.0
    img[rng.random(shape) < 0.02] = 0.0
    return img


def test_map_linear_matches_scipy():
    _say("[gpu-compare] map_linear vs scipy map_coordinates order=1 constant")
    rng = np.random.default_rng(1)
    img = rng.normal(size=(37, 53))
    h, w = img.shape
    ys = rng.uniform(-2, h + 1, size=50000)
    xs = rng.uniform(-2, w + 1, size=50000)
    ys[:6] = [0, h - 1, h - 1, 0, np.nextafter(h - 1, h), -1e-300]
    xs[:6] = [0, w - 1, 0, w - 1, 3.0, 2.0]
    want = map_coordinates(img, (ys, xs), order=1, mode='constant', cval=0.25)
    got = G._map_linear_dev(G._as_dev(img), G._as_dev(ys), G._as_dev(xs), 0.25).cpu().numpy()
    err = np.max(np.abs(got - want)) / np.max(np.abs(img))
    _say(f"  max |diff| / max|img| = {err:.2e}")
    assert err <= 1e-14
    assert np.array_equal(got == 0.25, want == 0.25)


def _geometry():
    det_h = np.linspace(-2.5, 2.5, 160)
    det_v = np.linspace(-2.5, 2.5, 160)
    exp_qxy = np.linspace(0.02, 2.1, 300)
    exp_qz = np.linspace(-0.03, 1.9, 260)
    return det_h, det_v, exp_qxy, exp_qz


def test_polar_warps_and_shift_peak_match_oracle():
    det_h, det_v, exp_qxy, exp_qz = _geometry()
    img = _image((160, 160), 2)
    sim, h_t, v_t = C.trim_sim_data(img, det_h, det_v, exp_qxy, exp_qz)
    sim_g, h_g, v_g = G.trim_sim_data(img, det_h, det_v, exp_qxy, exp_qz)
    assert np.array_equal(sim, sim_g) and np.array_equal(h_t, h_g) and np.array_equal(v_t, v_g)
    _say(f"[gpu-compare] trimmed image {sim.shape}: linear_polar, polar_linear, add_pad, shift_peak")
    scale = np.max(np.abs(sim))
    o = (int(np.argmin(np.abs(v_t))), int(np.argmin(np.abs(h_t))))
    radius = float(np.hypot(*sim.shape))
    p_ref = C.linear_polar(sim, o=o, r=radius)
    p_gpu = G.linear_polar(sim, o=o, r=radius)
    assert p_gpu.shape == p_ref.shape
    e1 = np.max(np.abs(p_gpu - p_ref)) / scale
    back_ref = C.polar_linear(p_ref, o=o, r=None, output=sim.shape)
    back_gpu = G.polar_linear(p_ref, o=o, r=None, output=sim.shape)
    e2 = np.max(np.abs(back_gpu - back_ref)) / scale
    r_axis = np.linspace(0, 2.9, p_ref.shape[1])
    pad_ref = C.add_pad(p_ref.copy(), r_axis, 0.06, (1.2, 1.9))
    pad_gpu = G.add_pad(p_ref.copy(), r_axis, 0.06, (1.2, 1.9))
    assert pad_gpu.shape == pad_ref.shape and np.array_equal(pad_gpu, pad_ref)
    sp_ref = C.shift_peak(sim.copy(), h_t, v_t, 0.06, (1.2, 1.9))
    sp_gpu = G.shift_peak(sim.copy(), h_t, v_t, 0.06, (1.2, 1.9))
    e3 = np.max(np.abs(sp_gpu - sp_ref)) / scale
    _say(f"  linear_polar {e1:.2e}  polar_linear {e2:.2e}  shift_peak {e3:.2e}")
    assert e1 <= 1e-12 and e2 <= 1e-12 and e3 <= 1e-12
    assert np.array_equal(sp_gpu == 0, sp_ref == 0)


def _experiment(h_t, v_t, seed):
    rng = np.random.default_rng(seed)
    exp_qxy = np.linspace(0.02, 2.1, 300)
    exp_qz = np.linspace(-0.03, 1.9, 260)
    X, Z = np.meshgrid(exp_qxy, exp_qz)
    exp_img = 5.0 + np.cos(4 * X) * np.sin(3 * Z) + 0.1 * rng.random(X.shape)
    mask = np.zeros(X.shape)
    mask[100:104, 50:60] = 1
    mask[rng.random(X.shape) < 0.001] = 1
    return exp_img, mask, exp_qxy, exp_qz


def test_rebin_and_fit_residual_match_oracle():
    det_h, det_v, exp_qxy, exp_qz = _geometry()
    img = _image((160, 160), 3)
    _, h_t, v_t = C.trim_sim_data(img, det_h, det_v, exp_qxy, exp_qz)
    exp_img, mask, exp_qxy, exp_qz = _experiment(h_t, v_t, 4)
    _say(f"[gpu-compare] rebin {exp_img.shape} -> ({len(v_t)}, {len(h_t)}); device fit residual")
    r_ref, m_ref = C.rebin_qmap_and_mask(h_t, v_t, exp_img, exp_qxy, exp_qz, mask)
    r_gpu, m_gpu = G.rebin_qmap_and_mask(h_t, v_t, exp_img, exp_qxy, exp_qz, mask)
    assert np.array_equal(np.isnan(r_gpu), np.isnan(r_ref))
    ok = ~np.isnan(r_ref)
    e_rb = np.max(np.abs(r_gpu[ok] - r_ref[ok])) / np.max(np.abs(r_ref[ok]))
    assert np.array_equal(m_gpu, m_ref)
    rebin_map = np.where(np.isnan(r_ref), 0.0, r_ref)
    rebin_map[m_ref == 1] = 0
    res_ref, (s_ref, b_ref) = C.fit_residual(img, det_h, det_v, rebin_map, m_ref, exp_qxy, exp_qz, 0.06,
                                             (1.2, 1.9))
    fc = G.FitComparison(det_h, det_v, rebin_map, m_ref, exp_qxy, exp_qz, 0.06, (1.2, 1.9))
    res_gpu, s_gpu, b_gpu = fc.residual(torch.as_tensor(img).to(engine.device()))
    res_gpu = res_gpu.cpu().numpy()
    e_res = np.max(np.abs(res_gpu - res_ref)) / np.max(np.abs(res_ref))
    _say(f"  rebin {e_rb:.2e}; residual {e_res:.2e}; scale {s_gpu:.9e} vs {s_ref:.9e}; "
         f"offset {b_gpu:.9e} vs {b_ref:.9e}")
    assert e_rb == 0.0
    assert e_res <= 1e-10
    assert abs(s_gpu - s_ref) <= 1e-10 * abs(s_ref) and abs(b_gpu - b_ref) <= 1e-10 * abs(b_ref)


def test_fullfit_model_matches_oracle_pipeline():
    """fit_slabsize.fullfit_model (slab -> GPU voxel grid -> GPU detector -> device
    comparison) vs the oracle pipeline (voxelgridmaker, detectormaker, fit_residual) on
    the graphite test cell; the simulated image is the FP32-transform path (1e-5 above
    1e-6 max), so the residual is compared to 1e-4 of the experimental map's scale."""
    import os
    import oracle as O
    from giwaxsim_amd.fit_slabsize import fullfit_model
    from giwaxsim_amd.tools import comparison as GC
    golden = os.path.join(os.path.dirname(__file__), 'golden')
    _say("[gpu-compare] fullfit_model on graphite_medium.xyz 20 x 20 x 14 A, N=256, 64 px x 19 psi")
    qv = 2 * np.pi / (0.25 * 256) * (1 + 1e-9)
    psis = np.linspace(0, 180, 19)
    one = np.array([0.0])
    fixed_slab = (f'{golden}/graphite_medium.xyz', 2.456, 2.456, 6.696, 90.0, 90.0, 120.0)
    fixed_vox = (0.25, qv, 2.0, 12700.0, True, 2)
    fixed_det = (64, (0.15, 0, 0), ('psi', 'None', 'None'), psis, None, one, None, one, None)
    # experimental map on its own grid, rebinned onto the trimmed simulated grid
    exp_qxy = np.linspace(0.05, 1.9, 90)
    exp_qz = np.linspace(0.0, 1.85, 80)
    X, Z = np.meshgrid(exp_qxy, exp_qz)
    exp_img = 1e-4 * (1.5 + np.cos(6 * np.hypot(X, Z)))
    mask = np.zeros(X.shape)
    mask[:, :5] = 1
    h_full = np.linspace(-2.0, 2.0, 64)
    _, h_t, v_t = C.trim_sim_data(np.zeros((64, 64)), h_full, h_full, exp_qxy, exp_qz)
    rebin_map, rebin_mask = C.rebin_qmap_and_mask(h_t, v_t, exp_img, exp_qxy, exp_qz, mask)
    rebin_map = np.where(np.isnan(rebin_map), 0.0, rebin_map)
    rebin_map[rebin_mask == 1] = 0
    fixed_exp = (rebin_map, rebin_mask, exp_qxy, exp_qz, 0.05, (1.0, 1.6))
    params = {'x_size': 20.0, 'y_size': 20.0, 'z_size': 14.0}
    r_gpu = fullfit_model(params, fixed_slab, fixed_vox, fixed_det, fixed_exp)
    # oracle pipeline
    coords, el = GC.slabmaker_fitting(fixed_slab[0], 20.0, 20.0, 14.0, *fixed_slab[1:])
    ids, ftab, _, _ = engine.element_table(el, 12700.0)
    ref = O.voxelgridmaker(coords, ftab[ids], 0.25, qv, 2.0, fill_bkg=True, smooth=2)
    d, h, v = O.detectormaker(ref['iq'], ref['qx'], ref['qy'], ref['qz'], 64, 2.0, (0.15, 0, 0),
                              ('psi', 'None', 'None'), psis, None, one, None, one, None, mirror=True)
    assert np.array_equal(h, h_full)
    r_ref, _ = C.fit_residual(d, h, v, rebin_map, rebin_mask, exp_qxy, exp_qz, 0.05, (1.0, 1.6))
    err = np.max(np.abs(r_gpu - r_ref)) / np.max(np.abs(rebin_map))
    _say(f"  residual: max |gpu - oracle| / max|map| = {err:.2e}; |r| {np.linalg.norm(r_ref):.3e}")
    assert r_gpu.shape == r_ref.shape
    assert err <= 1e-4
