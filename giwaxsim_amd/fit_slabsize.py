"""Slab-size fit of a simulated GIWAXS image to an experimental one, mirroring the
reference's ``fit_slabsize.py`` (fullfit_model 19-47).

``fullfit_model`` keeps the reference's signature and returns the residual vector
(numpy).  ``SlabSizeModel`` is the same model with its experimental side planned once
(``tools.comparison.FitComparison``) and its unit cell uploaded once
(``tools.comparison.SlabBuilder``): each evaluation tiles and crops the slab, runs the
voxel grid and the detector image, and trims, shift_peaks, fits scale and offset and
forms the masked residual, all on the device; only the residual vector comes back.  The reference drives it with lmfit's Nelder-Mead (lmfit 1.x,
``lmfit.minimize(..., method='nelder')``); any minimiser of the residual norm works
with ``SlabSizeModel.residual``.
"""

from __future__ import annotations

import numpy as np

from . import engine
from .tools.comparison import FitComparison, SlabBuilder, _weights

__all__ = ['fullfit_model', 'SlabSizeModel']


def _value(p):
    return float(getattr(p, 'value', p))


class SlabSizeModel:
    """fullfit_model with the experimental side resident on the device.

    fixed_* tuples are the reference's (fit_slabsize.py:142-145):
    fixed_slab_params = (input_filepath, a, b, c, alpha, beta, gamma),
    fixed_voxelgrid_params = (r_voxel_size, q_voxel_size, max_q, energy, fill_bkg, smooth),
    fixed_detectormaker_params = (num_pixels, angle_init_vals, angle_init_axs, psis,
    psi_weights_path, phis, phi_weights_path, thetas, theta_weights_path),
    fixed_exp_params = (rebin_map, rebin_mask, exp_qxy, exp_qz, pad_width, pad_range).
    """

    def __init__(self, fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params,
                 fixed_exp_params):
        self.slab = fixed_slab_params
        self.vox = fixed_voxelgrid_params
        (num_pixels, angle_init_vals, angle_init_axs, psis, psi_wp, phis, phi_wp, thetas,
         theta_wp) = fixed_detectormaker_params
        self.det_plan = engine.detector_plan(num_pixels, self.vox[2], angle_init_vals, angle_init_axs, psis,
                                             _weights(psi_wp, psis), phis, _weights(phi_wp, phis), thetas,
                                             _weights(theta_wp, thetas))
        self.exp = fixed_exp_params
        self.cmp = None
        self.builder = None
        self.f_table = None

    def image(self, x_size, y_size, z_size):
        """Device detector image [v, h] (mirrored) of the slab of the given size."""
        r_voxel_size, q_voxel_size, max_q, energy, fill_bkg, smooth = self.vox
        if self.builder is None:
            self.builder = SlabBuilder(*self.slab)
            # f of every element of the cell (an element the crop drops has count 0)
            self.f_table = engine.element_table(self.builder.uniq, energy)[1]
        x, y, z, el = self.builder.build(x_size, y_size, z_size)
        res = engine.run_voxelgrid((x, y, z), None, r_voxel_size, q_voxel_size, max_q, energy,
                                   fill_bkg=fill_bkg, smooth=int(smooth), f_table=self.f_table, elem_ids=el,
                                   want64=True)
        det, _ = engine.run_detector(res['iq64'], res['qx'], res['qy'], res['qz'], self.det_plan, True)
        return det

    def residual(self, x_size, y_size, z_size):
        """Residual vector (numpy) of one evaluation, and the fitted (scale, offset)."""
        det = self.image(x_size, y_size, z_size)
        if self.cmp is None:
            rebin_map, rebin_mask, exp_qxy, exp_qz, pad_width, pad_range = self.exp
            self.cmp = FitComparison(self.det_plan.h, self.det_plan.v, rebin_map, rebin_mask, exp_qxy, exp_qz,
                                     pad_width, pad_range)
        r, scale, offset = self.cmp.residual(det)
        return r.cpu().numpy(), (scale, offset)


_MODELS = {}


def fullfit_model(params, fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params, fixed_exp_params):
    """Residual of the simulated image of a slab (params 'x_size', 'y_size',
    'z_size': numbers or lmfit-style parameters with .value) against the rebinned
    experimental map (fit_slabsize.py:19-47)."""
    key = (id(fixed_exp_params[0]), repr(fixed_slab_params), repr(fixed_voxelgrid_params),
           repr([np.asarray(x).tobytes() if isinstance(x, np.ndarray) else x for x in fixed_detectormaker_params]))
    model = _MODELS.get(key)
    if model is None:
        if len(_MODELS) >= 4:
            _MODELS.pop(next(iter(_MODELS)))
        model = _MODELS[key] = SlabSizeModel(fixed_slab_params, fixed_voxelgrid_params,
                                             fixed_detectormaker_params, fixed_exp_params)
    r, _ = model.residual(_value(params['x_size']), _value(params['y_size']), _value(params['z_size']))
    return r
