"""Slab-size fit of a simulated GIWAXS image to an experimental one, mirroring the
reference's ``fit_slabsize.py`` (fullfit_model 19-47, main 63-233).

    python -m giwaxsim_amd.fit_slabsize --config fit_config.txt

``fullfit_model`` keeps the reference's signature and returns the residual vector
(numpy).  ``SlabSizeModel`` is the same model with its experimental side planned once
(``tools.comparison.FitComparison``) and its unit cell uploaded once
(``tools.comparison.SlabBuilder``): each evaluation tiles and crops the slab, runs the
voxel grid and the detector image, and trims, shift_peaks, fits scale and offset and
forms the masked residual, all on the device; only the residual vector comes back.

``main`` reads the reference's config keys, prepares the experimental map (image,
mask, mirror, trim, rebin) and fits x/y/z sizes by Nelder-Mead on the sum of squared
residuals, as the reference's ``lmfit.minimize(..., method='nelder')`` does: with
lmfit installed it calls lmfit itself; otherwise ``minimize_nelder`` restates lmfit's
bounded-parameter transform (MINUIT's arcsine mapping) around scipy's Nelder-Mead.
It writes the fit report, the convergence and parameter plots, the maps of the best
fit (evaluate_fit) and the config, under ``save_folder``.
"""

from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import engine
from .tools.comparison import FitComparison, SlabBuilder, _weights

__all__ = ['fullfit_model', 'SlabSizeModel', 'model_for', 'FitLogger', 'minimize_nelder', 'main']


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

    def _comparison(self):
        if self.cmp is None:
            rebin_map, rebin_mask, exp_qxy, exp_qz, pad_width, pad_range = self.exp
            self.cmp = FitComparison(self.det_plan.h, self.det_plan.v, rebin_map, rebin_mask, exp_qxy, exp_qz,
                                     pad_width, pad_range)
        return self.cmp

    def residual(self, x_size, y_size, z_size):
        """Residual vector (numpy) of one evaluation, and the fitted (scale, offset)."""
        r, scale, offset = self._comparison().residual(self.image(x_size, y_size, z_size))
        return r.cpu().numpy(), (scale, offset)

    def padded(self, x_size, y_size, z_size):
        """(trimmed and shift_peak-ed device image, scale, offset) of one evaluation."""
        from .tools.comparison import _scale_offset_dev
        cmp = self._comparison()
        sim = cmp.padded(self.image(x_size, y_size, z_size))
        scale, offset = _scale_offset_dev(sim, cmp.ref, cmp.mask)
        return sim, scale, offset


_MODELS = {}


def model_for(fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params, fixed_exp_params):
    """The SlabSizeModel of these fixed parameters (kept for the next call)."""
    key = (id(fixed_exp_params[0]), repr(fixed_slab_params), repr(fixed_voxelgrid_params),
           repr([np.asarray(x).tobytes() if isinstance(x, np.ndarray) else x for x in fixed_detectormaker_params]))
    model = _MODELS.get(key)
    if model is None:
        if len(_MODELS) >= 4:
            _MODELS.pop(next(iter(_MODELS)))
        model = _MODELS[key] = SlabSizeModel(fixed_slab_params, fixed_voxelgrid_params,
                                             fixed_detectormaker_params, fixed_exp_params)
    return model


def fullfit_model(params, fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params, fixed_exp_params):
    """Residual of the simulated image of a slab (params 'x_size', 'y_size',
    'z_size': numbers or lmfit-style parameters with .value) against the rebinned
    experimental map (fit_slabsize.py:19-47)."""
    model = model_for(fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params, fixed_exp_params)
    r, _ = model.residual(_value(params['x_size']), _value(params['y_size']), _value(params['z_size']))
    return r


class FitLogger:
    """Per-evaluation log of a fit (fit_slabsize.py:49-61): evaluation number,
    parameter values, residual norm."""

    def __init__(self):
        self.iterations = []
        self.param_logs = []
        self.residual_norms = []

    def callback(self, params, iter, resid, *args, **kwargs):
        self.iterations.append(iter)
        self.param_logs.append({name: _value(p) for name, p in params.items()})
        self.residual_norms.append(np.linalg.norm(resid))


class _Param:
    """A bounded fit parameter; like an lmfit Parameter, a value outside [min, max]
    is clipped to the bound (vary or not)."""

    def __init__(self, value, lo, hi, vary):
        self.min, self.max, self.vary = float(lo), float(hi), bool(vary)
        self.init_value = float(value)
        self.value = float(np.clip(self.init_value, self.min, self.max))

    # lmfit's transform of a parameter bounded on both sides (MINUIT's arcsine mapping):
    # internal = arcsin(2 (v - min) / (max - min) - 1), v = min + (sin(internal) + 1) (max - min) / 2
    def to_internal(self):
        return float(np.arcsin(2 * (self.value - self.min) / (self.max - self.min) - 1))

    def from_internal(self, u):
        return self.min + (np.sin(u) + 1) * (self.max - self.min) / 2


class FitResult:
    """What the fit script reads from a minimiser's result: best parameters, the
    number of evaluations, chi-square and reduced chi-square."""

    def __init__(self, params, nfev, chisqr, ndata, nvarys, aborted, message):
        self.params, self.nfev, self.chisqr = params, nfev, chisqr
        self.ndata, self.nvarys, self.nfree = ndata, nvarys, max(1, ndata - nvarys)
        self.redchi = chisqr / self.nfree
        self.aborted, self.message = aborted, message

    def report(self):
        lines = ['[[Fit Statistics]]', '    # fitting method   = Nelder-Mead',
                 f'    # function evals   = {self.nfev}', f'    # data points      = {self.ndata}',
                 f'    # variables        = {self.nvarys}', f'    chi-square         = {self.chisqr:.8g}',
                 f'    reduced chi-square = {self.redchi:.8g}', f'    message            = {self.message}',
                 '[[Variables]]']
        for name, p in self.params.items():
            tail = f'(init = {p.init_value:g})' if p.vary else '(fixed)'
            lines.append(f'    {name}: {p.value:.8g} {tail}  [{p.min:g}, {p.max:g}]')
        return '\n'.join(lines) + '\n'


def minimize_nelder(residual, params, max_nfev, iter_cb=None):
    """Nelder-Mead on sum(residual(values)^2) over the varying parameters in lmfit's
    bounded internal coordinates (scipy.optimize.minimize(method='Nelder-Mead'), its
    default initial simplex), stopping after ``max_nfev`` evaluations.  ``params``:
    name -> _Param; ``residual(dict of values)`` -> residual vector; ``iter_cb(params,
    nfev, resid)`` after every evaluation.  Returns a FitResult with the best
    parameters evaluated."""
    from scipy.optimize import minimize
    names = [k for k, p in params.items() if p.vary]
    best = {'chisqr': np.inf, 'values': {k: p.value for k, p in params.items()}, 'ndata': 0}
    nfev = [0]

    class _Stop(Exception):
        pass

    def values_of(u):
        vals = {k: p.value for k, p in params.items()}
        for k, ui in zip(names, u):
            vals[k] = params[k].from_internal(ui)
        return vals

    def objective(u):
        if nfev[0] >= max_nfev:
            raise _Stop()
        vals = values_of(u)
        r = np.asarray(residual(vals), dtype=np.float64)
        nfev[0] += 1
        chisqr = float(np.sum(r * r))
        if iter_cb is not None:
            shown = {k: _Param(vals[k], p.min, p.max, p.vary) for k, p in params.items()}
            iter_cb(shown, nfev[0], r)
        if chisqr < best['chisqr']:
            best.update(chisqr=chisqr, values=vals, ndata=r.size)
        return chisqr

    aborted, message = False, 'Optimization terminated successfully.'
    if names:
        u0 = np.array([params[k].to_internal() for k in names])
        try:
            # the evaluation cap is enforced in objective (scipy's own limit set beyond it)
            res = minimize(objective, u0, method='Nelder-Mead', options={'maxfev': max_nfev + 1,
                                                                          'maxiter': 2 * max_nfev + 2})
            message = str(res.message)
        except _Stop:
            aborted, message = True, f'Fit aborted: number of function evaluations > {max_nfev}'
    else:
        objective(np.zeros(0))
    out = {k: _Param(best['values'][k], p.min, p.max, p.vary) for k, p in params.items()}
    for k, p in params.items():
        out[k].init_value = p.init_value
    return FitResult(out, nfev[0], best['chisqr'], best['ndata'], len(names), aborted, message)


def _fit(params_init, fixed, num_evals, logger):
    """lmfit.minimize(fullfit_model, ..., method='nelder') when lmfit is installed,
    else minimize_nelder on the same residual."""
    try:
        import lmfit
    except ImportError:
        lmfit = None
    if lmfit is not None:
        params = lmfit.Parameters()
        for name, (value, lo, hi, vary) in params_init.items():
            params.add(name, value=value, min=lo, max=hi, vary=vary)
        result = lmfit.minimize(fullfit_model, params, args=fixed, method='nelder', max_nfev=num_evals,
                                iter_cb=logger.callback)
        return result, lmfit.fit_report(result)
    params = {name: _Param(*v) for name, v in params_init.items()}
    result = minimize_nelder(lambda vals: fullfit_model(vals, *fixed), params, num_evals, logger.callback)
    return result, result.report()


def _plot_trace(savepath, x, y, label, ylabel, title):
    from .tools.comparison import _pyplot
    _, plt = _pyplot()
    plt.figure()
    plt.plot(x, y, label=label)
    plt.xlabel('Iteration')
    plt.ylabel(ylabel)
    plt.title(title)
    plt.legend()
    plt.savefig(savepath, dpi=300)
    plt.close()


def main(config):
    """fit_slabsize.py:63-233: read the config, build the experimental map, fit the
    slab sizes (num_evals > 0) or evaluate the initial sizes, save the results."""
    from .tools.comparison import (evaluate_fit, load_exp_image, mirror_qmap_positive_qxy_only, plot_fit,
                                   plot_linecuts, rebin_qmap_and_mask, trim_sim_data)
    from .tools.utilities import save_config_to_txt, str_to_bool
    input_filepath = config.get('input_filepath')
    x_size_init = float(config.get('x_size_init', 150))
    y_size_init = float(config.get('y_size_init', 150))
    z_size_init = float(config.get('z_size_init', 150))
    fit_x = str_to_bool(config.get('fit_x', 'False'))
    fit_y = str_to_bool(config.get('fit_y', 'False'))
    fit_z = str_to_bool(config.get('fit_z', 'False'))
    a, b, c = float(config.get('a')), float(config.get('b')), float(config.get('c'))
    alpha, beta, gamma = float(config.get('alpha')), float(config.get('beta')), float(config.get('gamma'))
    r_voxel_size = float(config.get('r_voxel_size', 0.3))
    q_voxel_size = float(config.get('q_voxel_size', 0.04))
    max_q = float(config.get('max_q', 2.5))
    energy = float(config.get('energy', 10000))
    fill_bkg = str_to_bool(config.get('fill_bkg', 'False'))
    smooth = int(config.get('smooth', 0))
    num_pixels = int(config.get('num_pixels', max_q / q_voxel_size))
    angle_init_vals = tuple(float(config.get(f'angle_init_val{k}', 0)) for k in (1, 2, 3))
    angle_init_axs = tuple(config.get(f'angle_init_ax{k}', 'None') for k in (1, 2, 3))
    psis = np.linspace(float(config.get('psi_start')), float(config.get('psi_end')), int(config.get('psi_num')))
    phis = np.linspace(float(config.get('phi_start')), float(config.get('phi_end')), int(config.get('phi_num')))
    thetas = np.linspace(float(config.get('theta_start')), float(config.get('theta_end')),
                         int(config.get('theta_num')))
    psi_wp = config.get('psi_weights_path', None)
    phi_wp = config.get('phi_weights_path', None)
    theta_wp = config.get('theta_weights_path', None)
    mask_path, img_path = config.get('mask_path'), config.get('img_path')
    qxy_path, qz_path = config.get('qxy_path'), config.get('qz_path')
    pad_width = float(config.get('pad_width'))
    pad_range = (float(config.get('pad_range_min')), float(config.get('pad_range_max')))
    save_folder = config.get('save_folder')
    os.makedirs(save_folder, exist_ok=True)
    num_evals = int(config.get('num_evals'))

    fixed_slab_params = (input_filepath, a, b, c, alpha, beta, gamma)
    fixed_voxelgrid_params = (r_voxel_size, q_voxel_size, max_q, energy, fill_bkg, smooth)
    fixed_detectormaker_params = (num_pixels, angle_init_vals, angle_init_axs, psis, psi_wp, phis, phi_wp, thetas,
                                  theta_wp)
    # the detector axes of the (mirrored) simulated image depend on the geometry only
    plan = engine.detector_plan(num_pixels, max_q, angle_init_vals, angle_init_axs, psis, _weights(psi_wp, psis),
                                phis, _weights(phi_wp, phis), thetas, _weights(theta_wp, thetas))
    exp_img = load_exp_image(img_path)
    exp_qxy = np.loadtxt(qxy_path)
    exp_qz = np.loadtxt(qz_path)
    mask = np.load(mask_path, allow_pickle=False)
    exp_img, exp_qxy, exp_qz = mirror_qmap_positive_qxy_only(exp_img, exp_qxy, exp_qz)
    empty = np.zeros((len(plan.v), len(plan.h)))
    _, det_h_trim, det_v_trim = trim_sim_data(empty, plan.h, plan.v, exp_qxy, exp_qz)
    rebin_map, rebin_mask = rebin_qmap_and_mask(det_h_trim, det_v_trim, exp_img, exp_qxy, exp_qz, mask)
    rebin_map[rebin_mask == 1] = 0
    fixed_exp_params = (rebin_map, rebin_mask, exp_qxy, exp_qz, pad_width, pad_range)
    fixed = (fixed_slab_params, fixed_voxelgrid_params, fixed_detectormaker_params, fixed_exp_params)

    params_init = {'x_size': (x_size_init, 25, x_size_init * 3, fit_x),
                   'y_size': (y_size_init, 25, y_size_init * 3, fit_y),
                   'z_size': (z_size_init, 25, z_size_init * 3, fit_z)}
    result = logger = None
    if num_evals > 0:
        logger = FitLogger()
        result, report = _fit(params_init, fixed, num_evals, logger)
        best_params = tuple(_value(result.params[k]) for k in ('x_size', 'y_size', 'z_size'))
        with open(f'{save_folder}/fit_result.txt', 'w') as fh:
            fh.write(report)
        _plot_trace(f'{save_folder}/fit_residuals.png', logger.iterations, logger.residual_norms,
                    'Residual Norm', 'Residual Norm', 'Convergence of Residual Norm')
        for name in logger.param_logs[0]:
            _plot_trace(f'{save_folder}/fit_{name}_evolution.png', logger.iterations,
                        [p[name] for p in logger.param_logs], name, f'{name} Value', f'Evolution of {name}')
        suptitle = f'Reduced Chi2 = {result.redchi:0.0f}'
    else:
        best_params = (x_size_init, y_size_init, z_size_init)
        suptitle = 'No fit done, direct generation'

    rebin_map, sim_comp_map, diff_map = evaluate_fit(best_params, *fixed)
    np.save(f'{save_folder}/det_h.npy', det_h_trim)
    np.save(f'{save_folder}/det_v.npy', det_v_trim)
    np.save(f'{save_folder}/rebin_map.npy', rebin_map)
    np.save(f'{save_folder}/sim_comp_map.npy', sim_comp_map)
    plot_fit(f'{save_folder}/fit_result.png', rebin_map, sim_comp_map, diff_map, det_h_trim, det_v_trim, max_q,
             suptitle)
    plot_linecuts(f'{save_folder}/fit_result_linecut.png', rebin_map, sim_comp_map, det_h_trim, det_v_trim, suptitle)
    save_config_to_txt(config, f'{save_folder}/fit_config.txt')
    return {'best_params': best_params, 'result': result, 'logger': logger, 'rebin_map': rebin_map,
            'sim_comp_map': sim_comp_map, 'diff_map': diff_map, 'det_h': det_h_trim, 'det_v': det_v_trim}


def cli(argv=None):
    from .tools.utilities import parse_config_file
    start = time.time()
    parser = argparse.ArgumentParser(description="Process a configuration file.")
    parser.add_argument('--config', type=str, required=True, help='Path to the configuration file')
    args = parser.parse_args(argv)
    main(parse_config_file(args.config))
    print(f'\nTotal Time: {str(time.time() - start)}')


if __name__ == "__main__":
    cli()
