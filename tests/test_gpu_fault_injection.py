"""GX_FAULT_INJECT, the library's fault-injection hook for tests of callers' error
handling: an injected launch failure (all checks, or only those of one kernel) or HIP
runtime failure reaches the Python caller as GxError naming the check, and once the
variable is cleared the same calls succeed again with the same results (the injection
leaves no state behind)."""

import sys

import numpy as np
import pytest
import torch

from giwaxsim_amd import engine
from giwaxsim_amd._lib import GxError
from giwaxsim_amd.adjoint import DifferentiableSlab, fit_scale_background
from giwaxsim_amd.synth import graphite_slab

pytestmark = pytest.mark.gpu

R = 0.25


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _forward(coords, el):
    qv = 2 * np.pi / (R * 256) * (1 + 1e-9)
    ids, ftab, _, _ = engine.element_table(el, 12700.0)
    res = engine.run_voxelgrid(coords, el, R, qv, 2.5, 12700.0, fill_bkg=True, smooth=3, f_table=ftab, elem_ids=ids)
    return res['iq64'].cpu().numpy().copy()


def test_injected_failures_reach_the_caller(monkeypatch):
    _say("[gpu-fault-injection] graphite 16x16x12 N=256: launch, launch:<kernel>, hip:<call>")
    coords, el = graphite_slab(size=(16.0, 16.0, 12.0), seed=21)
    want = _forward(coords, el)
    monkeypatch.setenv("GX_FAULT_INJECT", "launch")
    with pytest.raises(GxError, match="injected launch failure"):
        _forward(coords, el)
    monkeypatch.delenv("GX_FAULT_INJECT")
    assert np.array_equal(_forward(coords, el), want)

    qv = 2 * np.pi / (R * 256) * (1 + 1e-9)
    det = dict(num_pixels=48, max_q=2.5, psis=list(np.arange(0.0, 91.0, 10.0)), phis=[0.0], thetas=[0.0],
               angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    m = DifferentiableSlab(el, coords, R, qv, 2.5, det, frames=(0, 6), fill_bkg=True, smooth=3)

    def grads():
        c = torch.tensor(coords, device='cuda', requires_grad=True)
        m(c).sum().backward()
        return c.grad.cpu().numpy().copy()

    g0 = grads()
    # only the column adjoint's launch check fails: the forward runs, the backward raises
    monkeypatch.setenv("GX_FAULT_INJECT", "launch:k_adj_colpass")
    c = torch.tensor(coords, device='cuda', requires_grad=True)
    img = m(c)
    with pytest.raises(GxError, match="k_adj_colpass: injected launch failure"):
        img.sum().backward()
    monkeypatch.delenv("GX_FAULT_INJECT")
    assert np.array_equal(grads(), g0)

    # a checked HIP runtime call (the synchronisation of the scale/offset moments)
    img = img.detach()
    monkeypatch.setenv("GX_FAULT_INJECT", "hip:hipStreamSynchronize")
    with pytest.raises(GxError, match="hipStreamSynchronize.*injected"):
        fit_scale_background(img, img * 2.0 + 1.0)
    monkeypatch.delenv("GX_FAULT_INJECT")
    s, b = fit_scale_background(img, img * 2.0 + 1.0)
    assert abs(s - 2.0) < 1e-9 and abs(b - 1.0) < 1e-9
