"""Pins of the FP64 autograd reference of the differentiable variant
(oracle/cic_model.py), on the CPU.

* deposit='ngp' collapses the model to the reference's nearest-grid-point
  path: iq and the detector image must equal oracle.voxelgridmaker /
  oracle.detectormaker, which pins every index map (atom nodes, pixel -> bin,
  crop, f0, detector voxels, mirror, floor) the model shares with the
  cloud-in-cell variant;
* two atoms: the frame intensity has the closed form
  |fA + fB Wv(kz) Wu(ky)|^2 with W(k) = (1 - t) e^{-2 pi i k n0 / N} + t e^{-2 pi i k (n0 + 1) / N},
  and its derivative in B's fractional position is checked against the
  model's autograd;
* torch.autograd of the full model (also with fill_bkg and smoothing, including
  the bounding-box atoms that carry the avg_voxel_f term) against FP64 central
  differences.
"""

import numpy as np
import pytest
import torch

import oracle as O
from oracle.cic_model import CicModel
from giwaxsim_amd.synth import graphite_slab

R = 0.25
N = 128
QV = 2 * np.pi / (R * N) * (1 + 1e-9)
MAXQ = 2.5
DET = dict(num_pixels=24, max_q=MAXQ, psis=[0.0, 30.0, 60.0, 90.0], phis=[0.0], thetas=[0.0],
           angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)


def _slab():
    c, e = graphite_slab(size=(12.0, 12.0, 9.0), seed=11)
    return c, np.zeros(len(c), dtype=np.int64)


def test_ngp_mode_reproduces_oracle():
    coords, ids = _slab()
    f = np.array([6.0 + 0.0159, 0.0093])
    frames = (0, 7)
    model = CicModel(coords, ids, R, QV, MAXQ, DET, frames=frames, deposit='ngp')
    t = torch.as_tensor(coords)
    fr = torch.tensor([f[0]], dtype=torch.float64)
    fi = torch.tensor([f[1]], dtype=torch.float64)
    iq = model.iq(t, fr, fi).numpy().reshape((model.c_n,) * 3)
    plan = O.voxelgrid_plan(coords, R, QV, MAXQ)
    fvals = np.full(len(coords), f[0] + 1j * f[1])
    ref = O.voxelgridmaker(coords, fvals, R, QV, MAXQ, phis=plan['phis'][frames[0]:frames[1]])
    assert iq.shape == ref['iq'].shape
    assert np.allclose(iq, ref['iq'], rtol=1e-11, atol=1e-12 * ref['iq'].max())
    det = model.detector(torch.as_tensor(ref['iq'].ravel())).numpy()
    want, _, _ = O.detectormaker(ref['iq'], ref['qx'], ref['qy'], ref['qz'], DET['num_pixels'], MAXQ,
                                 DET['angle_init_vals'], DET['angle_init_axs'], np.array(DET['psis']),
                                 None, np.array([0.0]), None, np.array([0.0]), None, mirror=True)
    assert np.allclose(det, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("fill,smooth", [(True, 0), (False, 2), (True, 2)])
def test_ngp_mode_reproduces_oracle_fill_smooth(fill, smooth):
    coords, ids = _slab()
    f = np.array([6.0 + 0.0159, 0.0093])
    frames = (2, 6)
    model = CicModel(coords, ids, R, QV, MAXQ, DET, frames=frames, deposit='ngp', fill_bkg=fill,
                     smooth=smooth)
    fr = torch.tensor([f[0]], dtype=torch.float64)
    fi = torch.tensor([f[1]], dtype=torch.float64)
    iq = model.iq(torch.as_tensor(coords), fr, fi).numpy().reshape((model.c_n,) * 3)
    plan = O.voxelgrid_plan(coords, R, QV, MAXQ)
    fvals = np.full(len(coords), f[0] + 1j * f[1])
    ref = O.voxelgridmaker(coords, fvals, R, QV, MAXQ, fill_bkg=fill, smooth=smooth,
                           phis=plan['phis'][frames[0]:frames[1]])
    assert np.allclose(iq, ref['iq'], rtol=1e-10, atol=1e-12 * ref['iq'].max())


def _two_atoms(uB, vB):
    """Atom A at the origin (it is min y' and min z), B at grid coordinates
    (uB, vB) in the phi = 0 frame (y' = y)."""
    return np.array([[0.3, 0.0, 0.0], [0.7, uB * R, vB * R]])


def _closed_form(fA, fB, uB, vB):
    u0, v0 = np.floor(uB), np.floor(vB)
    tu, tv = uB - u0, vB - v0
    k = np.arange(N)
    e = lambda n: np.exp(-2j * np.pi * k * n / N)
    Wu = (1 - tu) * e(u0) + tu * e(u0 + 1)
    Wv = (1 - tv) * e(v0) + tv * e(v0 + 1)
    F = fA + fB * np.outer(Wv, Wu)
    # d/du of |F|^2 = 2 Re(conj(F) fB Wv dWu/du),  dWu/du = e(u0 + 1) - e(u0)
    dF = fB * np.outer(Wv, e(u0 + 1) - e(u0))
    return np.abs(F) ** 2, 2 * np.real(np.conj(F) * dF)


@pytest.mark.parametrize("uB,vB", [(3.3, 2.6), (17.0625, 40.75), (0.5, 0.25)])
def test_two_atom_closed_form(uB, vB):
    c = _two_atoms(uB, vB)
    model = CicModel(np.array([[0, 0, 0], [30.0, 30.0, 30.0]]), np.array([0, 1]), R, QV, MAXQ, DET,
                     frames=(0, 1))
    fA, fB = 6.1 + 0.2j, 7.3 - 0.4j
    t = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    fr = torch.tensor([fA.real, fB.real], dtype=torch.float64)
    fi = torch.tensor([fA.imag, fB.imag], dtype=torch.float64)
    model.elem = torch.tensor([0, 1])
    I = model.frame_intensity(t, fr, fi, 0.0, 1.0)
    want, dwant = _closed_form(fA, fB, uB, vB)
    assert np.allclose(I.detach().numpy().reshape(N, N), want, rtol=1e-11, atol=1e-9)
    # derivative of sum(I * M) for a fixed random weight M, along B's y (u = y / r)
    M = torch.as_tensor(np.random.default_rng(0).normal(size=N * N))
    (I * M).sum().backward()
    got = t.grad[1, 1].item()
    ref = float(np.sum(dwant.ravel() * M.numpy())) / R
    assert abs(got - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("fill,smooth", [(False, 0), (True, 2)])
def test_autograd_matches_central_differences(fill, smooth):
    coords, ids = _slab()
    model = CicModel(coords, ids, R, QV, MAXQ, DET, frames=(3, 6), fill_bkg=fill, smooth=smooth)
    rng = np.random.default_rng(5)
    G = torch.as_tensor(rng.normal(size=(DET['num_pixels'],) * 2))
    fr = torch.tensor([6.0159], dtype=torch.float64, requires_grad=True)
    fi = torch.tensor([0.0093], dtype=torch.float64, requires_grad=True)
    t = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
    loss = (model(t, fr, fi) * G).sum()
    loss.backward()
    h = 1e-4

    def L(c, a=6.0159, b=0.0093):
        with torch.no_grad():
            return float((model(torch.as_tensor(c), torch.tensor([a], dtype=torch.float64),
                                torch.tensor([b], dtype=torch.float64)) * G).sum())

    picks = list(rng.choice(len(coords) * 3, size=12, replace=False))
    if fill:  # the atoms that fix the bounding box carry the avg_voxel_f term
        picks += [3 * int(np.argmax(coords[:, 0])), 3 * int(np.argmin(coords[:, 1])) + 1,
                  3 * int(np.argmax(coords[:, 2])) + 2]
    gmax = t.grad.abs().max().item()
    for p in picks:
        i, k = divmod(int(p), 3)
        cp, cm = coords.copy(), coords.copy()
        cp[i, k] += h
        cm[i, k] -= h
        fd = (L(cp) - L(cm)) / (2 * h)
        assert abs(fd - t.grad[i, k].item()) <= 2e-5 * gmax, (i, k, fd, t.grad[i, k].item())
    fd = (L(coords, a=6.0159 + h) - L(coords, a=6.0159 - h)) / (2 * h)
    assert abs(fd - fr.grad.item()) <= 1e-7 * abs(fd)
    fd = (L(coords, b=0.0093 + h) - L(coords, b=0.0093 - h)) / (2 * h)
    assert abs(fd - fi.grad.item()) <= 1e-6 * abs(fr.grad.item())
