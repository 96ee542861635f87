"""GPU checks of the detector adjoint (dL/d iq of the orientation gather and finalize),
through the C ABI, on the geometry of a small differentiable slab (48^2 px, psi 0..90
step 10, a 0.15 deg psi tilt, mirror):

* the transpose-map path (gx_detector_adjoint_mapped) and the scatter path with
  copies of the accumulator (gx_detector_adjoint, n_rep = 3) give the same int64
  fixed-point sums bit for bit (both sum exact integer terms under the same scale),
  with uniform psi weights and with weights of which only some equal the first;
* the adjoint identity <G, image(iq + e) - image(iq)> = <dL/d iq, e> for a random
  cotangent G and a random perturbation e of a positive iq grid (the image is linear
  in iq where the floor is inactive), to 1e-9 relative;
* the transpose map, expanded run by run, against the oracle's gather indices.
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

from giwaxsim_amd import _lib
from giwaxsim_amd._lib import call
from giwaxsim_amd.adjoint import DifferentiableSlab
from giwaxsim_amd.engine import ptr, stream_handle
from giwaxsim_amd.synth import graphite_slab

pytestmark = pytest.mark.gpu


# psi weights of which some equal the first and some do not: the mapped adjoint takes
# its prefix table for the former and sums pixel by pixel for the latter
PSI_W = [0.15, 0.05, 0.15, 0.1, 0.15, 0.02, 0.08, 0.15, 0.1, 0.05]


def _model(psi_weights=None):
    coords, el = graphite_slab(size=(16.0, 16.0, 12.0), seed=21)
    N, R = 256, 0.25
    qv = 2 * np.pi / (R * N) * (1 + 1e-9)
    det = dict(num_pixels=48, max_q=2.5, psis=list(np.arange(0.0, 91.0, 10.0)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True,
               psi_weights=psi_weights)
    return DifferentiableSlab(el, coords, R, qv, 2.5, det, frames=(0, 4))


def _image(m):
    st = stream_handle()
    call("gx_detector_accumulate", ctypes.byref(m.detector_desc), st)
    call("gx_detector_finalize", ptr(m.det_raw), m.n_pix, m.n_pix, m._flags(), ptr(m.det_out), st)
    return m.det_out.clone()


@pytest.mark.parametrize("psi_weights", [None, PSI_W])
def test_mapped_and_scatter_adjoints_agree_and_transpose_the_image(psi_weights):
    print(f"[gpu-detector-adjoint] 48^2 px x 10 psi, mapped vs scatter (n_rep=3), adjoint identity, "
          f"psi weights {psi_weights}", flush=True)
    sys.stdout.flush()
    m = _model(psi_weights)
    assert m.mapped, "this geometry's map fits; the test compares both paths"
    dev = torch.device('cuda')
    gen = torch.Generator(device='cpu').manual_seed(3)
    nv = m.iq64.numel()
    iq = (0.5 + torch.rand(nv, generator=gen, dtype=torch.float64)).to(dev)
    m.iq64.copy_(iq)
    img0 = _image(m)
    G = torch.randn(m.n_pix ** 2, generator=gen, dtype=torch.float64).to(dev)
    st = stream_handle()
    # transpose-map path
    fx1 = torch.empty(nv, dtype=torch.int64, device=dev)
    sc1 = torch.empty(1, dtype=torch.float64, device=dev)
    call("gx_detector_adjoint_mapped", ctypes.byref(m.detector_desc), ptr(G), m._flags(), ptr(m.det_work),
         ptr(m.det_map[0]), ptr(m.det_map[1]), ptr(m.det_map[2]), m.det_map[3], ptr(fx1), ptr(sc1), st)
    # scatter path into three copies
    n_rep = 3
    fx2 = torch.empty(n_rep * nv, dtype=torch.int64, device=dev)
    sc2 = torch.empty(1, dtype=torch.float64, device=dev)
    call("gx_detector_adjoint", ctypes.byref(m.detector_desc), ptr(G), m._flags(), ptr(m.det_work), ptr(fx2),
         n_rep, ptr(sc2), st)
    torch.cuda.synchronize()
    assert torch.equal(sc1, sc2)
    # wrapping int64 sum of the copies (torch sums int64 exactly)
    assert torch.equal(fx2.view(n_rep, nv).sum(0), fx1)
    # adjoint identity
    giq = fx1.double() / sc1
    e = 1e-3 * torch.randn(nv, generator=gen, dtype=torch.float64).to(dev)
    m.iq64.copy_(iq + e)
    img1 = _image(m)
    lhs = float((G * (img1 - img0)).sum())
    rhs = float((giq * e).sum())
    print(f"  <G, dimage> = {lhs:.12e}, <dL/diq, e> = {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-9 * abs(rhs)


def test_map_runs_transpose_the_oracle_gather():
    """The transpose map (gx_detector_map_count / gx_detector_map_fill), expanded run by
    run, holds every (pixel, orientation) pair exactly once, at the voxel the oracle's
    detector gather reads for it; the runs are voxel-sorted, at most 64 pixels long and
    never cross a wave of 64 pixels."""
    from oracle.cic_model import _detector_maps
    print("[gpu-detector-adjoint] 48^2 px x 10 psi transpose map vs the oracle's gather", flush=True)
    sys.stdout.flush()
    m = _model()
    assert m.mapped
    run_voxel, runs, run_len, n_runs = m.det_map
    rv = run_voxel.cpu().numpy().astype(np.int64)
    ids = runs.cpu().numpy().view(np.uint32).astype(np.int64)
    ln = run_len.cpu().numpy().astype(np.int64)
    npx = m.n_pix ** 2
    n_orient = int(m.detector_desc.n_orient)
    pbits = int(np.ceil(np.log2(npx)))
    p0, o = ids & ((1 << pbits) - 1), ids >> pbits
    assert len(rv) == n_runs and np.all(np.diff(rv) >= 0)
    assert np.all((ln >= 1) & (ln <= 64)) and np.all(p0 // 64 == (p0 + ln - 1) // 64)
    assert int(ln.sum()) == npx * n_orient
    rep = np.repeat(np.arange(n_runs), ln)
    j = np.arange(int(ln.sum())) - np.repeat(np.cumsum(ln) - ln, ln)
    P, Ov, V = p0[rep] + j, o[rep], rv[rep]
    assert len(np.unique(Ov * npx + P)) == npx * n_orient
    got = np.full((n_orient, npx), -1, dtype=np.int64)
    got[Ov, P] = V
    ax = m.plan.grid.crop_axis
    one = np.ones(1)
    want, _ = _detector_maps(m.n_pix, 2.5, (0.15,), ('psi',), np.arange(0.0, 91.0, 10.0), np.ones(10) / 10,
                             np.zeros(1), one, np.zeros(1), one, ax, ax, ax)
    assert np.array_equal(got, want)
    print(f"  {n_runs} runs for {npx * n_orient} pairs ({npx * n_orient / n_runs:.2f} pixels a run)", flush=True)
