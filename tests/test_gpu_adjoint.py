"""GPU parity of the differentiable variant and its hand-written adjoint.

Checks, through the C ABI (giwaxsim_amd.adjoint):
* forward image vs the FP64 autograd reference model (oracle/cic_model.py) on
  the same inputs, plain and with fill_bkg and/or smoothing: max |diff| / max |image| <= 1e-12
  with FP64 transforms (1e-10 for the filled / smoothed grids, whose transforms are dominated by
  the k = 0 term), <= 2e-6 with FP32;
* gradients dL/d coords, dL/d f_re, dL/d f_im of L = <G, image> vs the
  reference's torch.autograd: max |diff| / max |grad| <= 1e-11 (FP64),
  <= 1e-4 (FP32);
* size-independent properties: rigid translations leave the image unchanged
  (sum of the coordinate gradient over atoms ~ 0), the image is homogeneous of
  degree 2 in f (Euler: sum f dL/df = 2 L on unclipped pixels), and two
  backward passes are bitwise identical;
* the backward that reuses the forward's last batch of row transforms
  (gx_adj_backward_stored) gives bitwise the gradients of the one that recomputes
  every batch, in one batch and in four;
* the differentiation workload (graphite 40 x 40 x 30 A, N = 512, 453 frames,
  256^2 detector, psi 0..90 step 2, target with c x 1.02, fitted scale and
  background, taken 5% away from the least-squares optimum so that dL/ds and
  dL/db are not zero): gradients vs FP64 central differences (h = 1e-4 A) of
  the GPU forward on 64 random coordinates, and of f, scale and background;
  and the same loss, image and gradients vs the FP64 autograd reference over all
  453 frames (FP64 and FP32 transforms, the bars above).
"""

import sys

import numpy as np
import pytest
import torch

from giwaxsim_amd import _lib
from giwaxsim_amd.adjoint import DifferentiableSlab, fit_scale_background, image_loss
from giwaxsim_amd.synth import graphite_slab
from oracle.cic_model import CicModel

pytestmark = pytest.mark.gpu

R = 0.25
MAXQ = 2.5


# more elements than the kernels reduce per element in registers (GX_ADJ_ELEM_REG):
# the per-atom dL/df path
MANY = ('C', 'N', 'O', 'S', 'F', 'H', 'Si')


def _small(precision, fill=False, smooth=0, many=False, N=256):
    coords, el = graphite_slab(size=(16.0, 16.0, 12.0), seed=21)
    # a second element so per-element gradients are exercised
    el = el.copy()
    el[::7] = 'N'
    if many:
        el = np.array([MANY[i % len(MANY)] for i in range(len(el))])
    qv = 2 * np.pi / (R * N) * (1 + 1e-9)
    det = dict(num_pixels=48, max_q=MAXQ, psis=list(np.arange(0.0, 91.0, 10.0)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    frames = (0, 10)
    model = DifferentiableSlab(el, coords, R, qv, MAXQ, det, frames=frames, precision=precision,
                               fill_bkg=fill, smooth=smooth)
    ids = np.array([model.elements.index(e) for e in el])
    ref = CicModel(coords, ids, R, qv, MAXQ, det, frames=frames, fill_bkg=fill, smooth=smooth)
    return coords, model, ref


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision,fill,smooth,tol_img,tol_grad,many,N",
                         [('fp64', False, 0, 1e-12, 1e-11, False, 256), ('fp32', False, 0, 2e-6, 1e-4, False, 256),
                          ('fp64', True, 0, 1e-10, 1e-11, False, 256), ('fp64', False, 3, 1e-10, 1e-11, False, 256),
                          ('fp64', True, 3, 1e-10, 1e-11, False, 256), ('fp32', True, 3, 2e-6, 1e-4, False, 256),
                          ('fp64', True, 3, 1e-10, 1e-11, True, 256), ('fp32', False, 0, 2e-6, 1e-4, True, 256),
                          # edges that are not powers of two: chirp-z band transforms and their adjoints
                          ('fp64', False, 0, 1e-10, 1e-10, False, 300), ('fp64', True, 3, 1e-10, 1e-10, False, 300),
                          ('fp32', True, 3, 2e-6, 1e-4, False, 300), ('fp64', True, 3, 1e-10, 1e-10, True, 200)])
def test_forward_and_gradients_match_autograd(precision, fill, smooth, tol_img, tol_grad, many, N):
    _say(f"[gpu-adjoint] small graphite N={N} frames 0..10 precision={precision} fill={fill} "
         f"smooth={smooth} elements={len(MANY) if many else 2}")
    coords, model, ref = _small(precision, fill, smooth, many, N)
    assert (model.n_elem > _lib.GX_ADJ_ELEM_REG) == many
    assert model.plan.N == N and (model.bs is None) == (N == 256)
    dev = torch.device('cuda')
    G = np.random.default_rng(1).normal(size=(48, 48))
    # HIP
    c = torch.tensor(coords, dtype=torch.float64, device=dev, requires_grad=True)
    fr = model.f_re0.clone().requires_grad_(True)
    fi = model.f_im0.clone().requires_grad_(True)
    img = model(c, fr, fi)
    (img * torch.as_tensor(G, device=dev)).sum().backward()
    # FP64 autograd reference
    ct = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
    frt = model.f_re0.cpu().clone().requires_grad_(True)
    fit = model.f_im0.cpu().clone().requires_grad_(True)
    want = ref(ct, frt, fit)
    (want * torch.as_tensor(G)).sum().backward()
    w = want.detach().numpy()
    e_img = np.max(np.abs(img.detach().cpu().numpy() - w)) / np.max(np.abs(w))
    gc, gw = c.grad.cpu().numpy(), ct.grad.numpy()
    e_c = np.max(np.abs(gc - gw)) / np.max(np.abs(gw))
    e_fr = np.max(np.abs(fr.grad.cpu().numpy() - frt.grad.numpy())) / np.max(np.abs(frt.grad.numpy()))
    e_fi = np.max(np.abs(fi.grad.cpu().numpy() - fit.grad.numpy())) / np.max(np.abs(fit.grad.numpy()))
    _say(f"  image {e_img:.2e}  d/dcoords {e_c:.2e}  d/df_re {e_fr:.2e}  d/df_im {e_fi:.2e}")
    assert e_img <= tol_img
    assert e_c <= tol_grad and e_fr <= tol_grad and e_fi <= tol_grad


@pytest.mark.parametrize("fill,smooth", [(False, 0), (True, 3)])
def test_invariances_and_determinism(fill, smooth):
    _say(f"[gpu-adjoint] translation invariance, Euler identity in f, determinism fill={fill} "
         f"smooth={smooth}")
    coords, model, _ = _small('fp64', fill, smooth)
    dev = torch.device('cuda')
    G = torch.as_tensor(np.random.default_rng(2).normal(size=(48, 48)), device=dev)
    grads = []
    for _ in range(2):
        c = torch.tensor(coords, dtype=torch.float64, device=dev, requires_grad=True)
        fr = model.f_re0.clone().requires_grad_(True)
        fi = model.f_im0.clone().requires_grad_(True)
        img = model(c, fr, fi)
        L = (img * G).sum()
        L.backward()
        grads.append((c.grad.clone(), fr.grad.clone(), fi.grad.clone(), img.detach().clone(), L.item()))
    assert all(torch.equal(a, b) for a, b in zip(grads[0][:4], grads[1][:4]))
    gc, gfr, gfi, img, _ = grads[0]
    # rigid translation: the deposit follows the atoms, the image does not change
    scale = gc.abs().sum(0)
    assert torch.all(gc.sum(0).abs() <= 1e-10 * scale), (gc.sum(0), scale)
    # Euler: image(lambda f) = lambda^2 image(f) on unclipped pixels (clipped ones are 1e-12)
    live = img != 1e-12
    lhs = float((model.f_re0 * gfr).sum() + (model.f_im0 * gfi).sum())
    rhs = 2.0 * float((img * G)[live].sum())
    assert abs(lhs - rhs) <= 1e-10 * float((img * G).abs().sum())


@pytest.mark.parametrize("precision,batch", [('fp64', None), ('fp32', 3)])
def test_stored_frames_backward_is_bitwise_recompute(precision, batch):
    _say(f"[gpu-adjoint] stored vs recomputed row pass in the backward, {precision}, batch={batch}")
    coords, el = graphite_slab(size=(16.0, 16.0, 12.0), seed=21)
    el = el.copy()
    el[::7] = 'N'
    N = 256
    qv = 2 * np.pi / (R * N) * (1 + 1e-9)
    det = dict(num_pixels=48, max_q=MAXQ, psis=list(np.arange(0.0, 91.0, 10.0)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    model = DifferentiableSlab(el, coords, R, qv, MAXQ, det, frames=(0, 10), precision=precision,
                               batch=batch, fill_bkg=True, smooth=3)
    dev = torch.device('cuda')
    c = torch.tensor(coords, dtype=torch.float64, device=dev)
    fr, adj = model.descriptors(c, model.f_re0, model.f_im0)
    G = torch.as_tensor(np.random.default_rng(5).normal(size=48 * 48), device=dev)
    from giwaxsim_amd.engine import stream_handle
    st = stream_handle()
    out = []
    for stored in (True, False):
        model.forward_enqueue(fr, adj, st)
        if not stored:
            model._inter_fr = None   # force the recompute of every batch
        model.backward_enqueue(fr, adj, G, st)
        torch.cuda.synchronize()
        model._check_status()
        out.append((model.g_coords.clone(), model.g_f.clone()))
    assert len(model.batches(fr)) == (1 if batch is None else 4)
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    assert float(out[0][0].abs().max()) > 0


def _workload():
    """Graphite 40 x 40 x 30 A (a = 2.46, c = 6.71, jitter 0.02 A), N = 512; the
    target is the same lattice with c scaled by 1.02."""
    coords, el = graphite_slab(size=(40.0, 40.0, 30.0), seed=4)
    target, tel = graphite_slab(size=(40.0, 40.0, 30.0), seed=4, c=6.71 * 1.02)
    N = 512
    qv = 2 * np.pi / (R * N) * (1 + 1e-9)
    det = dict(num_pixels=256, max_q=MAXQ, psis=list(np.arange(0.0, 91.0, 2.0)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    return coords, el, target, tel, qv, det


def test_workload_gradients_vs_central_differences():
    coords, el, target, tel, qv, det = _workload()
    _say(f"[gpu-adjoint] graphite 40x40x30 ({len(coords)} atoms) N=512 all frames, FD h=1e-4 on 64 coords")
    dev = torch.device('cuda')
    tmodel = DifferentiableSlab(tel, target, R, qv, MAXQ, det)
    with torch.no_grad():
        T = tmodel(torch.tensor(target, device=dev))
    model = DifferentiableSlab(el, coords, R, qv, MAXQ, det)
    with torch.no_grad():
        img0 = model(torch.tensor(coords, device=dev))
    s_fit, b_fit = fit_scale_background(img0, T)
    # gradients are taken away from the least-squares optimum (where dL/ds = dL/db = 0)
    s0, b0 = 1.05 * s_fit, b_fit + 0.05 * float(T.mean())
    c = torch.tensor(coords, dtype=torch.float64, device=dev, requires_grad=True)
    fr = model.f_re0.clone().requires_grad_(True)
    fi = model.f_im0.clone().requires_grad_(True)
    s = torch.tensor(s0, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(b0, dtype=torch.float64, device=dev, requires_grad=True)
    loss = image_loss(model(c, fr, fi), T, s, b)
    loss.backward()
    gc = c.grad.cpu().numpy()
    h = 1e-4

    def L(cc, a=None, bi=None, ss=s0, bb=b0):
        with torch.no_grad():
            img = model(torch.as_tensor(cc, device=dev), model.f_re0 if a is None else a,
                        model.f_im0 if bi is None else bi)
            return float(image_loss(img, T, ss, bb))

    rng = np.random.default_rng(64)
    picks = rng.choice(coords.size, size=64, replace=False)
    errs = []
    gmax = np.abs(gc).max()
    for p in picks:
        i, k = divmod(int(p), 3)
        cp, cm = coords.copy(), coords.copy()
        cp[i, k] += h
        cm[i, k] -= h
        fd = (L(cp) - L(cm)) / (2 * h)
        errs.append(abs(fd - gc[i, k]) / gmax)
    errs = np.array(errs)
    _say(f"  coords: max|g| {gmax:.3e}; |fd - g| / max|g|: median {np.median(errs):.2e} "
         f"95% {np.quantile(errs, 0.95):.2e} max {errs.max():.2e}")
    # f, scale, background
    ef = []
    for e in range(len(model.elements)):
        for which, grad in ((0, fr.grad), (1, fi.grad)):
            d = torch.zeros_like(model.f_re0)
            d[e] = h
            if which == 0:
                fd = (L(coords, a=model.f_re0 + d) - L(coords, a=model.f_re0 - d)) / (2 * h)
            else:
                fd = (L(coords, bi=model.f_im0 + d) - L(coords, bi=model.f_im0 - d)) / (2 * h)
            ef.append(abs(fd - grad[e].item()) / max(abs(fd), 1e-300))
    fds = (L(coords, ss=s0 + h * abs(s0)) - L(coords, ss=s0 - h * abs(s0))) / (2 * h * abs(s0))
    fdb = (L(coords, bb=b0 + h) - L(coords, bb=b0 - h)) / (2 * h)
    _say(f"  f: max rel err {max(ef):.2e}; scale grad {s.grad.item():.3e} (fd {fds:.3e}); "
         f"background grad {b.grad.item():.3e} (fd {fdb:.3e})")
    # the loss is piecewise quartic in one coordinate; a cell crossing inside
    # [x - h, x + h] in one of the 453 frames costs the difference its exactness
    assert np.median(errs) <= 1e-7
    assert np.quantile(errs, 0.95) <= 1e-5
    assert errs.max() <= 1e-4
    assert max(ef) <= 1e-7
    assert abs(fds - s.grad.item()) <= 1e-7 * abs(fds)
    assert abs(fdb - b.grad.item()) <= 1e-7 * abs(fdb)


_WORKLOAD_REF = {}


@pytest.mark.parametrize("precision,tol_img,tol_grad", [('fp64', 1e-10, 1e-10), ('fp32', 2e-6, 1e-4)])
def test_workload_matches_autograd(precision, tol_img, tol_grad):
    """The differentiation workload end to end (all 453 frames, the 256^2 detector, the
    image loss at a scale and background 5% off the least-squares optimum) against the
    FP64 autograd reference: image and dL/d coords, dL/d f_re, dL/d f_im, dL/ds, dL/db."""
    coords, el, target, tel, qv, det = _workload()
    dev = torch.device('cuda')
    tmodel = DifferentiableSlab(tel, target, R, qv, MAXQ, det, precision='fp64')
    with torch.no_grad():
        T = tmodel(torch.tensor(target, device=dev))
    model = DifferentiableSlab(el, coords, R, qv, MAXQ, det, precision=precision)
    assert model.plan.N == 512 and model.plan.n_frames == 453
    with torch.no_grad():
        img0 = model(torch.tensor(coords, device=dev))
    s_fit, b_fit = fit_scale_background(img0, T)
    s0, b0 = 1.05 * s_fit, b_fit + 0.05 * float(T.mean())
    _say(f"[gpu-adjoint] workload graphite 40x40x30 ({len(coords)} atoms) N=512 all frames, "
         f"{precision} vs FP64 autograd")
    Tc = T.cpu()
    if 'ref' not in _WORKLOAD_REF:
        ids = np.array([model.elements.index(e) for e in el])
        ref = CicModel(coords, ids, R, qv, MAXQ, det)
        ct = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
        frt = model.f_re0.cpu().clone().requires_grad_(True)
        fit = model.f_im0.cpu().clone().requires_grad_(True)
        st = torch.tensor(s0, dtype=torch.float64, requires_grad=True)
        bt = torch.tensor(b0, dtype=torch.float64, requires_grad=True)
        want = ref(ct, frt, fit)
        image_loss(want, Tc, st, bt).backward()
        _WORKLOAD_REF['ref'] = (s0, b0, want.detach().numpy(), ct.grad.numpy(), frt.grad.numpy(),
                                fit.grad.numpy(), st.grad.item(), bt.grad.item())
    # the scale and background of the first call for both (the fit of an FP32 image differs
    # by ~1e-7)
    s0, b0, w, gw, gfr, gfi, gs, gb = _WORKLOAD_REF['ref']
    _say(f"  scale {s0:.6e} background {b0:.6e}")
    c = torch.tensor(coords, dtype=torch.float64, device=dev, requires_grad=True)
    fr = model.f_re0.clone().requires_grad_(True)
    fi = model.f_im0.clone().requires_grad_(True)
    s = torch.tensor(s0, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.tensor(b0, dtype=torch.float64, device=dev, requires_grad=True)
    img = model(c, fr, fi)
    image_loss(img, T, s, b).backward()
    e_img = np.max(np.abs(img.detach().cpu().numpy() - w)) / np.max(np.abs(w))
    e_c = np.max(np.abs(c.grad.cpu().numpy() - gw)) / np.max(np.abs(gw))
    e_fr = np.max(np.abs(fr.grad.cpu().numpy() - gfr)) / np.max(np.abs(gfr))
    e_fi = np.max(np.abs(fi.grad.cpu().numpy() - gfi)) / np.max(np.abs(gfi))
    e_s = abs(s.grad.item() - gs) / abs(gs)
    e_b = abs(b.grad.item() - gb) / abs(gb)
    _say(f"  image {e_img:.2e}  d/dcoords {e_c:.2e}  d/df_re {e_fr:.2e}  d/df_im {e_fi:.2e}  "
         f"d/ds {e_s:.2e}  d/db {e_b:.2e}")
    assert e_img <= tol_img
    assert e_c <= tol_grad and e_fr <= tol_grad and e_fi <= tol_grad
    assert e_s <= tol_grad and e_b <= tol_grad


_LARGE_REF = {}


def _large_case(N):
    """A 40 x 40 x 30 A polymer slab (C/H/N/O/S/F) on the production grids: N = 1024
    (r 0.2 A, max_q 2.0, smooth 10), N = 2048 (r 0.15 A, max_q 3.0, smooth 20) and the
    reference's fit geometry N = 1048 (r 0.3 A, q voxel 0.02, smooth 5: chirp-z bands), all
    with fill_bkg; three phi-frames and a 64^2 detector over psi 0..90 in 15 degree steps.
    The FP64 autograd reference's image and gradients are computed once per N."""
    from giwaxsim_amd.synth import polymer_slab
    r, mq, smooth = {1024: (0.2, 2.0, 10), 2048: (0.15, 3.0, 20), 1048: (0.3, 2.0, 5)}[N]
    coords, el = polymer_slab(size=(40.0, 40.0, 30.0), seed=31)
    # N = 1048: the reference's example / fit geometry (r 0.3 A, q voxel 0.02), chirp-z bands
    qv = 0.02 if N == 1048 else 2 * np.pi / (r * N) * (1 + 1e-9)
    det = dict(num_pixels=64, max_q=mq, psis=list(np.arange(0.0, 91.0, 15.0)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    frames = (100, 103)
    G = np.random.default_rng(7).normal(size=(64, 64))
    if N not in _LARGE_REF:
        _say(f"[gpu-adjoint] FP64 autograd reference N={N}, {len(coords)} atoms, frames {frames}")
        uniq = sorted(set(el))
        ids = np.array([uniq.index(e) for e in el])
        from giwaxsim_amd.engine import element_table
        _, ftab, _, _ = element_table(el, 12700.0)
        ref = CicModel(coords, ids, r, qv, mq, det, frames=frames, fill_bkg=True, smooth=smooth)
        ct = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
        frt = torch.tensor(ftab.real.copy(), requires_grad=True)
        fit = torch.tensor(ftab.imag.copy(), requires_grad=True)
        want = ref(ct, frt, fit)
        (want * torch.as_tensor(G)).sum().backward()
        _LARGE_REF[N] = (want.detach().numpy(), ct.grad.numpy(), frt.grad.numpy(), fit.grad.numpy())
    return coords, el, r, qv, mq, det, frames, smooth, G, _LARGE_REF[N]


@pytest.mark.parametrize("N,precision,tol_img,tol_grad", [(1024, 'fp64', 1e-10, 1e-10),
                                                          (1024, 'fp32', 2e-6, 1e-4),
                                                          (2048, 'fp64', 1e-10, 1e-10),
                                                          (2048, 'fp32', 2e-6, 1e-4),
                                                          (1048, 'fp64', 1e-10, 1e-10),
                                                          (1048, 'fp32', 2e-6, 1e-4)])
def test_production_grids_match_autograd(N, precision, tol_img, tol_grad):
    coords, el, r, qv, mq, det, frames, smooth, G, (w, gw, gfr, gfi) = _large_case(N)
    _say(f"[gpu-adjoint] N={N} {precision} fill smooth {smooth}, frames {frames}")
    model = DifferentiableSlab(el, coords, r, qv, mq, det, frames=frames, precision=precision,
                               fill_bkg=True, smooth=smooth)
    assert model.plan.N == N
    if N == 1048:
        # the row band sized to the slab's projected width (188 rows) and the column band
        # to its 120 rows (the k = 0 column's fill constant beyond them in closed form):
        # 1024 each, not 2048
        assert model.bs['M'] == 1024 and model.bs['Mc'] == 1024
    dev = torch.device('cuda')
    c = torch.tensor(coords, dtype=torch.float64, device=dev, requires_grad=True)
    fr = model.f_re0.clone().requires_grad_(True)
    fi = model.f_im0.clone().requires_grad_(True)
    img = model(c, fr, fi)
    (img * torch.as_tensor(G, device=dev)).sum().backward()
    e_img = np.max(np.abs(img.detach().cpu().numpy() - w)) / np.max(np.abs(w))
    e_c = np.max(np.abs(c.grad.cpu().numpy() - gw)) / np.max(np.abs(gw))
    e_fr = np.max(np.abs(fr.grad.cpu().numpy() - gfr)) / np.max(np.abs(gfr))
    e_fi = np.max(np.abs(fi.grad.cpu().numpy() - gfi)) / np.max(np.abs(gfi))
    _say(f"  image {e_img:.2e}  d/dcoords {e_c:.2e}  d/df_re {e_fr:.2e}  d/df_im {e_fi:.2e}")
    assert e_img <= tol_img
    assert e_c <= tol_grad and e_fr <= tol_grad and e_fi <= tol_grad


def test_chirpz_row_band_support_is_checked():
    """The differentiable path's chirp-z row band is sized to the slab the model was built
    for (here 188 rows of projected width: length 1024 at N = 1048, rows below 738
    covered).  Called with atoms stretched to ~250 A along x, frames near phi = 90 degrees
    reach ~830 rows: the kernels skip those frames and the call raises with the reason,
    instead of returning an image with aliased rows."""
    from giwaxsim_amd.engine import qgrid
    from giwaxsim_amd.synth import polymer_slab
    coords, el = polymer_slab(size=(40.0, 40.0, 30.0), seed=31)
    det = dict(num_pixels=64, max_q=2.0, psis=[0.0, 45.0], phis=[0.0], thetas=[0.0],
               angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    mid = len(qgrid(0.02, 2.0).phis) // 2
    frames = (mid - 1, mid + 2)
    _say(f"[gpu-adjoint] chirp-z row band support check N=1048, frames {frames}")
    model = DifferentiableSlab(el, coords, 0.3, 0.02, 2.0, det, frames=frames, precision='fp32',
                               fill_bkg=True, smooth=5)
    assert model.plan.N == 1048 and model.bs['M'] == 1024
    dev = torch.device('cuda')
    fr, fi = model.f_re0.clone(), model.f_im0.clone()
    img = model(torch.tensor(coords, dtype=torch.float64, device=dev), fr, fi)
    assert torch.isfinite(img).all()
    wide = coords.copy()
    wide[:, 0] = wide[:, 0] * 6.0
    with pytest.raises(RuntimeError, match="chirp-z row band"):
        model(torch.tensor(wide, dtype=torch.float64, device=dev), fr, fi)


@pytest.mark.parametrize("smooth", [0, 5])
def test_chirpz_short_bands_without_fill_match_autograd(smooth):
    """N = 1048 (the fit geometry's edge) without fill: the row band is sized to the slab's
    projected width (1024); without smoothing the column band is short too (rows >= Z are
    empty, no fill tail), with smoothing only it keeps every row (2048).  FP64 transforms
    against the FP64 autograd reference on three frames, the bars of the production grids."""
    from giwaxsim_amd.engine import element_table
    from giwaxsim_amd.synth import polymer_slab
    coords, el = polymer_slab(size=(30.0, 30.0, 20.0), seed=41)
    r, qv, mq = 0.3, 0.02, 2.0
    det = dict(num_pixels=64, max_q=mq, psis=list(np.arange(0.0, 91.0, 15.0)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    frames = (100, 103)
    _say(f"[gpu-adjoint] N=1048 short bands, no fill, smooth {smooth}, {len(coords)} atoms, frames {frames}")
    uniq = sorted(set(el))
    ids = np.array([uniq.index(e) for e in el])
    _, ftab, _, _ = element_table(el, 12700.0)
    ref = CicModel(coords, ids, r, qv, mq, det, frames=frames, fill_bkg=False, smooth=smooth)
    ct = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
    frt = torch.tensor(ftab.real.copy(), requires_grad=True)
    fit = torch.tensor(ftab.imag.copy(), requires_grad=True)
    G = np.random.default_rng(9).normal(size=(64, 64))
    want = ref(ct, frt, fit)
    (want * torch.as_tensor(G)).sum().backward()
    model = DifferentiableSlab(el, coords, r, qv, mq, det, frames=frames, precision='fp64',
                               fill_bkg=False, smooth=smooth)
    assert model.plan.N == 1048 and model.bs['M'] == 1024
    assert model.bs['Mc'] == (2048 if smooth else 1024)
    dev = torch.device('cuda')
    c = torch.tensor(coords, dtype=torch.float64, device=dev, requires_grad=True)
    fr = model.f_re0.clone().requires_grad_(True)
    fi = model.f_im0.clone().requires_grad_(True)
    img = model(c, fr, fi)
    (img * torch.as_tensor(G, device=dev)).sum().backward()
    w = want.detach().numpy()
    e_img = np.max(np.abs(img.detach().cpu().numpy() - w)) / np.max(np.abs(w))
    e_c = np.max(np.abs(c.grad.cpu().numpy() - ct.grad.numpy())) / np.max(np.abs(ct.grad.numpy()))
    e_fr = np.max(np.abs(fr.grad.cpu().numpy() - frt.grad.numpy())) / np.max(np.abs(frt.grad.numpy()))
    e_fi = np.max(np.abs(fi.grad.cpu().numpy() - fit.grad.numpy())) / np.max(np.abs(fit.grad.numpy()))
    _say(f"  image {e_img:.2e}  d/dcoords {e_c:.2e}  d/df_re {e_fr:.2e}  d/df_im {e_fi:.2e}")
    assert e_img <= 1e-10
    assert e_c <= 1e-10 and e_fr <= 1e-10 and e_fi <= 1e-10
