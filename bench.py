"""Benchmark of the phi-frame forward path (project + FFT + 3D bin) on one GPU.

Workload (the production-size synthetic config): a semicrystalline polymer
slab of 100 x 100 x 100 A (~110k C/H/N/O/S/F atoms), r_voxel_size 0.2 A, real
grid 1024 x 1024 (q_voxel = 2 pi / (0.2 * 1024)), energy 12700 eV, fill_bkg,
smooth 10 voxels (2 A), max_q 2.0 -> 581 phi-frames into a 185^3 bin grid
cropped to 131^3.  One step = every phi-frame of the slab, starting from
atoms resident in HBM, through the cropped FP64 I(q) on the device.  The
detector stage (512 x 512 px, 0.15 deg psi tilt, psi 0..180 in 0.5 deg steps,
361 orientations) is timed separately and reported as images/s.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the path runs on one GPU.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-frames", type=int, default=581,
                    help="phi-frames of the same slab timed on the host (CPU baseline sample)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="repeat the CPU baseline sample until this much CPU time has been measured")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-iters", type=int, default=20)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the ensemble (64 slabs) and large-slab (N=2048) figures")
    ap.add_argument("--only", choices=["adjoint", "adjoint_fp64", "adjoint_polymer", "adjoint_large",
                                       "ensemble", "large_slab", "compare", "slab_build", "fit_eval"],
                    help="run one of the extra workloads alone and print its JSON (for profiling)")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="one step, then one 256 MiB device copy (counter calibration) and one "
                         "launch of each frame kernel; for rocprofv3 --pmc passes")
    return ap.parse_args()


def measured_traffic_per_frame():
    """Calibrated HBM bytes per frame of the frame kernels, measured with two
    rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) on this workload and kept in
    profiles/frames_polymer100_N1024_counters.md."""
    path = os.path.join(ROOT, "profiles", "frames_polymer100_N1024_counters.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_frame:"):
                return {k: float(v) for k, v in (t.split("=") for t in line.split(":", 1)[1].split())}
    except OSError:
        pass
    return {}


def workload():
    from giwaxsim_amd.synth import polymer_slab
    coords, elements = polymer_slab(size=(100.0, 100.0, 100.0), seed=0)
    r = 0.2
    return dict(coords=coords, elements=elements, r=r, qv=2 * np.pi / (r * 1024) * (1 + 1e-9),
                max_q=2.0, energy=12700.0, fill=True, smooth=10)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        sys.exit(2)
    import torch
    from giwaxsim_amd import engine, _lib
    from giwaxsim_amd._lib import call

    if args.only:
        if args.only == "adjoint_polymer":
            print(json.dumps(adjoint_polymer_config(engine, torch, reps=args.steps)))
        elif args.only == "adjoint_large":
            print(json.dumps(adjoint_large_config(engine, torch, n_slabs=args.steps)))
        elif args.only.startswith("adjoint"):
            prec = "fp64" if args.only == "adjoint_fp64" else "fp32"
            print(json.dumps(adjoint_config(engine, torch, reps=args.steps, precision=prec)))
        else:
            print(json.dumps(extra_configs(args, engine, torch, which=args.only)[args.only]))
        return

    W = workload()
    dev = engine.device()
    ids, ftab, counts, _ = engine.element_table(W['elements'], W['energy'])
    coords_d = torch.as_tensor(W['coords']).to(dev)
    ids_d = torch.as_tensor(ids).to(dev)

    def step():
        return engine.run_voxelgrid(coords_d, None, W['r'], W['qv'], W['max_q'], W['energy'],
                                    fill_bkg=W['fill'], smooth=W['smooth'], f_table=ftab,
                                    elem_ids=ids_d, want64=True)

    if args.profile_kernels:
        res = step()
        torch.cuda.synchronize()
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)                      # 256 MiB read + 256 MiB write
        runner = res['runner']
        nb = runner.batch
        for kname in ("gx_frames_stats", "gx_frames_columns", "gx_frames_rowpass", "gx_frames_colpass"):
            call(kname, ctypes.byref(runner.desc), 0, nb, engine.stream_handle())
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernels": True, "frames_per_launch": nb, "copy_bytes": 2 * 4 * src.numel()}))
        return

    for _ in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    plan = res['plan']
    F = plan.n_frames
    N = plan.N
    n_atoms = len(W['coords'])

    # ---- dominant kernel, timed live with HIP events on the stream it runs on
    runner = res['runner']
    st = torch.cuda.current_stream()
    nb = runner.batch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    call("gx_frames_stats", ctypes.byref(runner.desc), 0, nb, engine.stream_handle())
    call("gx_frames_columns", ctypes.byref(runner.desc), 0, nb, engine.stream_handle())
    times = {}
    for kname in ("gx_frames_rowpass", "gx_frames_colpass"):
        call(kname, ctypes.byref(runner.desc), 0, nb, engine.stream_handle())
        ev[0].record(st)
        for _ in range(args.kernel_iters):
            call(kname, ctypes.byref(runner.desc), 0, nb, engine.stream_handle())
        ev[1].record(st)
        torch.cuda.synchronize()
        times[kname] = ev[0].elapsed_time(ev[1]) / args.kernel_iters / 1e3  # s per launch
    runner.bins.clear()
    dom = max(times, key=times.get)
    # algorithmic bytes per frame (fixed contract): 24 n_atoms + 16 N^2 + 16 P.
    P_mean = float(res['bins'].cnt.sum().item()) / F
    B_frame = 24 * n_atoms + 16 * N * N + 16 * P_mean
    # the row pass is charged the atom reads and the frame write (24 n_a + 8 N^2),
    # the column pass the frame read and the bin read-modify-writes (8 N^2 + 16 P)
    alg = {"gx_frames_rowpass": nb * (24 * n_atoms + 8 * N * N),
           "gx_frames_colpass": nb * (8 * N * N + 16 * P_mean)}
    achieved = alg[dom] / times[dom] / 1e9
    traffic = measured_traffic_per_frame()

    # ---- detector images/s
    from giwaxsim_amd.tools.comparison import _weights
    psis = np.linspace(0, 180, 361)
    one = np.array([0.0])
    dplan = engine.detector_plan(512, W['max_q'], (0.15, 0, 0), ('psi', 'None', 'None'), psis,
                                 _weights(None, psis), one, _weights(None, one), one,
                                 _weights(None, one))
    iq = res['iq64']
    engine.run_detector(iq, res['qx'], res['qy'], res['qz'], dplan, True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nd = max(3, args.steps // 4)
    for _ in range(nd):
        engine.run_detector(iq, res['qx'], res['qy'], res['qz'], dplan, True)
    torch.cuda.synchronize()
    det_s = (time.perf_counter() - t1) / nd

    # ---- other synthetic configurations, reported beside the headline figure
    extra = {}
    if not args.no_extra:
        extra = extra_configs(args, engine, torch)

    # ---- CPU baseline: the oracle in reference-faithful threaded mode on a bounded sample
    cpu = None
    if not args.no_cpu:
        import oracle as O
        phis = plan.grid.phis[:args.cpu_frames]
        workers = O.default_threads()
        # whole passes over the sample until at least --cpu-seconds of CPU work (at most 8)
        passes, tcpu = 0, 0.0
        while passes < 8 and (passes == 0 or tcpu < args.cpu_seconds):
            tc = time.perf_counter()
            O.voxelgridmaker(W['coords'], ftab[ids], W['r'], W['qv'], W['max_q'], fill_bkg=W['fill'],
                             smooth=W['smooth'], phis=phis, threads=None)
            tcpu += time.perf_counter() - tc
            passes += 1
        cpu = {"value": passes * len(phis) / tcpu, "unit": "phi-frames/s", "cores": workers, "kind": "port",
               "sample": f"{passes} pass(es) over {len(phis)} of {F} phi-frames of the same slab, FP64 numpy "
                         f"oracle with ThreadPoolExecutor() default workers ({workers}) on {os.cpu_count()} "
                         f"host CPUs; {tcpu:.1f} s"}

    out = {
        "metric": "phi-frames/s (project+FFT+3D bin)",
        "value": F / dt,
        "unit": "phi-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 FFT/|F|^2, f64 indices, int64 fixed-point bin sums",
        "data": "synthetic (seeded polymer slab, 100^3 A)",
        "config": {"workload": "polymer slab 100^3 A, 110k atoms, N=1024, 581 phi-frames, fill_bkg, "
                               "smooth 10 vox, max_q 2.0, bins 185^3 cropped to 131^3",
                   "n_atoms": n_atoms, "grid": N, "frames": F, "W": plan.W, "batch": runner.batch},
        "hbm_gbs_algorithmic_step": B_frame * F / dt / 1e9,
        "detector_images_per_s": 1.0 / det_s,
        "detector_ms_per_image": det_s * 1e3,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": (traffic[dom.split("_")[-1]] * nb) if dom.split("_")[-1] in traffic
                     else None,
                     "traffic_note": "calibrated rocprofv3 FETCH_SIZE+WRITE_SIZE bytes per launch, "
                                     "profiles/frames_polymer100_N1024_counters.md",
                     "algorithmic_bytes_per_launch": alg[dom],
                     "launch_ms": times[dom] * 1e3, "frames_per_launch": nb,
                     "kernel_ms": {k: v * 1e3 for k, v in times.items()}},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


def _time_adjoint(model, coords, engine, torch, reps):
    """(forward ms, forward + backward ms) of a DifferentiableSlab from device-resident
    inputs, HIP events on the stream the work is enqueued on."""
    ct = torch.as_tensor(coords).cuda()
    fr, adj = model.descriptors(ct, model.f_re0, model.f_im0)
    st = engine.stream_handle()
    G = torch.randn(model.n_pix ** 2, dtype=torch.float64, device='cuda')
    model.forward_enqueue(fr, adj, st)
    model.backward_enqueue(fr, adj, G, st)
    torch.cuda.synchronize()
    cs = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(cs)
    for _ in range(reps):
        model.forward_enqueue(fr, adj, st)
    ev[1].record(cs)
    for _ in range(reps):
        model.forward_enqueue(fr, adj, st)
        model.backward_enqueue(fr, adj, G, st)
    ev[2].record(cs)
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps, ev[1].elapsed_time(ev[2]) / reps


def _plain_forward_ms(engine, torch, coords, elements, r, qv, mq, det, reps, fill=False, smooth=0):
    """Wall time (ms) of the reference-parity forward (nearest-grid-point frames, bins,
    iq, detector image) of the same slab and detector, from device-resident atoms."""
    ct = torch.as_tensor(coords).cuda()
    ids, ftab, _, _ = engine.element_table(elements, 12700.0)
    idd = torch.as_tensor(ids).cuda()
    psis = np.asarray(det['psis'], dtype=np.float64)
    dplan = engine.detector_plan(det['num_pixels'], mq, det['angle_init_vals'], det['angle_init_axs'], psis,
                                 np.full(len(psis), 1.0 / len(psis)), [0.0], [1.0], [0.0], [1.0])

    def ngp():
        res = engine.run_voxelgrid(ct, None, r, qv, mq, 12700.0, fill_bkg=fill, smooth=smooth, f_table=ftab,
                                   elem_ids=idd)
        engine.run_detector(res['iq64'], res['qx'], res['qy'], res['qz'], dplan, det['mirror'])
    ngp()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        ngp()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps * 1e3


def adjoint_config(engine, torch, reps=10, precision='fp32'):
    """Differentiation workload: graphite 40 x 40 x 30 A (5.6k atoms), r 0.25 A,
    N = 512, 453 phi-frames, max_q 2.5, 256^2 detector with psi 0..90 step 2
    (46 orientations) and a 0.15 deg psi tilt.  Forward = cloud-in-cell frames
    + bins + iq + detector image; backward = the whole hand-written adjoint
    (detector, bins, the frames' row pass reused from the forward where it is still
    in the frame buffer and recomputed otherwise, coordinate and f gradients).
    Both timed from device-resident inputs with HIP events on the stream they
    are enqueued on."""
    from giwaxsim_amd.adjoint import DifferentiableSlab
    from giwaxsim_amd.synth import graphite_slab
    cg, eg = graphite_slab(size=(40.0, 40.0, 30.0), seed=4)
    r, mq = 0.25, 2.5
    qv = 2 * np.pi / (r * 512) * (1 + 1e-9)
    det = dict(num_pixels=256, max_q=mq, psis=list(np.arange(0.0, 91.0, 2.0)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    model = DifferentiableSlab(eg, cg, r, qv, mq, det, precision=precision)
    fwd, fb = _time_adjoint(model, cg, engine, torch, reps)
    ngp_ms = _plain_forward_ms(engine, torch, cg, eg, r, qv, mq, det, reps)
    F = model.plan.n_frames
    return {"workload": "graphite 40x40x30 A, 5.6k atoms, N=512, 453 phi-frames, 256^2 detector, "
                        f"psi 0..90 step 2; cloud-in-cell deposit, {precision.upper()} transforms; "
                        "gradients of the image w.r.t. all coordinates and per-element f",
            "precision": precision,
            "n_atoms": len(cg), "frames": F, "fwd_ms": fwd, "fwd_bwd_ms": fb, "bwd_ms": fb - fwd,
            "fwd_bwd_over_fwd": fb / fwd, "reference_parity_fwd_ms": ngp_ms,
            "fwd_bwd_over_reference_parity_fwd": fb / ngp_ms,
            "fwd_phi_frames_per_s": F / (fwd * 1e-3), "fwd_bwd_phi_frames_per_s": F / (fb * 1e-3)}


def adjoint_polymer_config(engine, torch, reps=3):
    """The production-size workload through the differentiable variant: polymer
    slab 100^3 A (110k atoms), r 0.2 A, N = 1024, fill_bkg, smooth 10 voxels,
    max_q 2.0 (581 frames), 512^2 detector, psi 0..180 step 0.5 (361
    orientations), 0.15 deg psi tilt; FP32 transforms."""
    from giwaxsim_amd.adjoint import DifferentiableSlab
    W = workload()
    det = dict(num_pixels=512, max_q=W['max_q'], psis=list(np.linspace(0, 180, 361)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    model = DifferentiableSlab(W['elements'], W['coords'], W['r'], W['qv'], W['max_q'], det,
                               fill_bkg=W['fill'], smooth=W['smooth'])
    fwd, fb = _time_adjoint(model, W['coords'], engine, torch, reps)
    ngp_ms = _plain_forward_ms(engine, torch, W['coords'], W['elements'], W['r'], W['qv'], W['max_q'], det,
                               max(reps, 5), fill=W['fill'], smooth=W['smooth'])
    F = model.plan.n_frames
    return {"workload": "polymer slab 100^3 A, 110k atoms, N=1024, 581 phi-frames, fill_bkg, smooth 10, "
                        "512^2 detector, psi 0..180 step 0.5; cloud-in-cell deposit, FP32 transforms; "
                        "gradients w.r.t. all coordinates and per-element f",
            "n_atoms": len(W['coords']), "frames": F, "batch": model.batch, "fwd_ms": fwd,
            "fwd_bwd_ms": fb, "bwd_ms": fb - fwd, "fwd_bwd_over_fwd": fb / fwd,
            "reference_parity_fwd_ms": ngp_ms, "fwd_bwd_over_reference_parity_fwd": fb / ngp_ms,
            "fwd_phi_frames_per_s": F / (fwd * 1e-3), "fwd_bwd_phi_frames_per_s": F / (fb * 1e-3)}


def adjoint_large_config(engine, torch, n_slabs=16):
    """Large slabs through the differentiable variant: n_slabs polymer slabs of
    200 x 200 x 60 A (~1.3M atoms each, seeds 0..n-1: seeded crystallite
    orientations), r 0.15 A, N = 2048, fill_bkg, smooth 20 voxels (3 A), max_q
    3.0 (1307 frames), 2048^2 detector, psi 0..180 step 0.25 (721 orientations).
    Per slab: forward, then backward w.r.t. the coordinates (frames in batches sized
    to the free device memory; the row pass of the batch the forward left in the
    frame buffer is reused, other batches are recomputed).  Slabs run one after another on one model's
    buffers' worth of memory."""
    from giwaxsim_amd.adjoint import DifferentiableSlab
    from giwaxsim_amd.synth import polymer_slab
    r, mq = 0.15, 3.0
    qv = 2 * np.pi / (r * 2048) * (1 + 1e-9)
    det = dict(num_pixels=2048, max_q=mq, psis=list(np.linspace(0, 180, 721)), phis=[0.0],
               thetas=[0.0], angle_init_vals=(0.15,), angle_init_axs=('psi',), mirror=True)
    fwd_tot = fb_tot = 0.0
    n_atoms = 0
    F = 0
    peak = 0
    for i in range(n_slabs):
        c, e = polymer_slab(size=(200.0, 200.0, 60.0), seed=i, density=0.54)
        model = DifferentiableSlab(e, c, r, qv, mq, det, fill_bkg=True, smooth=20)
        fwd, fb = _time_adjoint(model, c, engine, torch, 1)
        fwd_tot += fwd
        fb_tot += fb
        n_atoms += len(c)
        F = model.plan.n_frames
        peak = max(peak, torch.cuda.max_memory_allocated())
        del model
        if i == 0:   # the reference-parity forward of the first slab and the same detector
            ngp_ms = _plain_forward_ms(engine, torch, c, e, r, qv, mq, det, 2, fill=True, smooth=20)
    fb_slab = fb_tot / n_slabs
    return {"workload": f"{n_slabs} polymer slabs 200x200x60 A (~1.3M atoms each), N=2048, fill_bkg, "
                        "smooth 20, max_q 3.0, 2048^2 detector psi step 0.25; forward + backward "
                        "w.r.t. coordinates, forward's last frame batch reused, others recomputed",
            "n_slabs": n_slabs, "n_atoms": n_atoms, "frames_per_slab": F,
            "fwd_ms_total": fwd_tot, "fwd_bwd_ms_total": fb_tot, "fwd_bwd_over_fwd": fb_tot / fwd_tot,
            "fwd_bwd_ms_per_slab": fb_slab, "reference_parity_fwd_ms_per_slab": ngp_ms,
            "fwd_bwd_over_reference_parity_fwd": fb_slab / ngp_ms, "peak_hbm_gb": peak / 1e9}


def extra_configs(args, engine, torch, which=None):
    """Ensemble of 64 uneven slabs into one shared grid (HIP-graph replay), one
    large slab (200 x 200 x 60 A at ~1.3M atoms, N=2048, 1307 frames) and the
    differentiation workload (forward + hand-written adjoint).  ``which`` runs
    just one of the last two."""
    out = {}
    if which is None:
        out["adjoint"] = adjoint_config(engine, torch, precision='fp32')
        out["adjoint_fp64"] = adjoint_config(engine, torch, precision='fp64')
        out["adjoint_polymer"] = adjoint_polymer_config(engine, torch)
        out["adjoint_large"] = adjoint_large_config(engine, torch, n_slabs=16)
    reps = 3
    if which in (None, "compare"):
        out["compare"] = compare_config(engine, torch)
    if which in (None, "slab_build"):
        out["slab_build"] = slab_build_config(engine, torch)
    if which in (None, "fit_eval"):
        out["fit_eval"] = fit_eval_config(engine, torch)
    if which in (None, "ensemble"):
        out["ensemble"] = ensemble_config(engine, torch, reps)
    if which in (None, "large_slab"):
        out["large_slab"] = large_slab_config(engine, torch, reps)
    return out


def fit_eval_config(engine, torch, reps=5):
    """One evaluation of the slab-size fit model (fit_slabsize.py fullfit_model) at the
    reference's fit configuration (config_templates/fit_slabsize_config.txt geometry):
    the PM6 cell tiled and cropped to 176 x 43 x 186 A, r 0.3 A, q_voxel 0.02 (N = 1048,
    chirp-z transforms), max_q 2, fill_bkg, smooth 25, 250 px detector, psi 0..90 x 91,
    phi 0..179 x 180 (16,380 orientations, psi outer), 90-degree psi / phi / psi initial
    rotations, pad 0.12 in 1.05..1.55; a synthetic experimental map on 981 x 1043
    (qxy, qz) grids of the reference's TC47 shape.  Slab build, voxel grid, detector and
    comparison all on the device; only the residual vector returns."""
    from giwaxsim_amd.fit_slabsize import SlabSizeModel
    from giwaxsim_amd.tools.comparison import mirror_qmap_positive_qxy_only, rebin_qmap_and_mask, trim_sim_data
    path = os.path.join(ROOT, "tests", "golden", "PM6_sample.pdb")
    slab = (path, 44.456, 45.726, 40.097, 90.0, 90.0, 90.0)
    vox = (0.3, 0.02, 2.0, 12700.0, True, 25)
    psis = np.linspace(0, 90, 91)
    phis = np.linspace(0, 179, 180)
    thetas = np.linspace(0, 1, 1)
    det = (250, (90.0, 90.0, 90.0), ('psi', 'phi', 'psi'), psis, None, phis, None, thetas, None)
    exp_qxy = np.linspace(-1.304, 2.111, 981)
    exp_qz = np.linspace(-0.0287, 2.508, 1043)
    X, Z = np.meshgrid(exp_qxy, exp_qz)
    exp_img = 1e3 * (1.5 + np.cos(7 * np.hypot(X, Z)))
    e, q2, z2 = mirror_qmap_positive_qxy_only(exp_img, exp_qxy, exp_qz)
    h = np.linspace(-2.0, 2.0, 250)
    _, ht, vt = trim_sim_data(np.zeros((250, 250)), h, h, q2, z2)
    mask = np.zeros(e.shape)
    mask[:, :10] = 1
    rebin_map, rebin_mask = rebin_qmap_and_mask(ht, vt, e, q2, z2, mask)
    rebin_map[rebin_mask == 1] = 0
    model = SlabSizeModel(slab, vox, det, (rebin_map, rebin_mask, q2, z2, 0.12, (1.05, 1.55)))
    r, _ = model.residual(176.0, 43.0, 186.0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(reps):
        r, (sc, off) = model.residual(176.0 + 0.5 * k, 43.0, 186.0)   # a new slab each evaluation
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    return {"workload": "fit_slabsize evaluation at the reference fit geometry: PM6 slab 176 x 43 x 186 A, "
                        "N=1048 (chirp-z), fill_bkg, smooth 25, 250^2 detector x 16,380 orientations, "
                        "synthetic 981 x 1043 experimental map; slab, voxel grid, detector, comparison on the "
                        "device", "ms_per_eval": dt * 1e3, "n_pixels_residual": int(r.size),
            "orientations": int(len(psis) * len(phis) * len(thetas))}


def compare_config(engine, torch, reps=20, cpu_reps=3):
    """Experimental side of a fit evaluation (fit_slabsize.py fullfit_model after the
    detector image): trim to the experimental range, shift_peak, scale / offset,
    masked residual, on the device (FitComparison.residual) beside the oracle's numpy
    / scipy version on the host, for the production workload's 512^2 mirrored
    detector image and a synthetic 400 x 360 experimental map rebinned onto it."""
    import oracle.comparison as OC
    from giwaxsim_amd.tools.comparison import FitComparison, rebin_qmap_and_mask
    W = workload()
    ids, ftab, counts, _ = engine.element_table(W['elements'], W['energy'])
    res = engine.run_voxelgrid(torch.as_tensor(W['coords']).cuda(), None, W['r'], W['qv'], W['max_q'],
                               W['energy'], fill_bkg=W['fill'], smooth=W['smooth'], f_table=ftab,
                               elem_ids=torch.as_tensor(ids).cuda(), want64=True)
    from giwaxsim_amd.tools.comparison import _weights
    psis = np.linspace(0, 180, 361)
    one = np.array([0.0])
    dplan = engine.detector_plan(512, W['max_q'], (0.15, 0, 0), ('psi', 'None', 'None'), psis,
                                 _weights(None, psis), one, _weights(None, one), one, _weights(None, one))
    det, _ = engine.run_detector(res['iq64'], res['qx'], res['qy'], res['qz'], dplan, True)
    exp_qxy = np.linspace(0.0, 1.95, 400)
    exp_qz = np.linspace(0.0, 1.9, 360)
    X, Z = np.meshgrid(exp_qxy, exp_qz)
    exp_img = 1e-3 * (2.0 + np.cos(9 * np.hypot(X, Z)))
    mask = np.zeros(X.shape)
    mask[:, :12] = 1
    _, h_t, v_t = OC.trim_sim_data(np.zeros((len(dplan.v), len(dplan.h))), dplan.h, dplan.v, exp_qxy, exp_qz)
    rebin_map, rebin_mask = rebin_qmap_and_mask(h_t, v_t, exp_img, exp_qxy, exp_qz, mask)
    rebin_map = np.where(np.isnan(rebin_map), 0.0, rebin_map)
    rebin_map[rebin_mask == 1] = 0
    pad = (0.05, (1.2, 1.8))
    fc = FitComparison(dplan.h, dplan.v, rebin_map, rebin_mask, exp_qxy, exp_qz, *pad)
    fc.residual(det)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        r, s, b = fc.residual(det)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    det_np = det.cpu().numpy()
    tc = time.perf_counter()
    for _ in range(cpu_reps):
        ref, _ = OC.fit_residual(det_np, dplan.h, dplan.v, rebin_map, rebin_mask, exp_qxy, exp_qz, *pad)
    tcpu = (time.perf_counter() - tc) / cpu_reps
    err = float(np.max(np.abs(r.cpu().numpy() - ref)) / np.max(np.abs(ref)))
    return {"workload": "fit evaluation, experimental side: 512^2 mirrored detector image of the polymer "
                        "slab trimmed to qxy 0..1.95, qz 0..1.9 ({} x {}), shift_peak pad 0.05 in "
                        "1.2..1.8, scale/offset, masked residual".format(len(v_t), len(h_t)),
            "ms_per_eval": dt * 1e3, "cpu_ms_per_eval": tcpu * 1e3,
            "cpu_kind": "oracle (numpy + scipy.ndimage.map_coordinates, 1 thread)",
            "max_rel_diff_vs_cpu": err}


def slab_build_config(engine, torch, reps=20, cpu_reps=3):
    """Slab builder of a fit evaluation (tools/comparison.py slabmaker_fitting): the
    PM6 polymer cell (3840 atoms, 44.456 x 45.726 x 40.097 A) tiled 9 x 3 x 11 and
    cropped to 176 x 43 x 186 A, on the device (SlabBuilder.build: tile, bounds, crop,
    bounds, shift, two small read-backs) beside the oracle's numpy version."""
    import oracle.comparison as OC
    from giwaxsim_amd.tools.comparison import SlabBuilder
    from giwaxsim_amd.tools.utilities import calc_real_space_abc, load_pdb
    path = os.path.join(ROOT, "tests", "golden", "PM6_sample.pdb")
    cell = (44.456, 45.726, 40.097, 90.0, 90.0, 90.0)
    size = (176.0, 43.0, 186.0)
    sb = SlabBuilder(path, *cell)
    x, y, z, el = sb.build(*size)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        x, y, z, el = sb.build(*size)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    coords, elements = load_pdb(path)
    a, b, c = calc_real_space_abc(*cell)
    tc = time.perf_counter()
    for _ in range(cpu_reps):
        ref, ref_el = OC.slabmaker(coords, elements, *size, a, b, c)
    tcpu = (time.perf_counter() - tc) / cpu_reps
    got = np.stack([t_.cpu().numpy() for t_ in (x, y, z)], axis=1)
    n_rep = len(coords) * 9 * 3 * 11
    return {"workload": "PM6 cell (3840 atoms) tiled 9 x 3 x 11 ({} atoms), cropped to 176 x 43 x 186 A "
                        "({} atoms)".format(n_rep, len(ref)),
            "ms_per_build": dt * 1e3, "cpu_ms_per_build": tcpu * 1e3,
            "cpu_kind": "oracle (numpy, 1 thread)",
            "bit_exact_vs_cpu": bool(np.array_equal(got, ref) and
                                     np.array_equal(sb.uniq[el.cpu().numpy()], ref_el))}


def ensemble_config(engine, torch, reps):
    from giwaxsim_amd.synth import ensemble_slabs
    qv = 2 * np.pi / (0.2 * 1024) * (1 + 1e-9)
    slabs = ensemble_slabs(n_slabs=64, seed=0, q_voxel=qv)
    ens = engine.Ensemble(slabs, qv, 2.0, 12700.0, fill_bkg=True, smooth=10)
    ens.capture()
    ens.replay()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        ens.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    return {"workload": "64 polymer slabs, edges U[40,160] A, N in {256,512,1024,2048}, "
                        "one shared 185^3 grid (131^3 crop), fill_bkg, smooth 10",
            "n_atoms": int(sum(len(c) for c, _, _ in slabs)), "frames": ens.n_frames,
            "ms_per_pass": dt * 1e3, "phi_frames_per_s": ens.n_frames / dt}


def large_slab_config(engine, torch, reps):
    from giwaxsim_amd.synth import polymer_slab
    c, e = polymer_slab(size=(200.0, 200.0, 60.0), seed=5, density=0.54)
    r = 0.15
    ids, ftab, _, _ = engine.element_table(e, 12700.0)
    cd = torch.as_tensor(c).cuda()
    idd = torch.as_tensor(ids).cuda()
    run = lambda: engine.run_voxelgrid(cd, None, r, 2 * np.pi / (r * 2048) * (1 + 1e-9), 3.0, 12700.0,
                                       fill_bkg=True, smooth=20, f_table=ftab, elem_ids=idd)
    res = run()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        res = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    # its detector: 2048 x 2048 px, psi 0..180 in 0.25 deg steps (721 orientations)
    from giwaxsim_amd.tools.comparison import _weights
    psis = np.linspace(0, 180, 721)
    one = np.array([0.0])
    dplan = engine.detector_plan(2048, 3.0, (0.15, 0, 0), ('psi', 'None', 'None'), psis,
                                 _weights(None, psis), one, _weights(None, one), one,
                                 _weights(None, one))
    engine.run_detector(res['iq32'], res['qx'], res['qy'], res['qz'], dplan, True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    engine.run_detector(res['iq32'], res['qx'], res['qy'], res['qz'], dplan, True)
    torch.cuda.synchronize()
    det_ms = (time.perf_counter() - t) * 1e3
    return {"workload": "polymer slab 200x200x60 A, 1.3M atoms, N=2048, fill_bkg, smooth 20, "
                        "max_q 3.0, 417^3 grid (295^3 crop); detector 2048^2, psi step 0.25",
            "n_atoms": len(c), "frames": res['plan'].n_frames, "ms_per_slab": dt * 1e3,
            "phi_frames_per_s": res['plan'].n_frames / dt, "detector_ms_per_image": det_ms,
            "detector_images_per_s": 1e3 / det_ms}


if __name__ == "__main__":
    main()
