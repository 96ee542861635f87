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
    ap.add_argument("--cpu-frames-1t", type=int, default=48,
                    help="phi-frames of the single-threaded CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-iters", type=int, default=20)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the ensemble (64 slabs) and large-slab (N=2048) figures")
    ap.add_argument("--only", choices=["adjoint", "adjoint_fp64", "adjoint_polymer", "adjoint_large", "adjoint_fit",
                                       "ensemble", "large_slab", "compare", "slab_build", "fit_eval",
                                       "graphite50"],
                    help="run one of the extra workloads alone and print its JSON (for profiling)")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="one step, then one 256 MiB device copy (counter calibration) and one "
                         "launch of each frame kernel; for rocprofv3 --pmc passes")
    return ap.parse_args()


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
        elif args.only == "adjoint_fit":
            print(json.dumps(adjoint_fit_config(engine, torch, reps=args.steps)))
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
        d = runner.desc
        keep = (d.row_variant, d.inter_rm, d.fft_plan_row, d.fft_plan_col)
        plans = (1 << 4) | (1 << 6)   # radix 16.8.8 in both passes (fixed, so every pass profiles the same code)
        for shape in (0, 8):   # row-major and column-major intermediates, one launch of each kernel
            engine.set_row_variant(d, shape | plans)
            for kname in ("gx_frames_stats", "gx_frames_columns", "gx_frames_rowpass", "gx_frames_colpass"):
                call(kname, ctypes.byref(d), 0, nb, engine.stream_handle())
        d.row_variant, d.inter_rm, d.fft_plan_row, d.fft_plan_col = keep
        # one detector image at this workload's geometry (512^2, 361 orientations) and one
        # at the 1.3M-atom slab's (2048^2, 721 orientations, 295^3 I(q); values synthetic)
        detector_seconds(engine, torch, res, 512, W['max_q'], (0.15, 0, 0), ('psi', 'None', 'None'),
                         np.linspace(0, 180, 361), 0)
        g5 = engine.qgrid(2 * np.pi / (0.15 * 2048) * (1 + 1e-9), 3.0)
        ax5 = g5.crop_axis
        iq5 = torch.rand((g5.c_n,) * 3, dtype=torch.float64, device=dev)
        detector_seconds(engine, torch, {'iq64': iq5, 'qx': ax5, 'qy': ax5, 'qz': ax5}, 2048, 3.0,
                         (0.15, 0, 0), ('psi', 'None', 'None'), np.linspace(0, 180, 721), 0)
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernels": True, "frames_per_launch": nb, "copy_bytes": 2 * 4 * src.numel(),
                          "plans": [int(d.fft_plan_row), int(d.fft_plan_col)]}))
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

    # ---- frame kernels, timed live with HIP events on the stream they run on
    runner = res['runner']
    times = frame_kernel_times(runner, engine, torch, call, args.kernel_iters)
    roof = frame_roofline(res, times, n_atoms, engine)

    # ---- detector images/s
    det_s = detector_seconds(engine, torch, res, 512, W['max_q'], (0.15, 0, 0), ('psi', 'None', 'None'),
                             np.linspace(0, 180, 361), max(3, args.steps // 4))

    # ---- other synthetic configurations, reported beside the headline figure
    extra = {}
    if not args.no_extra:
        extra = extra_configs(args, engine, torch)

    # ---- CPU baseline: the oracle in reference-faithful threaded mode on a bounded sample
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(W, ftab, ids, plan, args)

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
                   "n_atoms": n_atoms, "grid": N, "frames": F, "W": plan.W, "batch": runner.batch,
                   "frame_pass_shape": shape_name(runner)},
        "detector_images_per_s": 1.0 / det_s,
        "detector_ms_per_image": det_s * 1e3,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))



PEAK_FP32_VALU_TFLOPS = 78.6   # MI355X unpacked FP32 vector rate (packed FP32 is not used)
CLOCK_GHZ = 2.4                # MI355X peak engine clock
N_SIMD = 256 * 4


def shape_name(runner):
    d = runner.desc
    rm = "row-major intermediate, per-wave rows" if d.inter_rm else (
        "column-major intermediate, " + ("per-wave rows" if (d.row_variant & 3) == 0 else
                                         f"block deposit over {4 * (d.row_variant & 3)} rows"))
    plans = {1024: ("16.16.4", "16.8.8", "8.16.8", "8.8.16"), 2048: ("16.16.8", "16.8.16", "8.16.16")}
    out = rm + (", XCD-aware order" if d.row_variant & 4 else ", frame-major order")
    if int(d.N) in plans and not d.bs_M:
        p = plans[int(d.N)]
        out += f", radix {p[d.fft_plan_row]} rows / {p[d.fft_plan_col]} columns"
    return out


def frame_kernel_times(runner, engine, torch, call, iters):
    """Seconds per launch of the row and column passes over the runner's first frame batch,
    from HIP events recorded on the stream the kernels are launched on."""
    st = torch.cuda.current_stream()
    nb = runner.batch
    h = engine.stream_handle()
    call("gx_frames_stats", ctypes.byref(runner.desc), 0, nb, h)
    if runner.desc.fill_bkg or runner.desc.smooth:
        call("gx_frames_columns", ctypes.byref(runner.desc), 0, nb, h)
    times = {}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for kname in ("gx_frames_rowpass", "gx_frames_colpass"):
        call(kname, ctypes.byref(runner.desc), 0, nb, h)
        ev[0].record(st)
        for _ in range(iters):
            call(kname, ctypes.byref(runner.desc), 0, nb, h)
        ev[1].record(st)
        torch.cuda.synchronize()
        times[kname] = ev[0].elapsed_time(ev[1]) / iters / 1e3
    runner.bins.clear()
    return times


def counters_profile():
    """Per-kernel figures of profiles/frames_polymer100_N1024_counters.md (rocprofv3 --pmc
    passes of `bench.py --profile-kernels`, each frame-pass shape with radix 16.8.8 in both
    passes): calibrated HBM bytes per frame and the VALU issue share of each row / column
    pass kernel."""
    path = os.path.join(ROOT, "profiles", "frames_polymer100_N1024_counters.md")
    out = {}
    try:
        for line in open(path):
            for key in ("traffic_bytes_per_frame:", "valu_issue_share:"):
                if line.startswith(key):
                    out[key[:-1]] = {k: float(v) for k, v in (t.split("=") for t in line.split(":", 1)[1].split())}
    except OSError:
        pass
    return out


def frame_roofline(res, times, n_atoms, engine):
    """The roofline record of the dominant frame kernel.  `achieved` / `frac`: the fixed
    algorithmic-byte contract (24 n_atoms + 16 N^2 + 16 P per phi-frame: the row pass is
    charged 24 n_atoms + 8 N^2, the column pass 8 N^2 + 16 P) over the live launch time;
    `hbm_traffic_frac`: calibrated FETCH+WRITE bytes of the same kernel (counter profile)
    over the live launch time; `issue`: what bounds it (VALU issue share from the counter
    profile, FFT FLOP/s against the unpacked FP32 vector rate)."""
    runner, plan = res['runner'], res['plan']
    F, N, nb, W = plan.n_frames, plan.N, runner.batch, plan.W
    P_mean = float(res['bins'].cnt.sum().item()) / F
    alg = {"gx_frames_rowpass": nb * (24 * n_atoms + 8 * N * N),
           "gx_frames_colpass": nb * (8 * N * N + 16 * P_mean)}
    dom = max(times, key=times.get)
    achieved = alg[dom] / times[dom] / 1e9
    fam = {"gx_frames_rowpass": "k_rowpass_rm" if runner.desc.inter_rm else "k_rowpass",
           "gx_frames_colpass": "k_colpass_rm" if runner.desc.inter_rm else "k_colpass"}
    prof = counters_profile()
    tr = prof.get("traffic_bytes_per_frame", {}).get(fam[dom])
    traffic = tr * nb if tr is not None else None
    # FFT flops: 5 N log2 N per transform; rows that hold atoms (row pass), kept columns (column pass)
    off = runner.slab.row_off[:plan.Z + 1].cpu().numpy()
    rows = int(np.count_nonzero(np.diff(off) > 0))
    fft_flops = {"gx_frames_rowpass": nb * rows * 5 * N * np.log2(N),
                 "gx_frames_colpass": nb * W * 5 * N * np.log2(N)}
    issue = {}
    for k in times:
        tf = fft_flops[k] / times[k] / 1e12
        issue[k] = {"fft_tflops": tf, "fft_frac_of_fp32_valu": tf / PEAK_FP32_VALU_TFLOPS,
                    "valu_issue_share": prof.get("valu_issue_share", {}).get(fam[k])}
    return {"bound": "hbm", "kernel": dom, "kernel_function": fam[dom], "achieved": achieved,
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
            "hbm_traffic_frac": (traffic / times[dom] / 1e9 / PEAK_HBM_GBS) if traffic else None,
            "traffic_note": "calibrated rocprofv3 FETCH_SIZE+WRITE_SIZE bytes per launch, "
                            "profiles/frames_polymer100_N1024_counters.md",
            "algorithmic_bytes_per_launch": alg[dom], "launch_ms": times[dom] * 1e3,
            "frames_per_launch": nb, "kernel_ms": {k: v * 1e3 for k, v in times.items()},
            "issue": issue}


def detector_seconds(engine, torch, res, n_pix, max_q, init_vals, init_axs, psis, reps):
    from giwaxsim_amd.tools.comparison import _weights
    one = np.array([0.0])
    dplan = engine.detector_plan(n_pix, max_q, init_vals, init_axs, psis, _weights(None, psis), one,
                                 _weights(None, one), one, _weights(None, one))
    iq = res['iq64']
    engine.run_detector(iq, res['qx'], res['qy'], res['qz'], dplan, True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        engine.run_detector(iq, res['qx'], res['qy'], res['qz'], dplan, True)
    torch.cuda.synchronize()
    return (time.perf_counter() - t1) / max(reps, 1)


def host_description():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    blas = None
    try:
        blas = np.show_config(mode="dicts")["Build Dependencies"]["blas"]["name"]
    except Exception:
        pass
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    return model, blas, env


def cpu_baseline(W, ftab, ids, plan, args):
    """The oracle timed on this host: the reference's ThreadPoolExecutor() fan-out over frames
    and over detector orientations (tools/comparison.py:759-762, 835-853) on bounded samples,
    and single-threaded samples of both."""
    import oracle as O
    F = plan.n_frames
    phis = plan.grid.phis[:args.cpu_frames]
    workers = O.default_threads()
    fv = ftab[ids]
    passes, tcpu = 0, 0.0
    res = None
    while passes < 8 and (passes == 0 or tcpu < args.cpu_seconds):
        tc = time.perf_counter()
        res = O.voxelgridmaker(W['coords'], fv, W['r'], W['qv'], W['max_q'], fill_bkg=W['fill'],
                               smooth=W['smooth'], phis=phis, threads=None)
        tcpu += time.perf_counter() - tc
        passes += 1
    # one thread, a sample of frames spread over phi
    ph1 = plan.grid.phis[::max(1, F // args.cpu_frames_1t)][:args.cpu_frames_1t]
    tc = time.perf_counter()
    O.voxelgridmaker(W['coords'], fv, W['r'], W['qv'], W['max_q'], fill_bkg=W['fill'], smooth=W['smooth'],
                     phis=ph1, threads=0)
    t1 = time.perf_counter() - tc
    # detector stage: 512^2 px, psi 0..180 step 0.5 (361 orientations), 0.15 deg psi tilt
    psis = np.linspace(0, 180, 361)
    one = np.array([0.0])
    dargs = (res['iq'], res['qx'], res['qy'], res['qz'], 512, W['max_q'], (0.15, 0, 0), ('psi', 'None', 'None'))
    tc = time.perf_counter()
    O.detectormaker(*dargs, psis, None, one, None, one, None, mirror=True, threads=None)
    tdet = time.perf_counter() - tc
    ps1 = psis[::12]
    tc = time.perf_counter()
    O.detectormaker(*dargs, ps1, None, one, None, one, None, mirror=True, threads=0)
    tdet1 = (time.perf_counter() - tc) * len(psis) / len(ps1)
    model, blas, env = host_description()
    return {"value": passes * len(phis) / tcpu, "unit": "phi-frames/s", "cores": os.cpu_count(),
            "kind": "port",
            "sample": f"{passes} pass(es) over {len(phis)} of {F} phi-frames of the same slab, FP64 numpy "
                      f"oracle with ThreadPoolExecutor() default workers ({workers}); {tcpu:.1f} s",
            "workers": workers, "value_1thread": len(ph1) / t1,
            "sample_1thread": f"{len(ph1)} phi-frames spread over 0..180 deg, threads=0; {t1:.1f} s",
            "detector_s_per_image": tdet, "detector_workers": workers,
            "detector_s_per_image_1thread": tdet1,
            "detector_sample_1thread": f"{len(ps1)} of 361 orientations, scaled to 361",
            "cpu_model": model, "numpy_blas": blas, "blas_thread_env": env}

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
    psis, phis, thetas = (np.asarray(det.get(k, [0.0]), dtype=np.float64) for k in ('psis', 'phis', 'thetas'))
    uni = lambda a: np.full(len(a), 1.0 / len(a))
    dplan = engine.detector_plan(det['num_pixels'], mq, det['angle_init_vals'], det['angle_init_axs'], psis,
                                 uni(psis), phis, uni(phis), thetas, uni(thetas))

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


def adjoint_fit_config(engine, torch, reps=3):
    """The differentiable variant at the reference's fit geometry
    (config_templates/fit_slabsize_config.txt): the PM6 cell tiled and cropped to
    176 x 43 x 186 A (66k atoms), r 0.3 A, q_voxel 0.02 -> N = 1048 (zoomed chirp-z
    band transforms and their adjoints), max_q 2, fill_bkg, smooth 25, 892 phi-frames,
    250^2 detector, psi 0..90 x 91 and phi 0..179 x 180 (16,380 orientations) after the
    90-degree psi / phi / psi initial rotations; FP32 transforms."""
    from giwaxsim_amd.adjoint import DifferentiableSlab
    from giwaxsim_amd.tools.comparison import slabmaker_fitting
    coords, el = slabmaker_fitting(os.path.join(ROOT, "tests", "golden", "PM6_sample.pdb"), 176.0, 43.0, 186.0,
                                   44.456, 45.726, 40.097, 90.0, 90.0, 90.0)
    det = dict(num_pixels=250, max_q=2.0, psis=list(np.linspace(0, 90, 91)), phis=list(np.linspace(0, 179, 180)),
               thetas=[0.0], angle_init_vals=(90.0, 90.0, 90.0), angle_init_axs=('psi', 'phi', 'psi'), mirror=True)
    model = DifferentiableSlab(el, coords, 0.3, 0.02, 2.0, det, fill_bkg=True, smooth=25)
    fwd, fb = _time_adjoint(model, coords, engine, torch, reps)
    ngp_ms = _plain_forward_ms(engine, torch, coords, el, 0.3, 0.02, 2.0, det, max(reps, 3), fill=True, smooth=25)
    F = model.plan.n_frames
    return {"workload": "PM6 slab 176 x 43 x 186 A (fit geometry), N=1048 (chirp-z), fill_bkg, smooth 25, "
                        "892 phi-frames, 250^2 detector x 16,380 orientations; cloud-in-cell deposit, FP32 "
                        "transforms; gradients w.r.t. all coordinates and per-element f",
            "n_atoms": len(coords), "grid": model.plan.N, "frames": F, "batch": model.batch, "fwd_ms": fwd,
            "fwd_bwd_ms": fb, "bwd_ms": fb - fwd, "fwd_bwd_over_fwd": fb / fwd,
            "reference_parity_fwd_ms": ngp_ms, "fwd_bwd_over_reference_parity_fwd": fb / ngp_ms}


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
        out["adjoint_fit"] = adjoint_fit_config(engine, torch)
        out["adjoint_large"] = adjoint_large_config(engine, torch, n_slabs=16)
    reps = 3
    if which in (None, "graphite50"):
        out["graphite50"] = graphite50_config(engine, torch, reps=10)
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


def graphite50_config(engine, torch, reps=10):
    """The small plain configuration: AB graphite 50^3 A (14.5k atoms, jitter 0.02 A), r 0.25 A,
    N = 512, no fill, no smoothing, max_q 2.5 (453 phi-frames, FP64 transforms: a sparse grid),
    and its 256^2 detector over psi 0..180 in 1 degree steps (181 orientations)."""
    from giwaxsim_amd._lib import call
    from giwaxsim_amd.synth import graphite_slab
    c, e = graphite_slab(size=(50.0, 50.0, 50.0), seed=0)
    r = 0.25
    qv = 2 * np.pi / (r * 512) * (1 + 1e-9)
    ids, ftab, _, _ = engine.element_table(e, 12700.0)
    cd, idd = torch.as_tensor(c).cuda(), torch.as_tensor(ids).cuda()
    run = lambda: engine.run_voxelgrid(cd, None, r, qv, 2.5, 12700.0, f_table=ftab, elem_ids=idd)
    res = run()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        res = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    times = frame_kernel_times(res['runner'], engine, torch, call, 20)
    roof = frame_roofline(res, times, len(c), engine)
    for k in ("traffic", "hbm_traffic_frac", "traffic_note"):   # counters are for the polymer workload
        roof.pop(k, None)
    det_s = detector_seconds(engine, torch, res, 256, 2.5, (0, 0, 0), ('None',) * 3, np.linspace(0, 180, 181), 10)
    F = res['plan'].n_frames
    return {"workload": "graphite 50^3 A, 14.5k atoms, r 0.25 A, N=512, no fill/smooth (FP64 transforms), "
                        "max_q 2.5, 453 phi-frames; detector 256^2, psi 0..180 step 1",
            "n_atoms": len(c), "frames": F, "ms_per_step": dt * 1e3, "phi_frames_per_s": F / dt,
            "fft_precision": "fp64" if res['runner'].f64 else "fp32",
            "frame_pass_shape": shape_name(res['runner']), "roofline": roof,
            "detector_ms_per_image": det_s * 1e3, "detector_images_per_s": 1.0 / det_s}


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


def ensemble_config(engine, torch, reps, cpu_frames=3):
    """64 polymer slabs of uneven sizes and grid edges into one shared grid: the whole pass
    captured in a HIP graph and replayed, against 64 sequential plain runs (each slab through
    run_voxelgrid on its own, as the reference's input_folder loop processes file after
    file) on the GPU, and against the oracle on the host (single thread), timed on a sample
    of frames of one slab per grid edge and scaled by the frames of each edge."""
    import oracle as O
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
    plans = {}
    for rn in ens.runners:
        p = plans.setdefault(int(rn.plan.N), {"slabs": 0, "frames_per_launch": set(), "shapes": set()})
        p["slabs"] += 1
        p["frames_per_launch"].add(int(rn.batch))
        p["shapes"].add(shape_name(rn))
    plans = {str(k): {"slabs": v["slabs"], "frames_per_launch": sorted(v["frames_per_launch"]),
                      "frame_pass_shapes": sorted(v["shapes"])} for k, v in sorted(plans.items())}
    # 64 sequential plain runs on the GPU (atoms already resident)
    dev = [(torch.as_tensor(c).cuda(), e, r) for c, e, r in slabs]
    prep = []
    for cd, e, r in dev:
        ids, ftab, _, _ = engine.element_table(e, 12700.0)
        prep.append((cd, torch.as_tensor(ids).cuda(), ftab, r))

    def sequential():
        for cd, idd, ftab, r in prep:
            engine.run_voxelgrid(cd, None, r, qv, 2.0, 12700.0, fill_bkg=True, smooth=10, f_table=ftab,
                                 elem_ids=idd)
    sequential()
    torch.cuda.synchronize()
    t = time.perf_counter()
    sequential()
    torch.cuda.synchronize()
    dt_seq = time.perf_counter() - t
    # the oracle, one thread: cpu_frames frames of one slab per grid edge
    per_edge = {}
    frames_per_edge = {}
    for (c, e, r), rn in zip(slabs, ens.runners):
        N = int(rn.plan.N)
        frames_per_edge[N] = frames_per_edge.get(N, 0) + rn.plan.n_frames
        if N in per_edge:
            continue
        ids, ftab, _, _ = engine.element_table(e, 12700.0)
        phis = rn.plan.grid.phis[::max(1, rn.plan.n_frames // cpu_frames)][:cpu_frames]
        tc = time.perf_counter()
        O.voxelgridmaker(c, ftab[ids], r, qv, 2.0, fill_bkg=True, smooth=10, phis=phis, threads=0)
        per_edge[N] = ((time.perf_counter() - tc) / len(phis), len(c))
    cpu_est = sum(per_edge[N][0] * frames_per_edge[N] for N in frames_per_edge)
    return {"workload": "64 polymer slabs, edges U[40,160] A, N in {256,512,1024,2048}, "
                        "one shared 185^3 grid (131^3 crop), fill_bkg, smooth 10",
            "n_atoms": int(sum(len(c) for c, _, _ in slabs)), "frames": ens.n_frames,
            "ms_per_pass": dt * 1e3, "phi_frames_per_s": ens.n_frames / dt, "plans_by_grid_edge": plans,
            "sequential_plain_runs_ms": dt_seq * 1e3, "speedup_vs_sequential_plain_runs": dt_seq / dt,
            "cpu_oracle_s_estimate": cpu_est, "speedup_vs_cpu_oracle": cpu_est / dt,
            "cpu_note": f"oracle single thread, {cpu_frames} frames of the first slab of each grid edge "
                        "(seconds per frame of that slab) times the frames of that edge; per-slab atom "
                        "counts differ, so this is an estimate",
            "cpu_s_per_frame_by_edge": {str(k): v[0] for k, v in sorted(per_edge.items())}}


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
