"""GPU parity at the benchmark configurations' own geometry and atom density.

Each case runs the HIP path and the FP64 oracle on the same seeded slab:

* graphite 50^3 A (14k atoms), r 0.25 A, N = 512, no fill, no smoothing, max_q 2.5:
  every one of the 453 phi-frames (a sparse grid, so FP64 transforms), then the
  256 x 256 detector over psi 0..180 in 1 degree steps fed the oracle's I(q);
* the 110k-atom polymer slab of 100^3 A, r 0.2 A, N = 1024, fill_bkg, smooth 10,
  max_q 2.0: 16 phi-frames in four windows spread over 0..180 degrees (FP32 transforms,
  the production atom density per z row);
* a 1.3M-atom polymer slab of 200 x 200 x 60 A, r 0.15 A, N = 2048, fill_bkg,
  smooth 20, max_q 3.0: 8 phi-frames in four windows over 0..180 degrees, into the
  417^3 grid cropped to 295^3 (the largest per-row atom density and fixed-point range).

Bar (as tests/test_gpu_frames.py): bin counts bit-exact; I(q) within 1e-5 relative
error wherever the oracle is above 1e-6 of its maximum; the detector image bit-exact
given the same I(q).
"""

import sys

import numpy as np
import pytest

import oracle as O
from giwaxsim_amd import engine
from giwaxsim_amd.synth import graphite_slab, polymer_slab
from giwaxsim_amd.tools.comparison import detectormaker_fitting

pytestmark = pytest.mark.gpu

RTOL = 1e-5
FLOOR = 1e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(coords, elements, r, qv, mq, fill, smooth, windows, precision='auto'):
    ids, ftab, _, _ = engine.element_table(elements, 12700.0)
    g = engine.qgrid(qv, mq)
    if windows is None:
        phis = g.phis
    else:
        phis = np.concatenate([g.phis[lo:hi] for lo, hi in windows])
    res = engine.run_voxelgrid(coords, elements, r, qv, mq, 12700.0, fill_bkg=fill, smooth=smooth,
                               f_table=ftab, elem_ids=ids, frames=windows, precision=precision)
    iq = res['iq64'].cpu().numpy()
    cnt = res['bins'].cnt.view((g.c_n,) * 3).cpu().numpy()
    _say(f"  GPU done: N={res['plan'].N}, {len(phis)} frames, fp64 transforms={res['runner'].f64}")
    ref = O.voxelgridmaker(coords, ftab[ids], r, qv, mq, fill_bkg=fill, smooth=smooth, phis=phis)
    s = slice(g.c_lo, g.c_lo + g.c_n)
    want_cnt = ref['count'][s, s, s]
    assert np.array_equal(cnt, want_cnt.astype(np.int32)), f"{np.sum(cnt != want_cnt)} bins differ in count"
    want = ref['iq']
    m = want > FLOOR * want.max()
    rel = np.abs(iq[m] - want[m]) / want[m]
    _say(f"  max rel err {rel.max():.3e} (mean {rel.mean():.3e}) over {m.sum()} voxels above the floor; "
         f"{int(cnt.sum())} binned pixels")
    assert rel.max() <= RTOL
    assert np.max(np.abs(iq[~m] - want[~m])) <= RTOL * FLOOR * want.max() * 10 + 1e-300
    return res, ref


def test_graphite50_N512_all_frames_and_detector():
    coords, el = graphite_slab(size=(50.0, 50.0, 50.0), seed=0)
    r = 0.25
    qv = 2 * np.pi / (r * 512) * (1 + 1e-9)
    _say(f"[gpu-configs] graphite 50^3 A ({len(coords)} atoms) N=512 sparse, all frames")
    res, ref = _compare(coords, el, r, qv, 2.5, False, 0, None)
    assert res['plan'].N == 512 and res['plan'].n_frames == 453 and res['runner'].f64
    _say("[gpu-configs] graphite 50^3 A detector 256^2, psi 0..180 step 1 (181 orientations)")
    psis = np.linspace(0, 180, 181)
    one = np.array([0.0])
    args = (ref['qx'], ref['qy'], ref['qz'], 256, 2.5, (0, 0, 0), ('None',) * 3, psis, None, one, None,
            one, None)
    got, h, v = detectormaker_fitting(ref['iq'], *args, mirror=True)
    want, wh, wv = O.detectormaker(ref['iq'], *args, mirror=True)
    assert np.array_equal(h, wh) and np.array_equal(v, wv)
    assert np.array_equal(got, want)


def test_polymer100_N1024_production_density():
    coords, el = polymer_slab(size=(100.0, 100.0, 100.0), seed=0)
    r = 0.2
    qv = 2 * np.pi / (r * 1024) * (1 + 1e-9)
    windows = [(0, 4), (145, 149), (290, 294), (435, 439)]
    _say(f"[gpu-configs] polymer 100^3 A ({len(coords)} atoms) N=1024 fill smooth 10, frames {windows}")
    res, _ = _compare(coords, el, r, qv, 2.0, True, 10, windows)
    assert res['plan'].N == 1024 and not res['runner'].f64


def test_polymer_1p3M_atoms_N2048():
    coords, el = polymer_slab(size=(200.0, 200.0, 60.0), seed=5, density=0.54)
    r = 0.15
    qv = 2 * np.pi / (r * 2048) * (1 + 1e-9)
    windows = [(0, 2), (327, 329), (653, 655), (980, 982)]
    _say(f"[gpu-configs] polymer 200x200x60 A ({len(coords)} atoms) N=2048 fill smooth 20 max_q 3.0, "
         f"frames {windows}")
    res, _ = _compare(coords, el, r, qv, 3.0, True, 20, windows)
    assert res['plan'].N == 2048 and res['plan'].n_frames == 1307 and res['plan'].grid.c_n == 295
