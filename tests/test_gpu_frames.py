"""GPU parity of the phi-frame path against the FP64 oracle.

Bar: per-voxel bin COUNTS bit-exact (they are fixed entirely by the FP64
index arithmetic); per-frame atom index bounds bit-exact; I(q) (mean, cropped,
f0-scaled) within 1e-5 relative error wherever the oracle value is above
1e-6 * max; bin sums bit-identical run to run.
"""

import sys

import numpy as np
import pytest

import oracle as O
from giwaxsim_amd import engine
from giwaxsim_amd.synth import graphite_slab, polymer_slab

pytestmark = pytest.mark.gpu

RTOL = 1e-5       # max relative error of I(q) above the floor
FLOOR = 1e-6      # ... relative to max I(q)


def _case(name):
    if name == 'graphite256':
        c, e = graphite_slab(size=(16.0, 16.0, 16.0), seed=1)
        return c, e, dict(r=0.25, qv=2 * np.pi / (0.25 * 256) * (1 + 1e-9), mq=2.5, fill=False, smooth=0)
    if name == 'polymer256_fill_smooth':
        c, e = polymer_slab(size=(24.0, 24.0, 24.0), seed=2)
        return c, e, dict(r=0.2, qv=2 * np.pi / (0.2 * 256) * (1 + 1e-9), mq=2.0, fill=True, smooth=3)
    if name == 'graphite512_fill':
        c, e = graphite_slab(size=(30.0, 30.0, 20.0), seed=3)
        return c, e, dict(r=0.25, qv=2 * np.pi / (0.25 * 512) * (1 + 1e-9), mq=2.5, fill=True, smooth=0)
    if name == 'polymer1024_smooth':
        c, e = polymer_slab(size=(40.0, 40.0, 30.0), seed=4)
        return c, e, dict(r=0.2, qv=2 * np.pi / (0.2 * 1024) * (1 + 1e-9), mq=2.0, fill=True, smooth=10)
    if name == 'polymer2048_fill_smooth':
        c, e = polymer_slab(size=(30.0, 30.0, 20.0), seed=8)
        return c, e, dict(r=0.1, qv=2 * np.pi / (0.1 * 2048) * (1 + 1e-9), mq=3.0, fill=True, smooth=20)
    if name == 'graphite300':          # not a power of two: chirp-z transforms, FP64 (sparse)
        c, e = graphite_slab(size=(16.0, 16.0, 16.0), seed=5)
        return c, e, dict(r=0.25, qv=2 * np.pi / (0.25 * 300) * (1 + 1e-9), mq=2.5, fill=False, smooth=0)
    if name == 'polymer700_fill':      # chirp-z, FP32 (dense)
        c, e = polymer_slab(size=(30.0, 30.0, 24.0), seed=6)
        return c, e, dict(r=0.25, qv=2 * np.pi / (0.25 * 700) * (1 + 1e-9), mq=2.0, fill=True, smooth=0)
    if name == 'polymer1048_fill_smooth':  # the reference's example settings (N = 1048 = 8 x 131)
        c, e = polymer_slab(size=(40.0, 30.0, 40.0), seed=7)
        return c, e, dict(r=0.3, qv=0.02, mq=2.0, fill=True, smooth=5)
    if name == 'graphite1048':         # N = 1048 without fill or smoothing: FP64, both bands short, no tail
        c, e = graphite_slab(size=(30.0, 30.0, 20.0), seed=9)
        return c, e, dict(r=0.3, qv=0.02, mq=2.0, fill=False, smooth=0)
    if name == 'polymer1048_smooth':   # smoothing without fill: the column band keeps every row
        c, e = polymer_slab(size=(40.0, 30.0, 40.0), seed=10)
        return c, e, dict(r=0.3, qv=0.02, mq=2.0, fill=False, smooth=5)
    raise KeyError(name)


def _run_both(name, frames=None):
    print(f"[gpu-frames] case={name} frames={frames}", flush=True)
    sys.stdout.flush()
    coords, elements, p = _case(name)
    ids, ftab, counts, _ = engine.element_table(elements, 12700.0)
    fvals = ftab[ids]
    g = engine.qgrid(p['qv'], p['mq'])
    phis = g.phis if frames is None else g.phis[frames[0]:frames[1]]
    ref = O.voxelgridmaker(coords, fvals, p['r'], p['qv'], p['mq'], fill_bkg=p['fill'],
                           smooth=p['smooth'], phis=phis)
    res = engine.run_voxelgrid(coords, elements, p['r'], p['qv'], p['mq'], 12700.0,
                               fill_bkg=p['fill'], smooth=p['smooth'], f_table=ftab, elem_ids=ids,
                               frames=frames)
    return coords, p, ref, res


def _crop(grid, g):
    s = slice(g.c_lo, g.c_lo + g.c_n)
    return grid[s, s, s]


def _check(ref, res):
    g = res['plan'].grid
    cnt = res['bins'].cnt.view(g.c_n, g.c_n, g.c_n).cpu().numpy()
    want_cnt = _crop(ref['count'], g)
    assert np.array_equal(cnt, want_cnt.astype(np.int32)), \
        f"{np.sum(cnt != want_cnt)} bins with different counts"
    iq = res['iq64'].cpu().numpy()
    want = ref['iq']
    assert iq.shape == want.shape
    assert np.array_equal(res['qx'], ref['qx'])
    m = want > FLOOR * want.max()
    rel = np.abs(iq[m] - want[m]) / want[m]
    print(f"  max rel err {rel.max():.3e} over {m.sum()} voxels; mean {rel.mean():.3e}", flush=True)
    assert rel.max() <= RTOL
    # below the floor the absolute error is bounded by the floor itself
    assert np.max(np.abs(iq[~m] - want[~m])) <= RTOL * FLOOR * want.max() * 10 + 1e-300
    return rel.max()


@pytest.mark.parametrize("name", ['graphite256', 'polymer256_fill_smooth'])
def test_full_small(name):
    _, _, ref, res = _run_both(name)
    _check(ref, res)


@pytest.mark.parametrize("name,frames", [('graphite512_fill', (0, 48)),
                                         ('polymer1024_smooth', (100, 116)),
                                         ('graphite300', (0, 40)),
                                         ('polymer700_fill', (30, 50)),
                                         ('polymer1048_fill_smooth', (200, 210)),
                                         ('graphite1048', (300, 306)),
                                         ('polymer1048_smooth', (500, 506))])
def test_frame_subsets(name, frames):
    _, _, ref, res = _run_both(name, frames)
    if name == 'graphite1048':
        d = res['runner'].desc
        assert res['runner'].f64 and d.bs_M == 1024 and d.bs_Mc == 1024 and d.bs_Mc < 1048 + d.bs_ck - 1
    if name == 'polymer1048_smooth':
        d = res['runner'].desc
        assert d.bs_M == 1024 and d.bs_Mc == 2048
    if name == 'polymer1048_fill_smooth':
        # the row band's convolution sized to the slab's projected width (166 rows):
        # length 1024 instead of the 2048 that N + K - 1 = 1334 would need; the column
        # band sized to the Z = 136 rows that hold atoms (the k = 0 column's fill
        # constant beyond it added in closed form): 1024 instead of 2048
        assert res['runner'].bs_M == 1024 and res['runner'].desc.bs_Mc == 1024
    _check(ref, res)


def test_chirpz_row_band_too_short_is_an_error():
    """A chirp-z row band shorter than the rows a frame's atoms reach is refused per
    frame (status, then RuntimeError), never binned with aliased rows: a 250 A long slab
    at N = 1048 needs length 2048 at phi near 90 degrees; forced to 1024 (support 100
    rows, so the tables hold rows < 1024 - K + 1) the run stops with the error."""
    coords, el = graphite_slab(size=(250.0, 10.0, 10.0), seed=11)
    ids, ftab, _, _ = engine.element_table(el, 12700.0)
    g = engine.qgrid(0.02, 2.0)
    mid = len(g.phis) // 2
    print(f"[gpu-frames] chirp-z support check, {len(coords)} atoms, frames {mid - 4}..{mid + 4}", flush=True)
    res = engine.run_voxelgrid(coords, el, 0.3, 0.02, 2.0, 12700.0, f_table=ftab, elem_ids=ids,
                               frames=(mid - 4, mid + 4))
    plan = res['plan']
    assert plan.N == 1048 and res['runner'].bs_M == 2048 and plan.max_voxels > 1024 - 287
    short = engine.FrameRunner(plan, res['slab'], res['bins'], False, 0, res['avg'], bs_support=100)
    assert short.bs_M == 1024
    short.run((mid - 4, mid + 4))
    with pytest.raises(RuntimeError, match="chirp-z row band"):
        short.check()
    # frames near phi = 0 project the slab onto its 10 A width and pass
    short.status.zero_()
    short.run((0, 4))
    short.check()


def test_frame_stats_bit_exact():
    coords, elements, p = _case('polymer256_fill_smooth')
    ids, ftab, counts, _ = engine.element_table(elements, 12700.0)
    res = engine.run_voxelgrid(coords, elements, p['r'], p['qv'], p['mq'], 12700.0, fill_bkg=True,
                               smooth=3, f_table=ftab, elem_ids=ids, frames=(0, 8), batch=8)
    runner = res['runner']
    b = runner.idx_bounds.cpu().numpy().reshape(-1, 4)[:8]
    g = res['plan'].grid
    fvals = ftab[ids]
    for f in range(8):
        _, st = O.frame_grid(coords, fvals, g.phis[f], res['plan'].N, p['r'], 0.0,
                             res['plan'].x_bound, res['plan'].y_bound, False, 0, return_stats=True)
        assert tuple(b[f]) == (st['y_min'], st['y_max'], st['z_min'], st['z_max'])


def test_row_grouping():
    coords, elements, p = _case('graphite256')
    ids, ftab, counts, _ = engine.element_table(elements, 12700.0)
    res = engine.run_voxelgrid(coords, elements, p['r'], p['qv'], p['mq'], 12700.0, f_table=ftab,
                               elem_ids=ids, frames=(0, 1))
    slab = res['slab']
    N = res['plan'].N
    off = slab.row_off.cpu().numpy()
    zi = O.floor_div(coords[:, 2] - coords[:, 2].min(), p['r']).astype(int)
    hist = np.bincount(np.minimum(zi, N), minlength=N + 1)
    assert np.array_equal(np.diff(off[:N + 2]), hist)
    rows = slab.szrow.cpu().numpy()
    assert np.array_equal(np.sort(rows), np.sort(np.minimum(zi, N)))


@pytest.mark.parametrize("name,frames", [('polymer256_fill_smooth', None), ('graphite512_fill', (0, 24)),
                                         ('polymer1024_smooth', (100, 108))])
def test_row_pass_shapes_identical(name, frames):
    """The row-pass shapes the autotuner chooses between give identical bins."""
    coords, elements, p = _case(name)
    ids, ftab, counts, _ = engine.element_table(elements, 12700.0)
    sums = []
    for v in engine.ROW_VARIANTS:
        print(f"[gpu-frames] case={name} row_variant={v}", flush=True)
        res = engine.run_voxelgrid(coords, elements, p['r'], p['qv'], p['mq'], 12700.0,
                                   fill_bkg=p['fill'], smooth=p['smooth'], f_table=ftab, elem_ids=ids,
                                   frames=frames, variant=v)
        sums.append((res['bins'].sum.cpu().numpy().copy(), res['bins'].cnt.cpu().numpy().copy()))
    for s_, c_ in sums[1:]:
        assert np.array_equal(s_, sums[0][0]) and np.array_equal(c_, sums[0][1])


_REFS = {}


_PLAN_CASES = [(name, frames, shape, pr, pc)
               for name, frames, n_plans in (('polymer1024_smooth', (300, 304), 4),
                                             ('polymer2048_fill_smooth', (500, 503), 3))
               for shape, pr, pc in ((0, 1, 2), (0, 2, 1), (8, 1, 1), (8, 2, 2), (1, 2, 1), (0, 3, 3),
                                     (8, 3, 0))
               if pr < n_plans and pc < n_plans]


@pytest.mark.parametrize("name,frames,shape,plan_row,plan_col", _PLAN_CASES)
def test_radix_plans_match_oracle(name, frames, shape, plan_row, plan_col):
    """Every radix plan of the N = 1024 / 2048 wave transforms, in the row-major and
    column-major frame-pass shapes, meets the oracle bar (the plans differ from each
    other only by rounding)."""
    coords, elements, p = _case(name)
    ids, ftab, counts, _ = engine.element_table(elements, 12700.0)
    if name not in _REFS:
        g = engine.qgrid(p['qv'], p['mq'])
        _REFS[name] = O.voxelgridmaker(coords, ftab[ids], p['r'], p['qv'], p['mq'], fill_bkg=p['fill'],
                                       smooth=p['smooth'], phis=g.phis[frames[0]:frames[1]])
    v = shape | (plan_row << 4) | (plan_col << 6)
    print(f"[gpu-frames] case={name} frames={frames} shape={shape} plans=({plan_row}, {plan_col})",
          flush=True)
    res = engine.run_voxelgrid(coords, elements, p['r'], p['qv'], p['mq'], 12700.0,
                               fill_bkg=p['fill'], smooth=p['smooth'], f_table=ftab, elem_ids=ids,
                               frames=frames, variant=v)
    d = res['runner'].desc
    assert (d.fft_plan_row, d.fft_plan_col) == (plan_row, plan_col)
    assert res['plan'].N == int(name[7:11]) and not res['runner'].f64
    _check(_REFS[name], res)


def test_stable_regrouping():
    """With the permutation and scratch requested, atoms keep their original order
    within each row (the differentiable path relies on it for reproducible sums)."""
    import torch
    from giwaxsim_amd import _lib
    from giwaxsim_amd._lib import GxSlab, call
    coords, elements, p = _case('polymer256_fill_smooth')
    ids, ftab, counts, _ = engine.element_table(elements, 12700.0)
    N, r = 256, p['r']
    zmin = float(coords[:, 2].min())
    dev = torch.device('cuda')
    c = torch.as_tensor(coords, device=dev)
    x, y, z = (c[:, k].contiguous() for k in range(3))
    el = torch.as_tensor(ids, device=dev).to(torch.int32)
    ft = torch.as_tensor(np.stack([ftab.real, ftab.imag], 1), device=dev).contiguous()
    n = len(coords)
    buf = {k: torch.empty(n, dtype=dt, device=dev) for k, dt in
           (('sx', torch.float64), ('sy', torch.float64), ('szrow', torch.int32), ('szf', torch.float64),
            ('sperm', torch.int32), ('selem', torch.int32))}
    sq = torch.empty(2 * n, dtype=torch.int32, device=dev)
    row_off = torch.empty(N + 2, dtype=torch.int32, device=dev)
    row_exp = torch.empty(N, dtype=torch.int32, device=dev)
    work = torch.empty(N * len(ftab) + 2 * (N + 1), dtype=torch.int32, device=dev)
    stable = torch.empty(-(-n // _lib.GX_STABLE_CHUNK) * (N + 1), dtype=torch.int32, device=dev)
    ptr = engine.ptr
    d = GxSlab(n_atoms=n, n_elem=len(ftab), N=N, r=r, zmin=zmin, x=ptr(x).value, y=ptr(y).value,
               z=ptr(z).value, elem=ptr(el).value, f_table=ptr(ft).value, sx=ptr(buf['sx']).value,
               sy=ptr(buf['sy']).value, sq=ptr(sq).value, szrow=ptr(buf['szrow']).value,
               row_off=ptr(row_off).value, row_exp=ptr(row_exp).value, work=ptr(work).value,
               szf=ptr(buf['szf']).value, sperm=ptr(buf['sperm']).value, selem=ptr(buf['selem']).value,
               stable_work=ptr(stable).value)
    call("gx_slab_sort_rows", ctypes_byref(d), engine.stream_handle())
    perm = buf['sperm'].cpu().numpy()
    zi = O.floor_div(coords[:, 2] - zmin, r).astype(int)
    rows = np.minimum(zi, N)
    want = np.lexsort((np.arange(n), rows))  # by row, then original index
    assert np.array_equal(perm, want)
    assert np.array_equal(buf['sx'].cpu().numpy(), coords[want, 0])
    tz = (coords[want, 2] - zmin) * (1.0 / r) - zi[want]
    assert np.allclose(buf['szf'].cpu().numpy(), np.clip(tz, 0, 1), rtol=0, atol=1e-12)
    assert np.array_equal(buf['selem'].cpu().numpy(), ids[want])


def ctypes_byref(d):
    import ctypes
    return ctypes.byref(d)


def test_deterministic():
    coords, elements, p = _case('polymer256_fill_smooth')
    ids, ftab, counts, _ = engine.element_table(elements, 12700.0)
    sums = []
    for _ in range(2):
        res = engine.run_voxelgrid(coords, elements, p['r'], p['qv'], p['mq'], 12700.0,
                                   fill_bkg=True, smooth=3, f_table=ftab, elem_ids=ids)
        sums.append(res['bins'].sum.cpu().numpy().copy())
    assert np.array_equal(sums[0], sums[1])
