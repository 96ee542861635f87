"""Ensembles (many slabs, uneven atom counts and grid edges, one shared I(q)
grid) and HIP-graph replay of the whole forward pass; also with slabs whose grid edge
is not a power of two (N = 300: zoomed chirp-z transforms) beside a N = 512 slab."""

import numpy as np
import pytest
import torch

import oracle as O
from giwaxsim_amd import engine
from giwaxsim_amd.synth import ensemble_slabs

pytestmark = pytest.mark.gpu

QV = 2 * np.pi / (0.4 * 512) * (1 + 1e-9)   # ~0.0307, shared by every slab
MQ = 2.0


def _slabs(kind='pow2'):
    if kind == 'chirpz':
        return (ensemble_slabs(n_slabs=2, seed=8, size_range=(16.0, 24.0), grids=(300,), q_voxel=QV) +
                ensemble_slabs(n_slabs=1, seed=9, size_range=(16.0, 24.0), grids=(512,), q_voxel=QV))
    return ensemble_slabs(n_slabs=3, seed=7, size_range=(16.0, 28.0), grids=(256, 512), q_voxel=QV)


@pytest.mark.parametrize("kind", ['pow2', 'chirpz'])
def test_ensemble_pooled_bins_exact(kind):
    print(f"[gpu-ensemble] pooled vs per-slab ({kind})", flush=True)
    slabs = _slabs(kind)
    ens = engine.Ensemble(slabs, QV, MQ, 12700.0, fill_bkg=True, smooth=2, batch=64)
    ens.run()
    ens.check()
    pooled_s = ens.bins.sum.clone()
    pooled_c = ens.bins.cnt.clone()
    # every slab alone into bins of the same scale: the pooled result is their exact sum
    tot_s = torch.zeros_like(pooled_s)
    tot_c = torch.zeros_like(pooled_c)
    for rn in ens.runners:
        ens.bins.clear()
        rn.run()
        tot_s += ens.bins.sum
        tot_c += ens.bins.cnt
    assert torch.equal(tot_s, pooled_s) and torch.equal(tot_c, pooled_c)


@pytest.mark.parametrize("kind", ['pow2', 'chirpz'])
def test_ensemble_vs_oracle(kind):
    print(f"[gpu-ensemble] vs oracle ({kind})", flush=True)
    slabs = _slabs(kind)
    ens = engine.Ensemble(slabs, QV, MQ, 12700.0, fill_bkg=True, smooth=2, batch=64)
    iq64, _ = ens.run()
    vg = vc = None
    for (coords, elements, r), rn in zip(slabs, ens.runners):
        idx = {str(u): i for i, u in enumerate(sorted(set(np.concatenate([e for _, e, _ in slabs]))))}
        fv = ens.f_table[np.array([idx[str(e)] for e in elements])]
        ref = O.voxelgridmaker(coords, fv, r, QV, MQ, fill_bkg=True, smooth=2)
        vg = ref['sum'] if vg is None else vg + ref['sum']
        vc = ref['count'] if vc is None else vc + ref['count']
    g = ens.grid
    mean = np.divide(vg, vc, out=np.zeros_like(vg), where=vc != 0)
    small, qx, qy, qz = O.downselect_voxelgrid(mean, g.q_axis, g.q_axis, g.q_axis, MQ)
    want = O.add_f0_q_3d(small, qx, qy, qz)
    s = slice(g.c_lo, g.c_lo + g.c_n)
    assert np.array_equal(ens.bins.cnt.view(g.c_n, g.c_n, g.c_n).cpu().numpy(), vc[s, s, s])
    got = iq64.cpu().numpy()
    m = want > 1e-6 * want.max()
    rel = np.abs(got[m] - want[m]) / want[m]
    print(f"  ensemble max rel err {rel.max():.3e}", flush=True)
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("kind", ['pow2', 'chirpz'])
def test_graph_replay_matches_eager(kind):
    print(f"[gpu-ensemble] graph replay ({kind})", flush=True)
    slabs = _slabs(kind)
    ens = engine.Ensemble(slabs, QV, MQ, 12700.0, fill_bkg=True, smooth=2, batch=64)
    iq_eager = ens.run()[0].clone()
    torch.cuda.synchronize()
    ens.capture()
    for k in range(3):
        junk = torch.randn(1 << 20, device='cuda').sum()   # eager work between replays
        torch.cuda.synchronize()
        iq, _ = ens.replay()
        torch.cuda.synchronize()
        assert torch.equal(iq, iq_eager), f"replay {k} differs"
        del junk
