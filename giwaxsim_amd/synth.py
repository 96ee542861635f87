"""Seeded synthetic slabs for the benchmark and test workloads.

* ``graphite_slab``: AB-stacked hexagonal carbon (a=2.46 A, c=6.71 A) cut to a
  box, positions jittered by N(0, jitter) from a seed.
* ``polymer_slab``: a semicrystalline conjugated-polymer-like slab of
  C/H/N/O/S/F atoms (ratio 60:40:2:2:4:2) built from crystallites of rigid-rod
  chains (15 A repeat along the backbone, pi-stacking 3.7 A, lamellar spacing
  22 A) with paracrystalline disorder g on the stacking distance and a
  Gaussian orientational spread per crystallite, embedded in an amorphous
  matrix of the same composition at the same mean density.

The reference builds slabs from files (tools/comparison.py:595-671); these
generators only provide seeded inputs of the stated shapes.
"""

from __future__ import annotations

import numpy as np

POLYMER_ELEMENTS = np.array(['C', 'H', 'N', 'O', 'S', 'F'])
POLYMER_RATIO = np.array([60, 40, 2, 2, 4, 2], dtype=np.float64)


def graphite_slab(size=(50.0, 50.0, 50.0), jitter=0.02, seed=0, a=2.46, c=6.71):
    """AB graphite cut to ``size`` (A).  Returns (coords (n,3) f64, elements)."""
    rng = np.random.default_rng(seed)
    a1 = np.array([a, 0.0, 0.0])
    a2 = np.array([a / 2.0, a * np.sqrt(3.0) / 2.0, 0.0])
    # two atoms per layer basis; layer B shifted by (a1+a2)/3
    basis_A = np.array([[0.0, 0.0, 0.0], (a1 + a2) / 3.0])
    basis_B = basis_A + (a1 + a2) / 3.0
    sx, sy, sz = size
    ni = int(np.ceil(sx / a)) + int(np.ceil(sy / (a * np.sqrt(3) / 2))) + 2
    nj = int(np.ceil(sy / (a * np.sqrt(3) / 2))) + 2
    nlay = int(np.ceil(sz / (c / 2.0))) + 1
    i, j = np.meshgrid(np.arange(-ni, ni + 1), np.arange(0, nj + 1), indexing='ij')
    lat = i.reshape(-1, 1) * a1 + j.reshape(-1, 1) * a2
    pts = []
    for L in range(nlay):
        b = basis_A if L % 2 == 0 else basis_B
        zz = L * c / 2.0
        for bb in b:
            p = lat + bb
            p = p.copy()
            p[:, 2] = zz
            pts.append(p)
    pts = np.concatenate(pts, axis=0)
    m = ((pts[:, 0] >= 0) & (pts[:, 0] < sx) & (pts[:, 1] >= 0) & (pts[:, 1] < sy) &
         (pts[:, 2] >= 0) & (pts[:, 2] < sz))
    pts = pts[m]
    pts = pts + rng.normal(0.0, jitter, size=pts.shape)
    pts -= pts.min(axis=0)
    return pts.astype(np.float64), np.full(len(pts), 'C')


def _rot_from_axis_angle(axis, ang):
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


def polymer_slab(size=(100.0, 100.0, 100.0), seed=0, density=0.11, crystal_fraction=0.6,
                 cryst_size=(10.0, 25.0), repeat=15.0, pi_stack=3.7, lamellar=22.0,
                 g=0.08, sigma_orient_deg=10.0):
    """Semicrystalline polymer slab of ``size`` A at ``density`` atoms/A^3.

    Returns (coords (n,3) f64, elements (n,) str).  About 110k atoms for the
    default 100^3 A box.
    """
    rng = np.random.default_rng(seed)
    size = np.asarray(size, dtype=np.float64)
    vol = float(np.prod(size))
    n_target = int(round(density * vol))
    probs = POLYMER_RATIO / POLYMER_RATIO.sum()
    # one repeat unit: 110 atoms over a 15 A x (lamellar ~ 8 A side chains) motif
    n_unit = 110
    unit_counts = np.round(probs * n_unit).astype(int)
    unit_counts[0] += n_unit - unit_counts.sum()
    unit_el = np.repeat(np.arange(6), unit_counts)
    # fixed motif: backbone atoms along x within a 15 A repeat, side chains along y
    mrng = np.random.default_rng(12345)
    motif = np.empty((n_unit, 3))
    nb = 40
    motif[:nb, 0] = np.linspace(0, repeat, nb, endpoint=False)
    motif[:nb, 1] = 0.7 * np.sin(2 * np.pi * motif[:nb, 0] / (repeat / 4))
    motif[:nb, 2] = 0.3 * np.cos(2 * np.pi * motif[:nb, 0] / (repeat / 4))
    ns = n_unit - nb
    motif[nb:, 0] = mrng.uniform(0, repeat, ns)
    motif[nb:, 1] = mrng.uniform(1.5, lamellar / 2 - 1.0, ns) * mrng.choice([-1, 1], ns)
    motif[nb:, 2] = mrng.normal(0, 0.4, ns)
    mrng.shuffle(unit_el)

    coords = []
    elems = []
    n_cryst_target = crystal_fraction * n_target
    n_cryst = 0
    # crystallites: boxes of edge U[cryst_size] with rods along local x,
    # pi-stacks along local z (3.7 A, paracrystalline g) and lamellae along local y
    while n_cryst < n_cryst_target:
        L = rng.uniform(cryst_size[0], cryst_size[1], size=3)
        center = rng.uniform(0, 1, 3) * size
        # orientation: edge-on (backbone in-plane, stacking in-plane) with Gaussian spread
        base = _rot_from_axis_angle(np.array([0.0, 0.0, 1.0]), rng.uniform(0, 2 * np.pi))
        tilt = _rot_from_axis_angle(rng.normal(size=3), np.deg2rad(rng.normal(0, sigma_orient_deg)))
        R = tilt @ base
        n_rep = max(1, int(L[0] // repeat))
        n_lam = max(1, int(L[1] // lamellar) + 1)
        # paracrystalline stacking: cumulative gaps with relative spread g
        n_stack = max(1, int(L[2] // pi_stack))
        gaps = pi_stack * (1.0 + g * rng.normal(size=n_stack))
        zs = np.concatenate([[0.0], np.cumsum(gaps[:-1])])
        for il in range(n_lam):
            for zk in zs:
                off = np.array([0.0, il * lamellar, zk])
                shift = rng.uniform(-0.3, 0.3)
                for ir in range(n_rep):
                    p = motif + off + np.array([ir * repeat + shift, 0.0, 0.0])
                    coords.append(p)
                    elems.append(unit_el)
        block = np.concatenate(coords[-(n_lam * len(zs) * n_rep):], axis=0)
        block -= block.mean(axis=0)
        block = block @ R.T + center
        coords[-(n_lam * len(zs) * n_rep):] = [block]
        elems[-(n_lam * len(zs) * n_rep):] = [np.concatenate(elems[-(n_lam * len(zs) * n_rep):])]
        n_cryst += len(block)
    cc = np.concatenate(coords, axis=0)
    ee = np.concatenate(elems)
    inside = np.all((cc >= 0) & (cc < size), axis=1)
    cc, ee = cc[inside], ee[inside]
    n_am = max(0, n_target - len(cc))
    am = rng.uniform(0, 1, size=(n_am, 3)) * size
    am_el = rng.choice(6, size=n_am, p=probs)
    cc = np.concatenate([cc, am], axis=0)
    ee = np.concatenate([ee, am_el])
    cc = cc - cc.min(axis=0)
    return cc.astype(np.float64), POLYMER_ELEMENTS[ee]


def ensemble_slabs(n_slabs=64, seed=0, size_range=(40.0, 160.0), grids=(256, 512, 1024, 2048),
                   q_voxel=2 * np.pi / (0.2 * 1024) * (1 + 1e-9)):
    """Slabs of the polymer generator with seeded random edge lengths per axis
    and a per-slab grid edge N drawn from ``grids``; r_voxel_size = 2 pi /
    (q_voxel N) so every slab shares one q grid.  Returns [(coords, elements, r)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_slabs):
        size = tuple(rng.uniform(size_range[0], size_range[1], 3))
        N = int(rng.choice(grids))
        # slightly above 2 pi / (q_voxel N) so ceil(2 pi / (q_voxel r)) is N, not N + 1
        r = 2 * np.pi / (q_voxel * N) * (1 + 1e-9)
        c, e = polymer_slab(size=size, seed=seed * 1000 + i)
        out.append((c, e, r))
    return out
