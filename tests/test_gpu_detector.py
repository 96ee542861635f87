"""GPU parity of the detector path (tools/detector.py, tools/comparison.py:790-870)
against the FP64 oracle.  Fed the same FP64 I(q) grid, the GPU image is
bit-identical: same rotation matrices, the same FP64 operations per pixel
(no FMA), exact floor-division indices and the same (psi, phi, theta)
summation order.  The bit-exact cases (<= 65^2 pixels, 222 orientations: six full
32-orientation chunks, in both of its LDS buffers, and a partial one) run the cooperative
gather used for small detectors with many orientations; the weights case (16 orientations) and the golden
500^2 image run the one-thread-per-pixel kernel.  A lattice case puts pixel values where the
one-multiply floor is off by one, for both kernels."""

import numpy as np
import pytest

import oracle as O
from giwaxsim_amd.tools import detector as D
from giwaxsim_amd.tools.comparison import detectormaker_fitting

pytestmark = pytest.mark.gpu


def _iq(seed=0, n=61):
    rng = np.random.default_rng(seed)
    q = np.linspace(-2.6, 2.6, n)
    iq = np.exp(rng.normal(size=(n, n, n)))
    return iq, q


@pytest.mark.parametrize("n_pix,mirror,init", [
    (64, True, ((0.15, 0, 0), ('psi', 'None', 'None'))),
    (65, False, ((90.0, 90.0, 90.0), ('psi', 'phi', 'psi'))),
    (48, True, ((0, 0, 0), ('None', 'None', 'None'))),
])
def test_detectormaker_bit_exact(n_pix, mirror, init):
    print(f"[gpu-detector] n_pix={n_pix} mirror={mirror} init={init}", flush=True)
    iq, q = _iq()
    psis = np.linspace(0, 180, 37)
    phis = np.linspace(0, 20, 3)
    thetas = np.linspace(0, 1, 2)
    got, h, v = detectormaker_fitting(iq, q, q, q, n_pix, 2.0, init[0], init[1], psis, None, phis,
                                      None, thetas, None, mirror=mirror)
    want, h2, v2 = O.detectormaker(iq, q, q, q, n_pix, 2.0, init[0], init[1], psis, None, phis,
                                   None, thetas, None, mirror=mirror)
    assert np.array_equal(h, h2) and np.array_equal(v, v2)
    assert np.array_equal(got, want), np.max(np.abs(got - want) / np.abs(want))


@pytest.mark.parametrize("n_psi", [16, 65])
def test_detector_lattice_floor_cases(n_psi):
    """Pixels and voxel edges on a 0.07 lattice (17 voxels, 49 pixels over the same range):
    h - q_min and v - q_min are near multiples of dq, and for four of the pixel values
    a * (1 / dq) rounds across an integer (floor(a * (1 / dq)) != a // dq), so those indices
    need the exact floor-division path, not the one-multiply one.  psi 0 / 90 / 180 keep the
    lattice; 16 orientations run the per-pixel kernel, 65 the cooperative one.  The image
    equals the oracle's bit for bit."""
    print(f"[gpu-detector] lattice floor cases, {n_psi} orientations", flush=True)
    q = np.arange(-8, 9) * 0.07
    mq = float(np.max(q))
    rng = np.random.default_rng(5)
    iq = np.exp(rng.normal(size=(len(q),) * 3))
    psis = np.linspace(0, 180, n_psi)
    assert 90.0 in psis or n_psi == 16
    one = np.array([0.0])
    got, h, v = detectormaker_fitting(iq, q, q, q, 49, mq, (0, 0, 0), ('None',) * 3, psis, None, one,
                                      None, one, None, mirror=False)
    want, _, _ = O.detectormaker(iq, q, q, q, 49, mq, (0, 0, 0), ('None',) * 3, psis, None, one, None,
                                 one, None, mirror=False)
    # the lattice really produces quotients whose one-multiply floor is off by one
    a = h - np.min(q)
    dq = np.diff(q)[0]
    assert np.sum(np.floor(a * (1.0 / dq)) != np.floor_divide(a, dq)) == 4
    assert np.array_equal(got, want)


def test_weights_from_files(tmp_path):
    iq, q = _iq(1)
    psis = np.linspace(75, 90, 16)
    w = np.random.default_rng(2).uniform(size=16)
    w /= w.sum()
    np.save(tmp_path / 'w.npy', w)
    one = np.array([0.0])
    got, _, _ = detectormaker_fitting(iq, q, q, q, 40, 2.0, (0, 0, 0), ('None',) * 3, psis,
                                      str(tmp_path / 'w.npy'), one, None, one, None, mirror=True)
    want, _, _ = O.detectormaker(iq, q, q, q, 40, 2.0, (0, 0, 0), ('None',) * 3, psis, w, one, None,
                                 one, None, mirror=True)
    assert np.array_equal(got, want)
    with pytest.raises(AssertionError, match='psi weights must sum to 1'):
        np.save(tmp_path / 'bad.npy', w * 2)
        detectormaker_fitting(iq, q, q, q, 40, 2.0, (0, 0, 0), ('None',) * 3, psis,
                              str(tmp_path / 'bad.npy'), one, None, one, None)


def test_rotations_and_intersect_bit_exact():
    iq, q = _iq(3)
    x, y, z, h, v = D.make_detector(2.0, 33, 2.0, 33)
    for fn, ofn, ang in ((D.rotate_about_normal, O.rotate_about_normal, 37.0),
                         (D.rotate_about_vertical, O.rotate_about_vertical, 12.5),
                         (D.rotate_about_horizontal, O.rotate_about_horizontal, 3.0)):
        got = fn(x, y, z, ang)
        want = ofn(x, y, z, ang)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        x, y, z = got
    got = D.intersect_detector(iq, q, q, q, x, y, z)
    want = O.intersect_detector(iq, q, q, q, x, y, z)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [(7, 9), (8, 8), (5, 5), (1, 4)])
def test_mirror(shape):
    q = np.random.default_rng(4).normal(size=shape)
    assert np.array_equal(D.mirror_vertical_horizontal(q), O.mirror_vertical_horizontal(q))
