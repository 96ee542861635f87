"""CPU tests of host helpers of giwaxsim_amd.tools.utilities that the simulation path
does not use but a caller of the reference's tools/utilities.py may import, by
closed forms:

* gaussian_kernel (tools/utilities.py:247-258): (2 half + 1)^3 samples, centre value
  1 / (sigma^3 (2 pi)^(3/2)), radially symmetric, and the samples of a wide kernel sum
  to 1 (the integer lattice sum of a unit-integral Gaussian with sigma >= 1);
* fft_gaussian (tools/utilities.py:260-274): 1 at q = 0, exp(-q^2 sigma^2 / 2) along
  each axis (the transform of a unit-integral Gaussian), separable, [qy, qx, qz] order;
* create_shared_array (tools/utilities.py:378-395): a fresh block of prod(shape)
  float64 values that a second handle of the same name sees;
* the saved-frame binning's host checks of frame and grid shapes.
"""

import numpy as np
import pytest

from giwaxsim_amd.tools.utilities import create_shared_array, fft_gaussian, gaussian_kernel


def test_gaussian_kernel_closed_form():
    g, half = gaussian_kernel(9, sigma=1.5)
    assert half == 4 and g.shape == (9, 9, 9)
    assert g[4, 4, 4] == 1 / (1.5 ** 3 * (2 * np.pi) ** 1.5)
    assert g[4, 4, 4 + 2] == g[4 + 2, 4, 4] == g[4, 4 - 2, 4]
    assert np.isclose(g[4, 4 + 1, 4 + 1], g[4, 4, 4] * np.exp(-2 / (2 * 1.5 ** 2)), rtol=1e-15)
    wide, _ = gaussian_kernel(31, sigma=2.0)
    assert abs(wide.sum() - 1) < 1e-12
    assert gaussian_kernel(4)[0].shape == (5, 5, 5)


def test_fft_gaussian_closed_form():
    qx = np.linspace(-1, 1, 7)
    qy = np.linspace(-2, 2, 5)
    qz = np.linspace(0, 3, 4)
    sigma = 1.3
    g = fft_gaussian(qx, qy, qz, sigma)
    assert g.shape == (5, 7, 4)
    assert g[2, 3, 0] == 1.0
    f = lambda q: np.exp(-q ** 2 * sigma ** 2 / 2)
    want = f(qy)[:, None, None] * f(qx)[None, :, None] * f(qz)[None, None, :]
    assert np.allclose(g, want, rtol=1e-13, atol=0)


def test_create_shared_array():
    from multiprocessing import shared_memory
    shm = create_shared_array((3, 4, 5), 'ignored')
    try:
        assert shm.size >= 3 * 4 * 5 * 8 and len(shm.name) == 29
        a = np.ndarray((3, 4, 5), dtype=np.float64, buffer=shm.buf)
        a[:] = np.arange(60).reshape(3, 4, 5)
        other = shared_memory.SharedMemory(name=shm.name)
        b = np.ndarray((3, 4, 5), dtype=np.float64, buffer=other.buf)
        assert np.array_equal(b, np.arange(60).reshape(3, 4, 5))
        del a, b
        other.close()
        second = create_shared_array((2,))
        assert second.name != shm.name
        second.close()
        second.unlink()
    finally:
        shm.close()
        shm.unlink()


def test_frame_binning_rejects_mismatched_shapes():
    """The saved-frame binning (giwaxsim_amd.tools.voxelgrids, behind process_file,
    process_file2 and frames_to_iq_parallel) checks the frame's axes against its image
    and the grids against the q axis on the host, before any device call."""
    import torch
    from giwaxsim_amd.tools import voxelgrids as V
    q = np.linspace(-1, 1, 5)
    g = torch.zeros(125, dtype=torch.float64)
    c = torch.zeros(125, dtype=torch.float64)
    with pytest.raises(ValueError, match='do not match'):
        V._bin_frame_dev(np.zeros((4, 6)), np.zeros(5), np.zeros(6), np.zeros(4), q, q, q, g, c, None)
    with pytest.raises(ValueError, match='2-D'):
        V._bin_frame_dev(np.zeros(6), np.zeros(6), np.zeros(6), np.zeros(4), q, q, q, g, c, None)
    with pytest.raises(ValueError, match='len\\(qx\\)\\^3'):
        V._bin_frame_dev(np.zeros((4, 6)), np.zeros(6), np.zeros(6), np.zeros(4), q, q, q, g[:64], c, None)
