"""gx_slab_crop_noalloc (the allocation-free variant of gx_slab_crop, block counts in a
library-owned device array) against gx_slab_crop with a caller-owned work buffer, through
the C ABI: kept count, coordinates, order and elements bit for bit, on two calls in a row
(the second with fewer atoms, so stale block counts of the first would show), and the
size limit rejected on the host before any launch."""

import ctypes

import numpy as np
import pytest
import torch

from giwaxsim_amd import _lib
from giwaxsim_amd.engine import ptr, stream_handle

pytestmark = pytest.mark.gpu


def _crop(n, seed, noalloc):
    rng = np.random.default_rng(seed)
    dev = torch.device('cuda')
    c = rng.uniform(-3.0, 57.0, size=(n, 3))
    x, y, z = (torch.as_tensor(np.ascontiguousarray(c[:, k])).to(dev) for k in range(3))
    el = torch.as_tensor(rng.integers(0, 5, size=n).astype(np.int32)).to(dev)
    mins = c.min(axis=0)
    ext = c.max(axis=0) - mins
    lo, hi = 0.2 * ext, 0.75 * ext
    lim = (ctypes.c_double * 9)(*mins, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
    ox, oy, oz = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3))
    oel = torch.empty(n, dtype=torch.int32, device=dev)
    kept = ctypes.c_int64(-1)
    if noalloc:
        _lib.call("gx_slab_crop_noalloc", ptr(x), ptr(y), ptr(z), ptr(el), n, lim, ptr(ox), ptr(oy), ptr(oz),
                  ptr(oel), ctypes.byref(kept), stream_handle())
    else:
        work = torch.empty((n + 2047) // 2048 + 1, dtype=torch.int64, device=dev)
        _lib.call("gx_slab_crop", ptr(x), ptr(y), ptr(z), ptr(el), n, lim, ptr(ox), ptr(oy), ptr(oz), ptr(oel),
                  ptr(work), ctypes.byref(kept), stream_handle())
    m = int(kept.value)
    return m, tuple(t[:m].cpu().numpy() for t in (ox, oy, oz, oel))


def test_noalloc_crop_matches_the_work_buffer_entry_point():
    print("[gpu-noalloc] gx_slab_crop_noalloc vs gx_slab_crop: 300000 then 20000 atoms", flush=True)
    for n, seed in ((300000, 1), (20000, 2)):
        m_ref, ref = _crop(n, seed, noalloc=False)
        m, got = _crop(n, seed, noalloc=True)
        assert m == m_ref and m > 0
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_noalloc_crop_rejects_sizes_beyond_its_array():
    big = _lib.GX_CROP_NOALLOC_MAX_ATOMS + 1
    lim = (ctypes.c_double * 9)(*([0.0] * 9))
    kept = ctypes.c_int64(0)
    with pytest.raises(_lib.GxError, match="outside"):
        # rejected by the host-side size check before anything is enqueued or read
        _lib.call("gx_slab_crop_noalloc", ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1),
                  ctypes.c_void_p(1), big, lim, ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1),
                  ctypes.c_void_p(1), ctypes.byref(kept), stream_handle())
