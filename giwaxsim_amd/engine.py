"""Host planner and device runner of the forward-scattering path.

Division of labour:

* host (numpy, FP64): geometry planning on 1-D axes and per-frame scalars,
  exactly as the reference computes them -- grid size, q axis, phi list,
  crop window (tools/comparison.py:705-731, tools/voxelgrids.py:16-48),
  per-frame trig scalars and linspace parameters (tools/voxelgrids.py:381-402),
  the branch constants of ``rectangular_collapse_lengths``
  (tools/voxelgrids.py:238-309) and the detector rotation matrices
  (tools/detector.py:33-162, tools/utilities.py:222-245).
* device (gfx950 HIP through the C ABI): every per-atom, per-voxel, per-pixel
  and per-bin operation.

Plans depend only on geometry (grid, q grid, angles, slab bounds), never on
atom positions, and are cached.
"""

from __future__ import annotations

import ctypes
import functools
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib
from ._lib import GxDetector, GxFrames, GxSlab, call
from .tools.ptable_dict import aff_dict, ptable

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the platform
    torch = None


# --------------------------------------------------------------------------------------
# device plumbing
# --------------------------------------------------------------------------------------
_DEVICE_SEEN = False
# torch's raw current-stream getter (0.3 us against 2.8 us for current_stream(), which
# builds a Stream object); the public call where a build lacks it
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None) if torch is not None else None


def device():
    """The HIP device; raises if none is visible (there is no CPU path).  Once a device
    has been seen the check is not repeated (a process does not lose its GPU)."""
    global _DEVICE_SEEN
    if not _DEVICE_SEEN:
        if torch is None or not torch.cuda.is_available():
            raise RuntimeError("giwaxsim_amd runs on a ROCm GPU (HIP device) and none is visible")
        _DEVICE_SEEN = True
    return torch.device("cuda")


def stream_handle():
    """torch's current stream on the current device, as the C ABI's gx_stream_t."""
    if _RAW_STREAM is not None:
        return ctypes.c_void_p(_RAW_STREAM(torch.cuda.current_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def to_dev(a, dtype):
    """numpy / torch -> contiguous device tensor of ``dtype``."""
    dev = device()
    if torch.is_tensor(a):
        return a.to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a)).to(device=dev, dtype=dtype).contiguous()


def _empty(n, dtype):
    return torch.empty(int(max(n, 1)), dtype=dtype, device=device())


# Device copies of constant tables (twiddles, chirp-z tables, Gaussian prefix sums,
# per-plan frame records, element f tables), keyed by what defines them, so that a
# repeated call with the same geometry uploads nothing.  Bounded; least recently
# used entries are dropped.
_DEV_CONST: dict = {}
_DEV_CONST_MAX = 64


def dev_const(key, make, dtype=None):
    """Cached device tensor of the host array ``make()`` for ``key``.  A key may
    hold an object whose identity defines the table (a cached plan): its entry is
    only reused while that same object is alive in the entry."""
    k = (key, torch.cuda.current_device())
    hit = _DEV_CONST.pop(k, None)
    if hit is not None:
        _DEV_CONST[k] = hit
        return hit[1]
    arr = np.ascontiguousarray(make())
    t = torch.as_tensor(arr).to(device=device(), dtype=dtype).contiguous()
    _DEV_CONST[k] = (key, t)
    while len(_DEV_CONST) > _DEV_CONST_MAX:
        _DEV_CONST.pop(next(iter(_DEV_CONST)))
    return t


class _Ident:
    """Hashable by identity: a dictionary key for an object that may not be
    hashable by value (a plan), kept alive by the key itself."""
    __slots__ = ("obj",)

    def __init__(self, obj):
        self.obj = obj

    def __hash__(self):
        return id(self.obj)

    def __eq__(self, other):
        return isinstance(other, _Ident) and other.obj is self.obj


# --------------------------------------------------------------------------------------
# element tables
# --------------------------------------------------------------------------------------
def element_table(elements, energy):
    """Unique elements -> (ids int32, f_table complex128, counts).

    f = Z + f1 + i f2 per element (tools/comparison.py:735-739); f1/f2 from
    the offline provider (``giwaxsim_amd.tools.f1f2``)."""
    from .tools.f1f2 import f1_f2
    elements = np.asarray(elements)
    uniq, inv, counts = np.unique(elements, return_inverse=True, return_counts=True)
    ftab = np.empty(len(uniq), dtype=np.complex128)
    for i, e in enumerate(uniq):
        f1, f2 = f1_f2(str(e), energy)
        ftab[i] = complex(f1, f2) + ptable[str(e)]
    return inv.astype(np.int32), ftab, counts.astype(np.int64), uniq


# --------------------------------------------------------------------------------------
# q-space grid (tools/comparison.py:705-731) and crop (tools/voxelgrids.py:16-48)
# --------------------------------------------------------------------------------------
@dataclass
class QGrid:
    q_voxel: float
    max_q: float
    max_q_diag: float
    q_num: int
    q_axis: np.ndarray
    phis: np.ndarray
    dq: float
    qmin: float
    qmax: float
    c_lo: int
    c_n: int

    @property
    def crop_axis(self):
        return self.q_axis[self.c_lo:self.c_lo + self.c_n]

    @property
    def linspace_params(self):
        """(start, step, stop) with q_axis[i] = i*step + start, last = stop."""
        start = self.q_axis[0]
        stop = self.q_axis[-1]
        step = (np.float64(self.max_q_diag) - np.float64(-self.max_q_diag)) / (self.q_num - 1)
        return float(start), float(step), float(stop)


@functools.lru_cache(maxsize=64)
def qgrid(q_voxel_size: float, max_q: float) -> QGrid:
    max_q_diag = np.sqrt(2) * max_q
    max_q_diag = max_q_diag + max_q_diag % q_voxel_size
    q_num = ((2 * max_q_diag / q_voxel_size) + 1).astype(int)
    if q_num % 2 == 0:
        q_num += 1
    q_axis = np.linspace(-max_q_diag, max_q_diag, q_num)
    delta_phi_rad = np.arctan(q_voxel_size / max_q_diag)
    phi_num = np.ceil(2 * np.pi / delta_phi_rad).astype(int)
    last_phi = 180 - (180 / phi_num)
    phis = np.linspace(0, last_phi, num=phi_num)
    voxel_size = np.abs(q_axis[1] - q_axis[0])
    max_val = max_q + voxel_size
    idx = np.where(np.abs(q_axis) < max_val)[0]
    return QGrid(q_voxel=q_voxel_size, max_q=max_q, max_q_diag=float(max_q_diag), q_num=int(q_num),
                 q_axis=q_axis, phis=phis, dq=float(np.diff(q_axis)[0]), qmin=float(np.min(q_axis)),
                 qmax=float(np.max(q_axis)), c_lo=int(idx[0]), c_n=int(idx[-1] - idx[0] + 1))


def check_physical(r_voxel_size, q_voxel_size, max_q, x_bound, y_bound, z_bound):
    """The two checks of tools/comparison.py:705-717; returns grid_size."""
    max_q_diag = np.sqrt(2) * max_q
    if max_q_diag > 2 * np.pi / r_voxel_size:
        raise Exception('Max_q is non-physical for given voxel size')
    grid_size = int(np.ceil(2 * np.pi / (q_voxel_size * r_voxel_size)))
    min_bound = np.min([x_bound, y_bound, z_bound])
    if grid_size * r_voxel_size < min_bound:
        raise Exception('Calculated real-space bounds smaller than simulation. '
                        'Please lower delta_q value')
    return grid_size


# --------------------------------------------------------------------------------------
# per-frame records
# --------------------------------------------------------------------------------------
def _linspace_params(start, stop, num):
    """(start, step, stop) of np.linspace(start, stop, num) (endpoint=True):
    y = arange(num)*step + start, y[-1] = stop."""
    start = np.float64(start)
    stop = np.float64(stop)
    delta = np.subtract(stop, start, dtype=np.float64)
    step = delta / (num - 1)
    if step == 0 and delta != 0:
        raise ValueError("linspace step underflows to zero")
    return start, step, stop


def collapse_record(phi, hor_length, ver_length):
    """Branch constants of rectangular_collapse_lengths (tools/voxelgrids.py:253-285),
    as (mode, hor, ver, stop1, stop2, s12, mid, A, tp, vc, tt, cp, ct)."""
    hor_length, ver_length = np.float64(hor_length), np.float64(ver_length)
    if phi > 90:
        phi = phi - 90
        hor_length, ver_length = ver_length, hor_length
    theta = 90 - phi
    theta_rad = np.deg2rad(theta)
    phi_rad = np.deg2rad(phi)
    if phi == 0:
        return (1, hor_length, ver_length) + (0.0,) * 10
    if phi == 90:
        return (2, hor_length, ver_length) + (0.0,) * 10
    stop1 = ver_length * np.cos(theta_rad)
    stop2 = hor_length * np.cos(phi_rad)
    if stop1 < stop2:
        mid = ver_length / np.sin(theta_rad)
    else:
        stop1, stop2 = stop2, stop1
        mid = hor_length / np.sin(phi_rad)
    A = np.sqrt(ver_length ** 2 - (ver_length * np.cos(theta_rad)) ** 2)
    return (0, hor_length, ver_length, stop1, stop2, stop1 + stop2, mid, A, np.tan(phi_rad),
            ver_length * np.cos(theta_rad), np.tan(theta_rad), np.cos(phi_rad), np.cos(theta_rad))


_COLLAPSE_TRIG: dict = {}
_COLLAPSE_TRIG_MAX = 16


def _collapse_trig(phis):
    """The bound-independent part of collapse_record for every phi, [F, 8]:
    (mode, swap, cos phi, cos theta, sin phi, sin theta, tan phi, tan theta) after the
    phi > 90 fold, from the same scalar calls as collapse_record.  Kept per phi list."""
    phis = np.asarray(phis, dtype=np.float64)
    key = phis.tobytes()
    hit = _COLLAPSE_TRIG.get(key)
    if hit is not None:
        return hit
    t = np.zeros((len(phis), 8), dtype=np.float64)
    for f, phi in enumerate(phis):
        swap = phi > 90
        if swap:
            phi = phi - 90
        theta_rad = np.deg2rad(90 - phi)
        phi_rad = np.deg2rad(phi)
        mode = 1 if phi == 0 else 2 if phi == 90 else 0
        t[f] = (mode, swap, np.cos(phi_rad), np.cos(theta_rad), np.sin(phi_rad), np.sin(theta_rad),
                np.tan(phi_rad), np.tan(theta_rad))
    t.setflags(write=False)
    if len(_COLLAPSE_TRIG) >= _COLLAPSE_TRIG_MAX:
        _COLLAPSE_TRIG.pop(next(iter(_COLLAPSE_TRIG)))
    _COLLAPSE_TRIG[key] = t
    return t


def collapse_records(phis, hor_length, ver_length):
    """collapse_record of every phi at once, [F, 13], bit-identical to the per-frame
    records: the trigonometry is cached per phi list, and what depends on the bounds
    is products, quotients, sums and square roots (correctly rounded elementwise) and
    squares taken as the same float64 scalar powers."""
    mode, swap, cp, ct, sp, st, tp, tt = _collapse_trig(phis).T
    hor0, ver0 = np.float64(hor_length), np.float64(ver_length)
    sw = swap != 0
    hor = np.where(sw, ver0, hor0)
    ver = np.where(sw, hor0, ver0)
    vc = ver * ct
    stop2 = hor * cp
    first = vc < stop2
    with np.errstate(divide='ignore', invalid='ignore'):
        mid = np.where(first, ver / st, hor / sp)
        ver2 = np.where(sw, hor0 ** 2, ver0 ** 2)
        vc2 = np.array([v ** 2 for v in vc], dtype=np.float64)
        A = np.sqrt(ver2 - vc2)
    s1 = np.where(first, vc, stop2)
    s2 = np.where(first, stop2, vc)
    rec = np.stack([mode, hor, ver, s1, s2, s1 + s2, mid, A, tp, vc, tt, cp, ct], axis=1)
    rec[mode != 0, 3:] = 0.0
    return rec


@dataclass
class FramePlan:
    """Atom-independent plan of the phi-frame loop of one slab."""
    N: int
    r: float
    grid: QGrid
    x_bound: float
    y_bound: float
    fparams: np.ndarray           # [F, 24]
    W: int                        # widest kept column band
    i_lo: int
    i_hi: int
    q_val: float
    max_voxels: float
    Z: int = 0                    # rows [0, Z) can hold atoms (multiple of 8)
    z_bound: float = None
    frame_cols: np.ndarray = field(repr=False, default=None)  # kept columns per frame

    @property
    def n_frames(self):
        return len(self.grid.phis)


def qgrid_from_axis(q_axis, phis, max_q=None):
    """QGrid of an explicit (qx = qy = qz) axis and phi list.  With ``max_q``
    the bins are cropped like downselect_voxelgrid, else the full grid is kept."""
    q_axis = np.asarray(q_axis, dtype=np.float64)
    q_num = len(q_axis)
    if max_q is None:
        c_lo, c_n = 0, q_num
    else:
        voxel_size = np.abs(q_axis[1] - q_axis[0])
        idx = np.where(np.abs(q_axis) < max_q + voxel_size)[0]
        c_lo, c_n = int(idx[0]), int(idx[-1] - idx[0] + 1)
    return QGrid(q_voxel=float(np.diff(q_axis)[0]), max_q=max_q, max_q_diag=float(q_axis[-1]),
                 q_num=q_num, q_axis=q_axis, phis=np.atleast_1d(np.asarray(phis, dtype=np.float64)),
                 dq=float(np.diff(q_axis)[0]), qmin=float(np.min(q_axis)), qmax=float(np.max(q_axis)),
                 c_lo=c_lo, c_n=c_n)


@functools.lru_cache(maxsize=8)
def qgrid_full(q_voxel_size: float, max_q: float) -> QGrid:
    """qgrid with every q_num^3 voxel binned (no downselect_voxelgrid crop), the grid
    generate_voxel_grid_low_mem returns (tools/voxelgrids.py:587-592, 636)."""
    g = qgrid(q_voxel_size, max_q)
    return replace(g, c_lo=0, c_n=g.q_num)


@functools.lru_cache(maxsize=32)
def frame_plan(r_voxel_size: float, q_voxel_size: float, max_q: float, x_bound: float,
               y_bound: float, z_bound: float, crop: bool = True) -> FramePlan:
    N = check_physical(r_voxel_size, q_voxel_size, max_q, x_bound, y_bound, z_bound)
    g = qgrid(q_voxel_size, max_q) if crop else qgrid_full(q_voxel_size, max_q)
    return build_frame_plan(N, r_voxel_size, g, x_bound, y_bound, z_bound)


def row_range(N, r, z_bound):
    """Rows [0, Z) that can hold atoms: the z row of an atom is
    (z - zmin) // r <= z_bound // r (floor division is monotone), rounded up to 8."""
    if z_bound is None:
        return N
    zr = int(np.floor_divide(np.float64(z_bound), np.float64(r))) + 1
    return int(min(N, max(8, -(-zr // 8) * 8)))


_FRAME_GEOM: dict = {}
_FRAME_GEOM_MAX = 16


def _frame_geometry(N, r, g: QGrid):
    """The slab-independent part of the frame records for grid edge N, voxel r and q
    grid g: per frame sin / cos, the horizontal q-axis linspaces and the kept column
    band, plus the kept detector rows.  Kept per (N, r, g) so that a fit loop, whose
    slab bounds change every evaluation, only recomputes the collapse records."""
    key = (int(N), float(r), id(g))
    hit = _FRAME_GEOM.get(key)
    if hit is not None and hit[0] is g:
        return hit[1]
    q_h = np.fft.fftfreq(N, d=r) * 2 * np.pi
    q_h_shifted = np.fft.fftshift(q_h)
    q_v_shifted = np.fft.fftshift(np.fft.fftfreq(N, d=r) * 2 * np.pi)
    Rq = np.max(q_h_shifted)
    c_lo, c_hi = g.c_lo, g.c_lo + g.c_n
    # rows (tools/voxelgrids.py:476, 499): in range and inside the crop
    vmask = (q_v_shifted <= g.qmax) & (q_v_shifted >= g.qmin)
    bz = np.floor_divide(q_v_shifted - g.qmin, g.dq).astype(int)
    rows = np.nonzero(vmask & (bz >= c_lo) & (bz < c_hi))[0]
    if len(rows) == 0 or rows[-1] - rows[0] + 1 != len(rows):
        raise RuntimeError("kept detector rows are not one contiguous band")
    F = len(g.phis)
    fp = np.zeros((F, _lib.GX_FP_STRIDE), dtype=np.float64)
    widths = np.zeros(F, dtype=np.int64)
    for f, phi in enumerate(g.phis):
        phi_rad = np.radians(phi)
        fp[f, _lib.FP_SIN] = np.sin(phi_rad)
        fp[f, _lib.FP_COS] = np.cos(phi_rad)
        right_qy = Rq * np.cos(np.deg2rad(-phi))
        left_qy = -right_qy
        right_qx = -Rq * np.sin(np.deg2rad(-phi))
        left_qx = -right_qx
        fp[f, _lib.FP_HX_START:_lib.FP_HX_STOP + 1] = _linspace_params(left_qx, right_qx, N)
        fp[f, _lib.FP_HY_START:_lib.FP_HY_STOP + 1] = _linspace_params(left_qy, right_qy, N)
        # kept columns (tools/voxelgrids.py:475, 497-498) inside the crop
        hqx = np.linspace(left_qx, right_qx, num=N)
        hqy = np.linspace(left_qy, right_qy, num=N)
        hmask = (hqx <= g.qmax) & (hqx >= g.qmin) & (hqy <= g.qmax) & (hqy >= g.qmin)
        with np.errstate(invalid='ignore'):
            bx = np.floor_divide(hqx - g.qmin, g.dq)
            by = np.floor_divide(hqy - g.qmin, g.dq)
        keep = hmask & (bx >= c_lo) & (bx < c_hi) & (by >= c_lo) & (by < c_hi)
        js = np.nonzero(keep)[0]
        if len(js):
            if js[-1] - js[0] + 1 != len(js):
                raise RuntimeError(f"kept columns of frame {f} are not one contiguous band")
            fp[f, _lib.FP_JLO], fp[f, _lib.FP_JHI] = js[0], js[-1]
            widths[f] = len(js)
        else:
            fp[f, _lib.FP_JLO], fp[f, _lib.FP_JHI] = 0, -1
    fp.setflags(write=False)
    value = (fp, int(max(1, widths.max())), int(rows[0]), int(rows[-1]))
    if len(_FRAME_GEOM) >= _FRAME_GEOM_MAX:
        _FRAME_GEOM.pop(next(iter(_FRAME_GEOM)))
    _FRAME_GEOM[key] = (g, value)
    return value


def build_frame_plan(N, r, g: QGrid, x_bound, y_bound, z_bound=None) -> FramePlan:
    fp0, W, i_lo, i_hi = _frame_geometry(N, r, g)
    fp = fp0.copy()
    max_voxels = np.sqrt(np.float64(x_bound) ** 2 + np.float64(y_bound) ** 2) // r
    fp[:, _lib.FP_LEN_MODE:_lib.FP_CT + 1] = collapse_records(g.phis, y_bound, x_bound)
    return FramePlan(N=N, r=r, grid=g, x_bound=float(x_bound), y_bound=float(y_bound), fparams=fp,
                     W=W, i_lo=i_lo, i_hi=i_hi,
                     q_val=1.0 / (N * r), max_voxels=float(max_voxels), Z=row_range(N, r, z_bound),
                     z_bound=None if z_bound is None else float(z_bound))


@functools.lru_cache(maxsize=16)
def twiddles(N: int, f64: bool = False):
    """exp(-2 pi i m / N), m < N, interleaved (re, im), computed in FP64."""
    m = np.arange(N, dtype=np.float64)
    ang = -2.0 * np.pi * m / N
    tw = np.empty(2 * N, dtype=np.float64 if f64 else np.float32)
    tw[0::2] = np.cos(ang)
    tw[1::2] = np.sin(ang)
    return tw


def is_pow2_grid(N):
    return N in (256, 512, 1024, 2048)


@functools.lru_cache(maxsize=32)
def chirpz_band_tables(N: int, k_lo: int, K: int, f64: bool = False, support: int = None):
    """Zoomed chirp-z of a length-N DFT kept at the signed frequencies k = k_lo + j,
    j < K, for inputs that vanish at n >= ``support`` (default N): the length
    M >= min(N, support) + K - 1 (a power of two, at least 1024), the kernel
    H = FFT_M(h) / M with h[t] = conj(c[k_lo + t]) = exp(+i pi ((k_lo + t)^2 mod 2N) / N)
    for the lags t in [-(L-1), K), L = min(N, M - K + 1), placed at t mod M (exact
    integer squares, FP64), and the length-M twiddles.  The tables depend on (N, k_lo,
    K, M) only and hold for any input supported on n < L: a row pass that uses them
    must see no input at n >= M - K + 1."""
    need = min(N, N if support is None else max(1, int(support)))
    M = next((m for m in (1024, 2048, 4096) if m >= need + K - 1), None)
    if M is None:
        raise NotImplementedError(f"chirp-z band of {K} frequencies at N={N} needs a transform longer than 4096")
    L = min(N, M - K + 1)
    t = np.arange(-(L - 1), K, dtype=np.int64)
    m = np.abs(k_lo + t)
    ph = np.pi * ((m * m) % (2 * N)).astype(np.float64) / N
    b = np.zeros(M, dtype=np.complex128)
    b[t % M] = np.cos(ph) + 1j * np.sin(ph)
    H = np.fft.fft(b) / M
    out = np.empty(2 * M, dtype=np.float64 if f64 else np.float32)
    out[0::2] = H.real
    out[1::2] = H.imag
    return M, out, twiddles(M, f64)


def chirpz_adjoint_tables(N: int, k_lo: int, K: int, M: int, f64: bool = False):
    """Kernel spectrum of the adjoint of the zoomed chirp-z band (the band DFT run from
    the K band values k = k_lo + j to the grid points o): with o k = (o^2 + k^2 -
    (k - o)^2) / 2 the sum over j is a convolution in the lag t = o - j of
    a_j = u_{k_j} c[|k_j|] with h[t] = conj(c[k_lo - t]) = exp(+i pi ((k_lo - t)^2 mod 2N) / N),
    t in [-(K-1), L), L = min(N, M - K + 1), placed at t mod M; H = FFT_M(h) / M (exact
    integer squares, FP64).  M is the forward band's length: the outputs o < L are exact
    (all N of them when M >= N + K - 1), the rows the forward band's input may reach."""
    if M < K:
        raise ValueError(f"chirp-z adjoint needs M >= K (M={M}, K={K})")
    L = min(N, M - K + 1)
    t = np.arange(-(K - 1), L, dtype=np.int64)
    m = np.abs(k_lo - t)
    ph = np.pi * ((m * m) % (2 * N)).astype(np.float64) / N
    b = np.zeros(M, dtype=np.complex128)
    b[t % M] = np.cos(ph) + 1j * np.sin(ph)
    H = np.fft.fft(b) / M
    out = np.empty(2 * M, dtype=np.float64 if f64 else np.float32)
    out[0::2] = H.real
    out[1::2] = H.imag
    return out


def chirpz_bands(plan):
    """(k_lo, K) of the row pass (every frame's kept columns) and of the column pass
    (kept rows [i_lo, i_hi]) as signed frequencies k = column (row) - N // 2."""
    fp = plan.fparams
    lo, hi = fp[:, _lib.FP_JLO], fp[:, _lib.FP_JHI]
    ok = hi >= lo
    if ok.any():
        jlo, jhi = int(lo[ok].min()), int(hi[ok].max())
    else:
        jlo = jhi = plan.N // 2
    h = plan.N // 2
    return (jlo - h, jhi - jlo + 1), (plan.i_lo - h, plan.i_hi - plan.i_lo + 1)


@functools.lru_cache(maxsize=16)
def chirpz_tables(N: int, f64: bool = False):
    """Chirp-z (Bluestein) tables of a length-N DFT through power-of-two
    transforms of length M >= 2N - 1: chirp c[n] = exp(-i pi (n^2 mod 2N) / N),
    kernel K = FFT_M(zero-padded conj(c) wrapped to negative lags) / M, and the
    length-M twiddles.  All computed in FP64 (n^2 mod 2N in exact integers)."""
    M = next(m for m in (1024, 2048, 4096) if m >= 2 * N - 1)
    n = np.arange(N, dtype=np.int64)
    ph = np.pi * ((n * n) % (2 * N)).astype(np.float64) / N
    c = np.cos(ph) - 1j * np.sin(ph)
    b = np.zeros(M, dtype=np.complex128)
    b[:N] = np.conj(c)
    b[M - N + 1:] = np.conj(c[1:])[::-1]
    K = np.fft.fft(b) / M
    dt = np.float64 if f64 else np.float32

    def inter(z):
        out = np.empty(2 * len(z), dtype=dt)
        out[0::2] = z.real
        out[1::2] = z.imag
        return out
    return M, inter(c), inter(K), twiddles(M, f64)


def fft_precision(fill_bkg, smooth, precision='auto'):
    """True for FP64 transforms.

    'auto': FP32 for dense grids (fill_bkg or smooth) and FP64 for sparse
    ones.  The FP32 transform error of a pixel is ~eps * sqrt(log2 N^2) *
    ||G||_2, while the smallest intensities that must be accurate are 1e-6 of
    |sum G|^2; for a sparse grid of n atoms |sum G| / ||G||_2 ~ sqrt(n), which
    leaves FP32 at ~1e-5 relative error at that floor for n ~ 1e4, whereas a
    background-filled grid has |sum G| / ||G||_2 ~ N and FP32 stays at a few
    1e-6 or below."""
    if precision == 'fp64':
        return True
    if precision == 'fp32':
        return False
    if precision != 'auto':
        raise ValueError(f"precision must be 'auto', 'fp32' or 'fp64', not {precision!r}")
    return not (fill_bkg or smooth)


def gauss_prefix(sigma):
    """Prefix sums cw[t] = sum_{d < t - R} w[d], t = 0..2R+1, of scipy's
    normalised Gaussian kernel w[-R..R] (truncate 4.0)."""
    half, R = gauss_half_kernel(sigma)
    w = np.concatenate([half[:0:-1], half])
    return np.concatenate([[0.0], np.cumsum(w)]), R


def gauss_half_kernel(sigma):
    """scipy.ndimage gaussian kernel (truncate 4.0), centre outwards:
    radius = int(4*sigma + 0.5); phi = exp(-0.5/sigma^2 x^2) / sum."""
    sd = float(sigma)
    radius = int(4.0 * sd + 0.5)
    sigma2 = sigma * sigma
    x = np.arange(-radius, radius + 1)
    phi_x = np.exp(-0.5 / sigma2 * x ** 2)
    phi_x = phi_x / phi_x.sum()
    return phi_x[radius:].copy(), radius


# --------------------------------------------------------------------------------------
# fixed-point scale of the bin sums
# --------------------------------------------------------------------------------------
def intensity_bound(sum_abs_f, N, fill_bkg, smooth, avg, max_voxels, hor, ver, r):
    """Upper bound B on |F(k)| = |sum G| <= sum |G| of one frame's grid."""
    a = abs(avg)
    per_voxel = 0.0
    if fill_bkg:
        nm_max = max_voxels + (hor + ver) / r + 2.0
        per_voxel += (nm_max + max_voxels) * a
    if smooth:
        per_voxel += a * max_voxels
    return 1.01 * (sum_abs_f + N * N * per_voxel) + 1.0


def pixels_per_bin(plan: FramePlan):
    """Max pixels of one frame that can share a bin."""
    g = plan.grid
    Rq = (plan.N // 2 - 1 if plan.N % 2 == 0 else (plan.N - 1) // 2) * plan.q_val * 2 * np.pi
    dh = 2 * Rq / (plan.N - 1)
    dv = plan.q_val * 2 * np.pi
    return (math.ceil(math.sqrt(2) * g.dq / dh) + 1) * (math.ceil(g.dq / dv) + 1)


def bin_scale_for(terms):
    """terms: iterable of (frames, pixels_per_bin, B).  Power-of-two scale with
    sum_frames k B^2 * scale <= 2^62."""
    total = sum(float(F) * float(k) * float(B) ** 2 for F, k, B in terms)
    e = math.floor(62 - math.log2(max(total, 1e-300)))
    e = max(min(e, 200), -200)
    return math.ldexp(1.0, e)


# --------------------------------------------------------------------------------------
# device runner of the phi-frame loop
# --------------------------------------------------------------------------------------
def f_table_dev(f_table):
    """Device [n_elem][2] (re, im) copy of a complex f table (cached by value)."""
    f_table = np.asarray(f_table, dtype=np.complex128)

    def make():
        ft = np.empty((len(f_table), 2), dtype=np.float64)
        ft[:, 0] = np.real(f_table)
        ft[:, 1] = np.imag(f_table)
        return ft
    return dev_const(("f_table", f_table.tobytes()), make)


class Slab:
    """Atoms of one slab on the device, regrouped by projected z row.
    ``columns``: the (x, y, z, elem) device columns when the caller already has
    them (slab_bounds_device returns them).  A slab can be pointed at new atoms of
    the same count (``bind``) and regrouped again (``sort``) without reallocating."""

    def __init__(self, coords, elem_ids, f_table, N, r, zmin, columns=None, sort=True):
        if columns is None:
            columns = device_columns(coords, elem_ids)
        n = int(columns[0].shape[0])
        self.n, self.n_elem = n, len(f_table)
        self.N, self.r = int(N), float(r)
        self.sx = _empty(n, torch.float64)
        self.sy = _empty(n, torch.float64)
        self.sq = _empty(2 * n, torch.int32)
        self.szrow = _empty(n, torch.int32)
        self.row_off = _empty(N + 2, torch.int32)
        self.row_exp = _empty(N, torch.int32)
        self.work = _empty(N * self.n_elem + 2 * (N + 1), torch.int32)
        self.desc = GxSlab(
            n_atoms=n, n_elem=self.n_elem, N=self.N, r=self.r,
            sx=ptr(self.sx).value, sy=ptr(self.sy).value, sq=ptr(self.sq).value,
            szrow=ptr(self.szrow).value, row_off=ptr(self.row_off).value,
            row_exp=ptr(self.row_exp).value, work=ptr(self.work).value)
        self.bind(columns, f_table, zmin)
        if sort:
            self.sort()

    def bind(self, columns, f_table, zmin):
        """Point the slab at the device columns (x, y, z, elem) of atoms of the same
        count, with their f table and min z."""
        x, y, z, el = columns
        if int(x.shape[0]) != self.n or len(f_table) != self.n_elem:
            raise ValueError("bind: atom count or element count differs from the slab's")
        self.x, self.y, self.z, self.elem = x, y, z, el
        self.f_table = f_table_dev(f_table)
        self.zmin = float(zmin)
        d = self.desc
        d.x, d.y, d.z, d.elem = ptr(x).value, ptr(y).value, ptr(z).value, ptr(el).value
        d.f_table, d.zmin = ptr(self.f_table).value, self.zmin

    def sort(self):
        call("gx_slab_sort_rows", ctypes.byref(self.desc), stream_handle())


def device_columns(coords, elem_ids):
    """(x, y, z, elem) contiguous device columns of host or device atoms; ``coords``
    may already be a tuple of three FP64 device columns (the slab builder's output)."""
    if isinstance(coords, tuple):
        x, y, z = (to_dev(c, torch.float64) for c in coords)
        return x, y, z, to_dev(elem_ids, torch.int32)
    x, y, z = to_dev(coords, torch.float64).t().contiguous()   # one transpose, three rows
    return x, y, z, to_dev(elem_ids, torch.int32)


def slab_bounds_device(coords, elem_ids, n_elem, with_columns=False):
    """(x_bound, y_bound, z_bound, zmin, elem_counts) of device-resident atoms, one
    read-back; with ``with_columns`` also the (x, y, z, elem) device columns."""
    x, y, z, el = device_columns(coords, elem_ids)
    out = _empty(6 + n_elem, torch.float64)   # min/max (FP64) | counts (int64 words)
    cnt = out[6:].view(torch.int64)
    d = GxSlab(n_atoms=int(x.shape[0]), n_elem=int(n_elem), x=ptr(x).value, y=ptr(y).value,
               z=ptr(z).value, elem=ptr(el).value)
    call("gx_slab_bounds", ctypes.byref(d), ptr(out), ptr(cnt), stream_handle())
    host = out.cpu()
    mm = host[:6].numpy()
    res = (mm[1] - mm[0], mm[3] - mm[2], mm[5] - mm[4], mm[4], host[6:].view(torch.int64).numpy())
    return res + ((x, y, z, el),) if with_columns else res


class Bins:
    """Cropped fixed-point bin sums + counts on the device."""

    def __init__(self, grid: QGrid, scale: float):
        self.grid = grid
        self.scale = float(scale)
        n = grid.c_n ** 3
        self.sum = _empty(n, torch.int64)
        self.cnt = _empty(n, torch.int32)
        self.clear()

    def clear(self):
        call("gx_bins_clear", ptr(self.sum), ptr(self.cnt), self.grid.c_n ** 3, stream_handle())

    def finalize(self, want64=True, want32=True, f0_element='C'):
        """(iq64, iq32) device tensors [c, c, c] indexed [qy, qx, qz], times
        (f0(|q|) / Z)^2 of ``f0_element`` (None: no f0 factor; the coefficients
        (0, ..., 0, c = 1) with Z = 1 make the factor exactly 1)."""
        g = self.grid
        n = g.c_n ** 3
        iq64 = _empty(n, torch.float64) if want64 else None
        iq32 = _empty(n, torch.float32) if want32 else None
        start, step, stop = g.linspace_params
        if f0_element is None:
            aff, z = _lib.dbl_array((0.0,) * 8 + (1.0,)), 1.0
        else:
            aff, z = _lib.dbl_array(aff_dict[f0_element]), float(ptable[f0_element])
        call("gx_bins_finalize", ptr(self.sum), ptr(self.cnt), g.c_n, self.scale, g.c_lo, g.q_num,
             start, step, stop, aff, z, ptr(iq64), ptr(iq32), stream_handle())
        shp = (g.c_n,) * 3
        return (iq64.view(shp) if iq64 is not None else None,
                iq32.view(shp) if iq32 is not None else None)


class FrameRunner:
    """Runs every phi-frame of one slab into shared bins, in batches."""

    def __init__(self, plan: FramePlan, slab: Slab, bins: Bins, fill_bkg, smooth, avg,
                 batch=None, precision='auto', variant=None, counts=None, volume=None,
                 bs_support=None):
        """``counts`` (atoms per element, all atoms) and ``volume`` (x_b y_b z_b): the
        fill and smoothing terms then use avg_voxel_f of the f values the fixed-point
        deposit adds (gx_slab_avg, computed on the device) instead of ``avg``.
        ``bs_support``: grid edges that are not powers of two only, the y rows the chirp-z
        row band is sized for (default: the slab's projected width, max_voxels + 2); a
        frame whose atoms reach beyond the tables' support stops the run with an error."""
        dev = device()
        N, W, Z = plan.N, plan.W, plan.Z
        self.plan, self.slab, self.bins = plan, slab, bins
        self.f64 = fft_precision(fill_bkg, smooth, precision)
        esz = 16 if self.f64 else 8
        F = plan.n_frames
        if batch is None:
            # keep the batch's intermediate (batch * W * Z complex) within ~160 MB
            batch = max(1, min(F, int(160e6 // (W * Z * esz))))
        self.batch = int(batch)
        self.fparams = dev_const(("fparams", _Ident(plan)), lambda: plan.fparams)
        self.tw = dev_const(("twiddles", N, self.f64), lambda: twiddles(N, self.f64))
        self.tw64 = dev_const(("twiddles", N, True), lambda: twiddles(N, True))
        if is_pow2_grid(N):
            self.bs_M, self.bs = 0, (None, None, None)
            self.bs_band = (0, 0, 0, 0, 0, None, None)
        else:
            _, ch, _, _ = chirpz_tables(N, self.f64)
            (rlo, rk), (clo, ck) = chirpz_bands(plan)
            # a row's deposit reaches y indices up to (x_b^2 + y_b^2)^(1/2) / r (the
            # projected width; max_voxels), so the row band's convolution can be shorter
            # than N + K - 1 (the kernel checks each frame's y bound against the tables)
            support = int(plan.max_voxels) + 2 if bs_support is None else int(bs_support)
            M, Kr, twr = chirpz_band_tables(N, rlo, rk, self.f64, support=support)
            # along z only the Z rows that hold atoms carry data, unless smoothing without fill
            # spreads the background over every row; the k = 0 column's fill constant on
            # the rows beyond the band's support is added in closed form by the column pass
            Mc, Kc, twc = chirpz_band_tables(N, clo, ck, self.f64,
                                             support=Z if (fill_bkg or not smooth) else None)
            self.bs_M = M
            dc = lambda key, a: dev_const(("chirpz", N, self.f64) + key, lambda: a)
            self.bs = (dc(("c",), ch), dc(("K", rlo, rk, M), Kr), dc(("tw", M), twr))
            self.bs_band = (rlo, rk, Mc, clo, ck, dc(("K", clo, ck, Mc), Kc), dc(("tw", Mc), twc))
        B = self.batch
        # per-frame statistics and column data for every frame (small: the statistics,
        # column masks and background spectra of all frames are computed in one launch
        # each); the row-transformed frames only for one batch
        self.ymin_key = _empty(3 * F, torch.int64)
        self.idx_bounds = _empty(4 * F, torch.int32)
        self.col_nm = _empty(F * N, torch.int32) if fill_bkg else None
        self.col_sy = _empty(F * N, torch.float64) if smooth else None
        self.col_sz = _empty(F * N, torch.float64) if smooth else None
        cdt = torch.float64 if self.f64 else torch.float32
        self.inter = _empty(B * W * Z * 2, cdt)
        self.col_spec = _empty(F * W * 2, cdt) if (fill_bkg or smooth) else None
        self.status = torch.zeros(4, dtype=torch.int32, device=dev)
        if smooth:
            cw, rad = gauss_prefix(smooth)
            self.gauss = dev_const(("gauss_prefix", int(smooth)), lambda: cw)
        else:
            self.gauss, rad = None, 0
        zr = int(np.floor_divide(np.float64(plan.z_bound), np.float64(plan.r))) \
            if plan.z_bound is not None else N
        zrow_max = zr if zr < N else -1
        g = plan.grid
        self.desc = GxFrames(
            N=N, r=plan.r, n_atoms=slab.n,
            ax=ptr(slab.sx).value, ay=ptr(slab.sy).value, aq=ptr(slab.sq).value,
            azrow=ptr(slab.szrow).value, row_off=ptr(slab.row_off).value,
            row_exp=ptr(slab.row_exp).value,
            fill_bkg=int(bool(fill_bkg)), smooth=int(smooth), max_voxels=plan.max_voxels,
            avg_re=float(np.real(avg)), avg_im=float(np.imag(avg)), gauss_cw=ptr(self.gauss).value,
            smooth_radius=rad, zrow_max=zrow_max, qmin=g.qmin, dq=g.dq, c_lo=g.c_lo, c_n=g.c_n, q_val=plan.q_val,
            i_lo=plan.i_lo, i_hi=plan.i_hi, W=W, Z=Z, bin_scale=bins.scale, fft_f64=int(self.f64),
            twiddle=ptr(self.tw).value, twiddle64=ptr(self.tw64).value, bs_M=self.bs_M,
            bs_chirp=ptr(self.bs[0]).value, bs_kernel=ptr(self.bs[1]).value,
            bs_twiddle=ptr(self.bs[2]).value, bs_rlo=self.bs_band[0], bs_rk=self.bs_band[1],
            bs_Mc=self.bs_band[2], bs_clo=self.bs_band[3], bs_ck=self.bs_band[4],
            bs_kernel_c=ptr(self.bs_band[5]).value, bs_twiddle_c=ptr(self.bs_band[6]).value,
            n_frames=F, fparams=ptr(self.fparams).value, batch=self.batch,
            ymin_key=ptr(self.ymin_key).value, idx_bounds=ptr(self.idx_bounds).value,
            col_nm=ptr(self.col_nm).value, col_sy=ptr(self.col_sy).value,
            col_sz=ptr(self.col_sz).value, inter=ptr(self.inter).value,
            col_spec=ptr(self.col_spec).value,
            bin_sum=ptr(bins.sum).value, bin_cnt=ptr(bins.cnt).value, status=ptr(self.status).value,
            row_variant=0)
        set_row_variant(self.desc, 0 if variant is None else int(variant))
        self.avg_q = _empty(2, torch.float64)
        if counts is not None:
            self._deposited_avg(counts, volume)
        if variant is None and is_pow2_grid(N):
            set_row_variant(self.desc, tuned_variant(self))

    def _deposited_avg(self, counts, volume):
        """Enqueue gx_slab_avg (after the slab's regrouping) into ``avg_q`` and point the
        descriptor at it."""
        counts = np.ascontiguousarray(np.asarray(counts, dtype=np.int64))
        cdev = dev_const(("elem_counts", counts.tobytes()), lambda: counts, torch.int64)
        call("gx_slab_avg", ctypes.byref(self.slab.desc), ptr(cdev), float(volume), float(self.plan.r ** 3),
             ptr(self.avg_q), stream_handle())
        self.desc.avg_q = ptr(self.avg_q).value

    def rebind(self, avg, counts=None, volume=None):
        """New avg_voxel_f and the bins' current scale, status cleared: the runner
        is ready for a slab that was re-bound to new atoms of the same geometry."""
        self.desc.avg_re, self.desc.avg_im = float(np.real(avg)), float(np.imag(avg))
        self.desc.avg_q = None
        if counts is not None:
            self._deposited_avg(counts, volume)
        self.desc.bin_scale = self.bins.scale
        self.status.zero_()

    def run(self, frames=None):
        """Enqueue frames [lo, hi) (default: all) on the current stream: the per-frame
        statistics and columns of all of them in one launch each, then the row and
        column passes batch by batch (each batch's descriptor addresses its frames'
        slots of the all-frame arrays)."""
        lo, hi = (0, self.plan.n_frames) if frames is None else frames
        n = hi - lo
        if n <= 0:
            return
        st = stream_handle()
        d = self.desc
        da = GxFrames.from_buffer_copy(d)
        da.batch = self.plan.n_frames
        call("gx_frames_stats", ctypes.byref(da), lo, n, st)
        if d.fill_bkg or d.smooth:
            call("gx_frames_columns", ctypes.byref(da), lo, n, st)
        N, W = self.plan.N, self.plan.W
        esz = 16 if self.f64 else 8
        for f0 in range(lo, hi, self.batch):
            nf = min(self.batch, hi - f0)
            k = f0 - lo
            db = GxFrames.from_buffer_copy(d)
            db.ymin_key = d.ymin_key + 8 * 3 * k
            db.idx_bounds = d.idx_bounds + 4 * 4 * k
            if d.col_nm:
                db.col_nm = d.col_nm + 4 * N * k
            if d.col_sy:
                db.col_sy = d.col_sy + 8 * N * k
                db.col_sz = d.col_sz + 8 * N * k
            if d.col_spec:
                db.col_spec = d.col_spec + esz * W * k
            call("gx_frames_rowpass", ctypes.byref(db), f0, nf, st)
            call("gx_frames_colpass", ctypes.byref(db), f0, nf, st)

    def check(self):
        s = self.status.cpu().numpy()
        if s[0]:
            raise ValueError("zero-size array to reduction operation minimum which has no identity "
                             f"({int(s[0])} phi-frame(s) without any atom inside the grid)")
        if s[1]:
            raise RuntimeError("internal: a frame's column band exceeded the planned width, or its "
                               "atoms the y rows of the chirp-z row band")


_WORKSPACES: dict = {}
_WORKSPACES_MAX = 4


def run_voxelgrid(coords, elements, r_voxel_size, q_voxel_size, max_q, energy, fill_bkg=False,
                  smooth=0, f_table=None, elem_ids=None, batch=None, want64=True, frames=None,
                  precision='auto', variant=None, f0_element='C', crop=True):
    """Full voxel-grid path of one slab (``coords``: host (n, 3), device (n, 3) or a tuple
    of device x, y, z columns).  Returns dict with device iq32/iq64,
    host axes and the objects that produced them.  ``frames=(lo, hi)``
    restricts the run to phi-frames [lo, hi) (bounded samples); a list of such
    windows bins the frames of every window into the same bins.

    The device buffers and descriptors (slab, bins, frame runner) are kept per
    (geometry plan, atom and element counts, options, device, stream) and reused
    by the next call with the same key; iq64 / iq32 are new tensors every call,
    while the returned 'slab', 'bins' and 'runner' objects are overwritten by
    that next call.  ``f0_element``: the element of the (f0/Z)^2 factor applied at the
    end (the reference's 'C'), None for none.  ``crop=False`` bins every voxel of the
    q_num^3 grid (the grid before downselect_voxelgrid) and returns the full axes."""
    if f_table is None:
        elem_ids, f_table, counts, uniq = element_table(elements, energy)
    elif not torch.is_tensor(elem_ids):
        counts = np.bincount(np.asarray(elem_ids), minlength=len(f_table))
    columns = None
    if torch.is_tensor(coords) or isinstance(coords, tuple):
        xb, yb, zb, zmin, counts, columns = slab_bounds_device(coords, elem_ids, len(f_table),
                                                              with_columns=True)
    else:
        coords = np.asarray(coords, dtype=np.float64)
        xb = np.max(coords[:, 0]) - np.min(coords[:, 0])
        yb = np.max(coords[:, 1]) - np.min(coords[:, 1])
        zb = np.max(coords[:, 2]) - np.min(coords[:, 2])
        zmin = np.min(coords[:, 2])
    plan = frame_plan(float(r_voxel_size), float(q_voxel_size), float(max_q), float(xb), float(yb),
                      float(zb), bool(crop))
    if not _lib.load().gx_supported_grid(plan.N):
        raise NotImplementedError(
            f"grid_size N={plan.N} (from 2*pi/(q_voxel_size*r_voxel_size)) is outside the "
            "supported range 16..2048")
    sum_f = np.sum(counts * f_table)
    volume = xb * yb * zb
    avg = (sum_f / volume) * r_voxel_size ** 3
    sum_abs = float(np.sum(counts * np.abs(f_table)))
    B = intensity_bound(sum_abs, plan.N, fill_bkg, smooth, avg, plan.max_voxels, yb, xb,
                        r_voxel_size)
    scale = bin_scale_for([(plan.n_frames, pixels_per_bin(plan), B)])
    if columns is None:
        columns = device_columns(coords, elem_ids)
    key = (_Ident(plan), int(columns[0].shape[0]), len(f_table), bool(fill_bkg), int(smooth),
           precision, batch, variant, torch.cuda.current_device(), stream_handle().value)
    ws = _WORKSPACES.pop(key, None)
    if ws is None:
        bins = Bins(plan.grid, scale)
        slab = Slab(None, None, f_table, plan.N, r_voxel_size, zmin, columns=columns)
        runner = FrameRunner(plan, slab, bins, fill_bkg, smooth, avg, batch=batch,
                             precision=precision, variant=variant, counts=counts, volume=volume)
    else:
        slab, bins, runner = ws
        bins.scale = float(scale)
        bins.clear()
        slab.bind(columns, f_table, zmin)
        slab.sort()
        runner.rebind(avg, counts=counts, volume=volume)
    _WORKSPACES[key] = (slab, bins, runner)
    while len(_WORKSPACES) > _WORKSPACES_MAX:
        _WORKSPACES.pop(next(iter(_WORKSPACES)))
    if frames is not None and len(frames) and not np.isscalar(frames[0]):
        for window in frames:
            runner.run(tuple(window))
    else:
        runner.run(frames)
    # the finalize is enqueued before the status read-back, so the device does not
    # wait on the host between the frames and the finalize
    iq64, iq32 = bins.finalize(want64=want64, want32=True, f0_element=f0_element)
    runner.check()
    ax = plan.grid.crop_axis
    return dict(iq64=iq64, iq32=iq32, qx=ax.copy(), qy=ax.copy(), qz=ax.copy(), plan=plan,
                bins=bins, runner=runner, slab=slab, avg=avg)


# --------------------------------------------------------------------------------------
# detector planning (tools/detector.py, tools/comparison.py:790-853)
# --------------------------------------------------------------------------------------
def rotation_matrix(u, theta):
    """Rodrigues matrix, tools/utilities.py:222-245."""
    ux, uy, uz = u[0], u[1], u[2]
    R = np.zeros((3, 3))
    c = np.cos(theta)
    s = np.sin(theta)
    R[0, 0] = c + ux ** 2 * (1 - c)
    R[0, 1] = ux * uy * (1 - c) - uz * s
    R[0, 2] = ux * uz * (1 - c) + uy * s
    R[1, 0] = uy * ux * (1 - c) + uz * s
    R[1, 1] = c + uy ** 2 * (1 - c)
    R[1, 2] = uy * uz * (1 - c) - ux * s
    R[2, 0] = uz * ux * (1 - c) - uy * s
    R[2, 1] = uz * uy * (1 - c) + ux * s
    R[2, 2] = c + uz ** 2 * (1 - c)
    return R


def _apply_host(R, pts):
    """Canonical R @ (x, y, z) on host arrays (same operations as the device)."""
    x, y, z = pts
    return (R[0, 0] * x + R[0, 1] * y + R[0, 2] * z,
            R[1, 0] * x + R[1, 1] * y + R[1, 2] * z,
            R[2, 0] * x + R[2, 1] * y + R[2, 2] * z)


def _axis_matrix(kind, corners, deg):
    """Rotation matrix of rotate_about_{normal,vertical,horizontal} from the
    detector's current corner pixels [0,0], [0,-1], [-1,0]
    (tools/detector.py:56-68, 104-110, 146-152)."""
    x, y, z = corners
    p1 = np.array([x[0], y[0], z[0]])
    p2 = np.array([x[1], y[1], z[1]])
    p3 = np.array([x[2], y[2], z[2]])
    if kind == 'psi':
        ax = np.cross(p2 - p1, p3 - p1)
    elif kind == 'phi':
        ax = p3 - p1
    elif kind == 'theta':
        ax = p2 - p1
    else:
        raise ValueError(kind)
    ax /= np.linalg.norm(ax)
    return rotation_matrix(ax, np.radians(deg))


@dataclass
class DetectorPlan:
    n_pix: int
    h: np.ndarray
    v: np.ndarray
    init_R: np.ndarray      # [n_init, 9]
    orient_R: np.ndarray    # [O, 27]
    orient_w: np.ndarray    # [O]


def detector_plan(num_pixels, max_q, angle_init_vals, angle_init_axs, psis, psi_w, phis, phi_w,
                  thetas, theta_w):
    n = int(num_pixels)
    h = np.linspace(-max_q, max_q, n)
    v = np.linspace(-max_q, max_q, n)
    corners = (np.zeros(3), np.array([h[0], h[-1], h[0]]), np.array([v[0], v[0], v[-1]]))
    init = []
    for val, ax in zip(angle_init_vals, angle_init_axs):
        # each slot applies at most one of psi/phi/theta (tools/comparison.py:799-818)
        if ax in ('psi', 'phi', 'theta'):
            R = _axis_matrix(ax, corners, val)
            corners = _apply_host(R, corners)
            init.append(R.ravel())
    mats, ws = [], []
    for ps, pw in zip(psis, psi_w):
        Rp = _axis_matrix('psi', corners, ps)
        c1 = _apply_host(Rp, corners)
        for ph, phw in zip(phis, phi_w):
            Rv = _axis_matrix('phi', c1, ph)
            c2 = _apply_host(Rv, c1)
            for th, tw in zip(thetas, theta_w):
                Rh = _axis_matrix('theta', c2, th)
                mats.append(np.concatenate([Rp.ravel(), Rv.ravel(), Rh.ravel()]))
                ws.append(pw * phw * tw)
    return DetectorPlan(n_pix=n, h=h, v=v,
                        init_R=np.array(init, dtype=np.float64).reshape(-1, 9),
                        orient_R=np.array(mats, dtype=np.float64).reshape(-1, 27),
                        orient_w=np.array(ws, dtype=np.float64))


def run_detector(iq_dev, qx, qy, qz, plan: DetectorPlan, mirror, floor_scale=True):
    """Orientation sum + finalize on the device.  ``iq_dev``: float32 or float64
    device tensor [cy, cx, cz].  Returns the FP64 device image [n, n] (mirrored when
    ``mirror``; with ``floor_scale`` NaN and <= 0 floored to 1e-6 and times 1e-6) and
    the accumulated image."""
    dev = device()
    n = plan.n_pix
    h = torch.as_tensor(plan.h).to(dev)
    v = torch.as_tensor(plan.v).to(dev)
    iR = torch.as_tensor(plan.init_R).to(dev).contiguous()
    oR = torch.as_tensor(plan.orient_R).to(dev).contiguous()
    ow = torch.as_tensor(plan.orient_w).to(dev).contiguous()
    iq_dev = iq_dev.contiguous()
    cy, cx, cz = iq_dev.shape
    det = _empty(n * n, torch.float64)
    dq = np.diff(qz)[0]
    d = GxDetector(n_pix=n, n_init=len(plan.init_R), n_orient=len(plan.orient_w),
                   h_axis=ptr(h).value, v_axis=ptr(v).value, init_R=ptr(iR).value,
                   orient_R=ptr(oR).value, orient_w=ptr(ow).value, iq=ptr(iq_dev).value,
                   iq_f64=int(iq_dev.dtype == torch.float64), cy=cy, cx=cx, cz=cz,
                   qx_min=float(np.min(qx)), qy_min=float(np.min(qy)), qz_min=float(np.min(qz)),
                   dq=float(dq), det=ptr(det).value)
    st = stream_handle()
    call("gx_detector_accumulate", ctypes.byref(d), st)
    out = _empty(n * n, torch.float64)
    flags = (_lib.GX_FIN_FLOOR_SCALE if floor_scale else 0) | (_lib.GX_FIN_MIRROR if mirror else 0)
    call("gx_detector_finalize", ptr(det), n, n, flags, ptr(out), st)
    return out.view(n, n), det.view(n, n)


# --------------------------------------------------------------------------------------
# ensembles: many slabs, uneven atom counts and grid sizes, one shared I(q) grid
# --------------------------------------------------------------------------------------
class Ensemble:
    """Slabs of one q grid (same q_voxel_size and max_q, hence the same q axis and
    phi-frames) binned into one shared cropped grid: pooled sums and counts over
    every frame of every slab (the input_folder ensemble of simulate_GIWAXS.py:84-118,
    averaged by pooling rather than per-file means).

    ``slabs`` is a list of (coords (n,3), elements (n,), r_voxel_size).  All
    geometry planning and the atom grouping happen at construction; ``run()``
    enqueues the device work of one full pass (bins cleared, every frame of
    every slab, finalize) and can be captured once in a HIP graph
    (``capture()``) and replayed (``replay()``).
    """

    def __init__(self, slabs, q_voxel_size, max_q, energy, fill_bkg=False, smooth=0,
                 precision='auto', batch=None):
        self.grid = qgrid(float(q_voxel_size), float(max_q))
        self.fill_bkg, self.smooth = fill_bkg, int(smooth)
        all_el = np.concatenate([np.asarray(e) for _, e, _ in slabs])
        _, ftab, _, uniq = element_table(all_el, energy)
        index = {str(u): i for i, u in enumerate(uniq)}
        self.f_table = ftab
        prepared, terms = [], []
        for coords, elements, r in slabs:
            coords = np.asarray(coords, dtype=np.float64)
            ids = np.array([index[str(e)] for e in elements], dtype=np.int32)
            counts = np.bincount(ids, minlength=len(ftab))
            xb, yb, zb = (np.max(coords[:, i]) - np.min(coords[:, i]) for i in range(3))
            plan = frame_plan(float(r), float(q_voxel_size), float(max_q), float(xb), float(yb),
                              float(zb))
            if not _lib.load().gx_supported_grid(plan.N):
                raise NotImplementedError(f"grid_size N={plan.N} has no compiled FFT")
            avg = (np.sum(counts * ftab) / (xb * yb * zb)) * r ** 3
            B = intensity_bound(float(np.sum(counts * np.abs(ftab))), plan.N, fill_bkg, smooth, avg,
                                plan.max_voxels, yb, xb, r)
            terms.append((plan.n_frames, pixels_per_bin(plan), B))
            prepared.append((coords, ids, plan, avg, np.min(coords[:, 2]), counts, xb * yb * zb))
        self.bins = Bins(self.grid, bin_scale_for(terms))
        self.runners = []
        for coords, ids, plan, avg, zmin, counts, vol in prepared:
            slab = Slab(coords, ids, ftab, plan.N, plan.r, zmin)
            f64 = fft_precision(fill_bkg, smooth, precision)
            self.runners.append(FrameRunner(plan, slab, self.bins, fill_bkg, smooth, avg,
                                            batch=batch if batch is not None else
                                            tuned_batch(plan.N, plan.W, plan.n_frames, f64, plan.Z),
                                            precision=precision, counts=counts, volume=vol))
        self.n_frames = sum(rn.plan.n_frames for rn in self.runners)
        self.graph = None
        self.iq64 = self.iq32 = None

    def run(self, regroup=True):
        """Enqueue one full pass; returns (iq64, iq32) device tensors."""
        self.bins.clear()
        for rn in self.runners:
            if regroup:
                call("gx_slab_sort_rows", ctypes.byref(rn.slab.desc), stream_handle())
            rn.run()
        self.iq64, self.iq32 = self.bins.finalize()
        return self.iq64, self.iq32

    def capture(self):
        """Capture run() in a HIP graph (after one eager warm-up pass)."""
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
        return self.iq64, self.iq32

    def check(self):
        for rn in self.runners:
            rn.check()


# --------------------------------------------------------------------------------------
# autotuner: frames per launch, per (grid edge, band width, device)
# --------------------------------------------------------------------------------------
_TUNED = {}


def tuned_batch(N, W, n_frames, f64=False, Z=None):
    """Frames per launch for a grid edge: the best of a few candidates timed once on
    a synthetic slab, cached per (N, W bucket, device name).  Falls back to the
    memory rule (intermediate <= ~160 MB) when tuning is disabled
    (GIWAXSIM_AUTOTUNE=0)."""
    import os
    esz = 16 if f64 else 8
    cap = max(1, int(160e6 // (W * (Z or N) * esz)))
    if os.environ.get("GIWAXSIM_AUTOTUNE", "1") == "0":
        return min(n_frames, cap)
    key = (N, (W + 63) // 64, bool(f64), torch.cuda.get_device_name())
    if key not in _TUNED:
        _TUNED[key] = _tune(N, cap, f64)
    return max(1, min(n_frames, _TUNED[key]))


_TUNED_VARIANT = {}
_SCRATCH_BINS = {}


# 0 / 4: row-major intermediate (per-wave rows, no staging); 8 / 12: the column-major
# intermediate staged per block; 1 / 5: block-wide deposit over 4 rows (column-major).
# Block-wide deposit over 8 rows (2 / 6) is never faster and does not fit N = 2048.
ROW_VARIANTS = (0, 4, 8, 12, 1, 5)


# radix plans of the wave transforms (gx_frames.fft_plan_row / fft_plan_col): N = 1024 has
# 16.16.4, 16.8.8, 8.16.8 and 8.8.16; N = 2048 has 16.16.8, 16.8.16 and 8.16.16; other edges one
FFT_PLANS = {1024: 4, 2048: 3}


def set_row_variant(desc, v):
    """Frame-pass shape ``v`` on a frame descriptor: bits 0-1 and 2 are gx_frames.row_variant
    (0 one wave per row, 1 / 2 block-wide deposit over 4 / 8 rows; + 4 XCD-aware order);
    with shape 0 on N = 512, 1024, 2048 the row-major intermediate (inter_rm) is used
    unless bit 3 asks for the column-major one; bits 4-5 and 6-7 are the row and column
    passes' radix plans (0 where the grid edge has one plan)."""
    v = int(v)
    desc.row_variant = v & 7
    desc.inter_rm = int(desc.N in (512, 1024, 2048) and desc.bs_M == 0 and (v & 3) == 0 and not (v & 8))
    n_plans = FFT_PLANS.get(int(desc.N), 1) if desc.bs_M == 0 else 1
    pr, pc = (v >> 4) & 3, (v >> 6) & 3
    desc.fft_plan_row = pr if pr < n_plans else 0
    desc.fft_plan_col = pc if pc < n_plans else 0


def tuned_variant(runner: FrameRunner):
    """Frame-pass shape (ROW_VARIANTS, see set_row_variant: the row-major intermediate
    with per-wave rows, or the column-major one with per-wave rows or a block-wide
    deposit; each in frame-major or XCD-aware tile order) and radix plans for this
    runner's grid edge, precision, band width, atoms per row and atom count (whether the
    atoms fit one XCD's L2).  The shapes are timed with plan 0, then for each of the two
    best shapes the row pass's plans and then the column pass's, each candidate five times
    (best of five: the column plans differ by a few per cent of a batch's time, less than
    one timing's spread) on the runner's own first frames (into scratch bins, so the real bins
    stay untouched); the fastest is cached.  All shapes give identical bins, the plans agree
    to rounding.  GIWAXSIM_AUTOTUNE=0 keeps shape 0 and plan 0."""
    import os
    if os.environ.get("GIWAXSIM_AUTOTUNE", "1") == "0":
        return 0
    p, d = runner.plan, runner.desc
    n_rowed = max(1, runner.slab.n)
    per_row = int(np.log2(max(1.0, n_rowed / max(1, p.Z))))
    key = (p.N, (p.W + 63) // 64, runner.f64, per_row, int(np.log2(n_rowed)),
           torch.cuda.get_device_name())
    if key in _TUNED_VARIANT:
        return _TUNED_VARIANT[key]
    g = p.grid
    sb = _SCRATCH_BINS.get(g.c_n)
    if sb is None:
        sb = _SCRATCH_BINS[g.c_n] = Bins(g, 1.0)
    keep = (d.bin_sum, d.bin_cnt, d.row_variant, d.inter_rm, d.fft_plan_row, d.fft_plan_col)
    d.bin_sum, d.bin_cnt = ptr(sb.sum).value, ptr(sb.cnt).value
    nfr = min(p.n_frames, runner.batch)

    def timed(v):
        set_row_variant(d, v)
        try:
            runner.run((0, nfr))
        except _lib.GxError:
            return float("inf")  # shape does not fit this grid (LDS)
        best_t = float("inf")
        for _ in range(5):
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            runner.run((0, nfr))
            e1.record()
            torch.cuda.synchronize()
            best_t = min(best_t, e0.elapsed_time(e1))
        return best_t

    try:
        times = {v: timed(v) for v in ROW_VARIANTS}
        n_plans = FFT_PLANS.get(p.N, 1) if runner.bs_M == 0 else 1
        # the plans move the shapes by different amounts (the row-major shape gains more from
        # a faster row transform), so the two best shapes each get a plan search
        for start in sorted(times, key=times.get)[:2 if n_plans > 1 else 1]:
            cur = start
            for shift in (4, 6):
                for pl in range(1, n_plans):
                    v = (cur & ~(3 << shift)) | (pl << shift)
                    times[v] = timed(v)
                cur = min((v for v in times if (v & 15) == (start & 15)), key=times.get)
        best = min(times, key=times.get)
    finally:
        d.bin_sum, d.bin_cnt, d.row_variant, d.inter_rm, d.fft_plan_row, d.fft_plan_col = keep
    _TUNED_VARIANT[key] = best
    return best


def _tune(N, cap, f64):
    from .synth import polymer_slab
    r = 0.2 * 1024 / N
    qv = 2 * np.pi / (r * N) * (1 + 1e-9)
    coords, elements = polymer_slab(size=(40.0, 40.0, 40.0), seed=123)
    cands = sorted({c for c in (16, 32, 64, 128, 256) if c <= cap} | {min(cap, 256)})
    ids, ftab, counts, _ = element_table(elements, 12700.0)
    xb, yb, zb = (np.max(coords[:, i]) - np.min(coords[:, i]) for i in range(3))
    plan = frame_plan(r, qv, 2.0, float(xb), float(yb), float(zb))
    avg = (np.sum(counts * ftab) / (xb * yb * zb)) * r ** 3
    bins = Bins(plan.grid, 1.0)
    slab = Slab(coords, ids, ftab, plan.N, r, np.min(coords[:, 2]))
    best, best_t = cands[0], float("inf")
    for b in cands:
        rn = FrameRunner(plan, slab, bins, True, 10, avg, batch=b,
                         precision='fp64' if f64 else 'fp32', variant=0)
        nfr = min(plan.n_frames, 4 * b)
        rn.run((0, nfr))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rn.run((0, nfr))
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / nfr
        if t < best_t:
            best, best_t = b, t
    return best
