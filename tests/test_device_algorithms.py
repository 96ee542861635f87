"""CPU checks of the arithmetic choices the kernels make where they deviate
from a literal restatement (the device code is mirrored here line by line in
numpy and compared against the oracle / library):

* exact floor division via floor(a * (1/b)) plus one fma-signed correction
  step (gx::floor_div) vs numpy's fmod-based ``//``;
* smoothed masks from prefix sums over runs of the wrapped window
  (smooth_mask) vs scipy.ndimage.gaussian_filter1d(mode='wrap');
* the linear split of the fill / blend terms into a per-frame background
  spectrum (k_frames_spectra + k_rowpass step 5) vs the FFT of the full grid.
"""

from fractions import Fraction

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter1d

import oracle as O
from giwaxsim_amd import engine


def fma_sign(q, b, a):
    """sign of the exact a - q*b (what the device obtains from one fma)."""
    return np.sign(float(Fraction(a) - Fraction(q) * Fraction(b)))


def dev_floor_div(a, b):
    q = np.floor(a * (1.0 / b))
    if fma_sign(q, b, a) < 0:
        q -= 1.0
    elif fma_sign(q + 1.0, b, a) >= 0:
        q += 1.0
    return q


def test_device_floor_div():
    rng = np.random.default_rng(0)
    for b in (0.2, 0.15, 0.3, 0.0307875, 0.0204399):
        k = rng.integers(0, 2500, 300)
        a = k * b
        a = np.concatenate([a, np.nextafter(a, np.inf), np.nextafter(a, -np.inf),
                            rng.uniform(0, 500, 300)])
        got = np.array([dev_floor_div(x, b) for x in a])
        assert np.array_equal(got, O.floor_div(a, b))


def dev_smooth_mask(cw, R, lo, hi, i, n):
    if hi <= lo:
        return 0.0
    acc = 0.0
    x0, x1 = i - R, i + R
    m0 = int((x0 - hi) / n) - 1
    m1 = int((x1 - lo) / n) + 1
    for m in range(m0, m1 + 1):
        a, b = lo + m * n, hi + m * n - 1
        a, b = max(a, x0), min(b, x1)
        if a > b:
            continue
        acc += cw[b - i + R + 1] - cw[a - i + R]
    return acc


@pytest.mark.parametrize("n,sigma,lo,hi", [(256, 3, 40, 200), (256, 3, 0, 256), (128, 10, 5, 20),
                                           (64, 25, 10, 50), (512, 20, 100, 480), (256, 2, 7, 7)])
def test_smooth_mask_prefix_sums(n, sigma, lo, hi):
    mask = np.zeros(n)
    mask[lo:hi] = 1
    want = gaussian_filter1d(mask, sigma=sigma, mode='wrap')
    cw, R = engine.gauss_prefix(sigma)
    got = np.array([dev_smooth_mask(cw, R, lo, hi, i, n) for i in range(n)])
    assert np.allclose(got, want, rtol=0, atol=4e-16 * max(1, R / 10))


def test_background_spectrum_split():
    """F_row = mz (FFT(S my) + U) + A N delta_0 with U = FFT(b my - A my) for a row
    inside the fill, equal to the FFT of the blended row of the oracle."""
    rng = np.random.default_rng(3)
    coords = rng.uniform(0, 20, size=(500, 3))
    f = np.full(500, 6.003 + 0.0016j)
    N, r, sig, phi = 128, 0.5, 3, 23.0
    xb, yb, zb = (np.ptp(coords[:, i]) for i in range(3))
    avg = O.avg_voxel_f_of(f, xb, yb, zb, r)
    g, st = O.frame_grid(coords, f, phi, N, r, avg, xb, yb, True, sig, return_stats=True)
    S = O.frame_grid(coords, f, phi, N, r, avg, xb, yb, False, 0)
    mv = st['max_voxels']
    A = avg * mv
    L = O.rectangular_collapse_lengths(np.arange(N) * r, yb, xb, phi)
    nm = (mv - np.floor_divide(L, r)).astype(int)
    y = np.arange(N)
    inside_y = (y > st['y_min']) & (y < st['y_max'])
    b = np.where(inside_y, nm * avg, A)
    my = np.zeros(N); my[st['y_min'] + sig:st['y_max'] - sig] = 1
    mz = np.zeros(N); mz[st['z_min'] + sig:st['z_max'] - sig] = 1
    sy = gaussian_filter1d(my, sig, mode='wrap')
    sz = gaussian_filter1d(mz, sig, mode='wrap')
    U = np.fft.fft(b * sy - A * sy)
    for z in range(N):
        want = np.fft.fft(g[z])
        if st['z_min'] <= z < st['z_max']:
            Sz = np.where(inside_y, S[z], 0)
            got = sz[z] * (np.fft.fft(Sz * sy) + U)
            got[0] += A * N
        else:
            got = np.zeros(N, complex)
            got[0] = A * N
        assert np.allclose(got, want, rtol=1e-12, atol=1e-9 * abs(want[0]))


@pytest.mark.parametrize("N", [131, 300, 1048, 2000])
def test_chirpz_tables(N):
    """The chirp-z identity with the product's tables reproduces a length-N DFT."""
    M, ch, K, _ = engine.chirpz_tables(N, True)
    assert M >= 2 * N - 1 and M in (1024, 2048, 4096)
    c = ch[0::2] + 1j * ch[1::2]
    Kc = K[0::2] + 1j * K[1::2]
    x = np.random.default_rng(N).normal(size=N) + 1j * np.random.default_rng(N + 1).normal(size=N)
    a = np.zeros(M, complex)
    a[:N] = x * c
    conv = np.conj(np.fft.fft(np.conj(np.fft.fft(a) * Kc)))   # IFFT via conj(FFT(conj)); 1/M in K
    X = c * conv[:N]
    assert np.allclose(X, np.fft.fft(x), rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("N,k_lo,K", [(131, -30, 61), (300, -150, 300), (1048, -95, 190),
                                      (1048, -141, 283), (700, 5, 40), (2000, -1000, 100)])
def test_chirpz_band_tables(N, k_lo, K):
    """The zoomed chirp-z with the product's band tables reproduces the length-N DFT
    at the signed frequencies k_lo .. k_lo + K - 1 (wrapped mod N), in the kernels'
    arithmetic: input chirp c[n], circular convolution of length M >= N + K - 1,
    output j times c[|k_lo + j|]."""
    M, Kt, _ = engine.chirpz_band_tables(N, k_lo, K, True)
    assert M >= N + K - 1 and M in (1024, 2048, 4096)
    _, ch, _, _ = engine.chirpz_tables(N, True)
    c = ch[0::2] + 1j * ch[1::2]
    H = Kt[0::2] + 1j * Kt[1::2]
    x = np.random.default_rng(N + K).normal(size=N) + 1j * np.random.default_rng(N - K).normal(size=N)
    a = np.zeros(M, complex)
    a[:N] = x * c
    conv = np.conj(np.fft.fft(np.conj(np.fft.fft(a) * H)))
    k = k_lo + np.arange(K)
    X = c[np.abs(k)] * conv[:K]
    assert np.allclose(X, np.fft.fft(x)[k % N], rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("N,k_lo,K,support", [(1048, -141, 283, 606), (1048, -95, 190, 190), (1500, -300, 600, 300),
                                              (700, 5, 40, 700), (1048, -141, 283, 900)])
def test_chirpz_band_tables_with_support(N, k_lo, K, support):
    """Row bands sized to the rows the atoms reach: M is the shortest length with
    M >= min(N, support) + K - 1, and the tables reproduce the band DFT exactly for any
    input that vanishes at n >= L = min(N, M - K + 1) -- the bound the row pass checks
    per frame (y_max <= M - K) -- while an input at n = L (when L < N) aliases."""
    M, Kt, _ = engine.chirpz_band_tables(N, k_lo, K, True, support=support)
    need = min(N, support) + K - 1
    assert M in (1024, 2048, 4096) and M >= need and (M == 1024 or M // 2 < need)
    L = min(N, M - K + 1)
    assert L >= min(N, support)
    _, ch, _, _ = engine.chirpz_tables(N, True)
    c = ch[0::2] + 1j * ch[1::2]
    H = Kt[0::2] + 1j * Kt[1::2]
    k = k_lo + np.arange(K)
    rng = np.random.default_rng(N + K + support)

    def band(x):
        a = np.zeros(M, complex)
        n = min(len(x), M)
        assert not np.any(x[n:])
        a[:n] = (x * c)[:n]
        conv = np.conj(np.fft.fft(np.conj(np.fft.fft(a) * H)))
        return c[np.abs(k)] * conv[:K]

    x = np.zeros(N, complex)
    x[:L] = rng.normal(size=L) + 1j * rng.normal(size=L)
    assert np.allclose(band(x), np.fft.fft(x)[k % N], rtol=1e-12, atol=1e-10)
    if L < N and L < M:
        y = np.zeros(N, complex)
        y[L] = 1.0
        assert not np.allclose(band(y), np.fft.fft(y)[k % N], rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("N,k_lo,K,Z", [(1048, -100, 202, 624), (1048, -100, 202, 136), (1500, -300, 600, 300)])
def test_chirpz_column_band_with_fill_tail(N, k_lo, K, Z):
    """The column pass's short band for fill_bkg: the k = 0 column holds data on z < Z and
    the fill constant A on every row z >= Z; the band takes the rows z < Lc = M - K + 1
    (>= Z) and the constant on [Lc, N) is added in closed form as the kernel does,
    A (exp(-2 pi i Lc k / N) - 1) / (1 - exp(-2 pi i k / N)) (A (N - Lc) at k = 0), with
    the angles reduced in integers.  Equals the full DFT of the column."""
    M, Kt, _ = engine.chirpz_band_tables(N, k_lo, K, True, support=Z)
    Lc = min(N, M - K + 1)
    assert Z <= Lc < N and M < N + K - 1
    _, ch, _, _ = engine.chirpz_tables(N, True)
    c = ch[0::2] + 1j * ch[1::2]
    H = Kt[0::2] + 1j * Kt[1::2]
    rng = np.random.default_rng(N + Z)
    A = complex(rng.normal() * 50, rng.normal() * 50)
    x = np.full(N, A)
    x[:Z] = rng.normal(size=Z) + 1j * rng.normal(size=Z)
    a = np.zeros(M, complex)
    a[:Lc] = x[:Lc] * c[:Lc]
    k = k_lo + np.arange(K)
    F = c[np.abs(k)] * np.conj(np.fft.fft(np.conj(np.fft.fft(a) * H)))[:K]
    kk = np.where(k < 0, k + N, k)
    s1, c1 = np.sin(-2 * np.pi * kk / N), np.cos(-2 * np.pi * kk / N)
    m = (Lc * kk) % N
    sL, cL = np.sin(-2 * np.pi * m / N), np.cos(-2 * np.pi * m / N)
    with np.errstate(divide='ignore', invalid='ignore'):
        q = ((cL - 1.0) + 1j * sL) / ((1.0 - c1) - 1j * s1)
    tail = np.where(kk == 0, A * (N - Lc), A * q)
    want = np.fft.fft(x)[k % N]
    assert np.allclose(F + tail, want, rtol=1e-11, atol=1e-9 * np.abs(want).max())
    # without the tail the short band misses the constant rows (the term is needed)
    assert not np.allclose(F, want, rtol=1e-6, atol=1e-6 * np.abs(want).max())


@pytest.mark.parametrize("N,k_lo,K,support", [(1048, -141, 283, 606), (1048, -95, 190, 191), (1500, -300, 600, 300)])
def test_chirpz_adjoint_tables_with_support(N, k_lo, K, support):
    """With the row band sized to the rows the deposit reaches (M < N + K - 1), the
    adjoint band transform is exact at the outputs o < L = min(N, M - K + 1), the rows
    the forward input may occupy, and the adjoint identity holds for inputs there."""
    M, Kf, _ = engine.chirpz_band_tables(N, k_lo, K, True, support=support)
    assert M < N + K - 1
    L = min(N, M - K + 1)
    Ha = engine.chirpz_adjoint_tables(N, k_lo, K, M, True)
    _, ch, _, _ = engine.chirpz_tables(N, True)
    c = ch[0::2] + 1j * ch[1::2]
    H = Ha[0::2] + 1j * Ha[1::2]
    Hf = Kf[0::2] + 1j * Kf[1::2]
    rng = np.random.default_rng(N * 3 + K)
    k = k_lo + np.arange(K)
    u = rng.normal(size=K) + 1j * rng.normal(size=K)
    a = np.zeros(M, complex)
    a[:K] = u * c[np.abs(k)]
    conv = np.conj(np.fft.fft(np.conj(np.fft.fft(a) * H)))
    o = np.arange(L)
    y = c[o] * conv[:L]
    want = np.exp(-2j * np.pi * np.outer(o, k % N) / N) @ u
    assert np.allclose(y, want, rtol=1e-11, atol=1e-9 * np.abs(want).max())
    # <S F x, u> = <x, S F^H u> for x on the rows [0, L)
    x = np.zeros(N, complex)
    x[:L] = rng.normal(size=L) + 1j * rng.normal(size=L)
    af = np.zeros(M, complex)
    af[:L] = x[:L] * c[:L]
    X = c[np.abs(k)] * np.conj(np.fft.fft(np.conj(np.fft.fft(af) * Hf)))[:K]
    assert np.allclose(X, np.fft.fft(x)[k % N], rtol=1e-12, atol=1e-10)
    a2 = np.zeros(M, complex)
    a2[:K] = np.conj(u) * c[np.abs(k)]
    xbar = np.conj(c[o] * np.conj(np.fft.fft(np.conj(np.fft.fft(a2) * H)))[:L])
    assert np.isclose(np.vdot(u, X), np.vdot(xbar, x[:L]), rtol=1e-11)


@pytest.mark.parametrize("N,k_lo,K", [(131, -30, 61), (300, -150, 300), (1048, -95, 190),
                                      (1048, -141, 283), (700, 5, 40)])
def test_chirpz_adjoint_tables(N, k_lo, K):
    """The adjoint band transform of the differentiable path, in the kernels' arithmetic
    (band values times c[|k|] at positions j < K, circular convolution of the forward
    band's length M with the adjoint kernel, output o times c[o]), equals the
    conjugate transpose of the band DFT: y_o = sum_j u_j exp(-2 pi i o k_j / N), and
    <S F x, u> = <x, conj(adjoint(conj(u)))> for random x, u."""
    M, _, _ = engine.chirpz_band_tables(N, k_lo, K, True)
    Ha = engine.chirpz_adjoint_tables(N, k_lo, K, M, True)
    _, ch, _, _ = engine.chirpz_tables(N, True)
    c = ch[0::2] + 1j * ch[1::2]
    H = Ha[0::2] + 1j * Ha[1::2]
    rng = np.random.default_rng(N * 7 + K)
    u = rng.normal(size=K) + 1j * rng.normal(size=K)
    k = k_lo + np.arange(K)
    a = np.zeros(M, complex)
    a[:K] = u * c[np.abs(k)]
    conv = np.conj(np.fft.fft(np.conj(np.fft.fft(a) * H)))
    y = c[np.arange(N)] * conv[:N]
    want = np.exp(-2j * np.pi * np.outer(np.arange(N), k % N) / N) @ u
    assert np.allclose(y, want, rtol=1e-11, atol=1e-9 * np.abs(want).max())
    # adjoint identity with the forward band transform
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    X = np.fft.fft(x)[k % N]
    a2 = np.zeros(M, complex)
    a2[:K] = np.conj(u) * c[np.abs(k)]
    xbar = np.conj(c[np.arange(N)] * np.conj(np.fft.fft(np.conj(np.fft.fft(a2) * H)))[:N])
    assert np.isclose(np.vdot(u, X), np.vdot(xbar, x), rtol=1e-11)
