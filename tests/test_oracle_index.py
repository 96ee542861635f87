"""Pins of the oracle's index arithmetic (reference: numpy float ``//`` at
tools/voxelgrids.py:328-329, 497-499 and tools/detector.py:217-219;
``np.linspace`` / ``np.fft.fftfreq`` axes at tools/voxelgrids.py:381-402,
tools/comparison.py:724).

The floor division is pinned against exact rational arithmetic on inputs
chosen so that the tempting ``np.floor(a / b)`` gives a different answer.
"""

from fractions import Fraction

import numpy as np
import pytest

from oracle import floor_div
from giwaxsim_amd import engine


def exact_floor(a, b):
    q = Fraction(float(a)) / Fraction(float(b))
    return q.numerator // q.denominator


def adversarial_pairs(rng, n=4000):
    """a = k*b nudged by a few ulps either way, for voxel sizes in use."""
    out = []
    for b in (0.2, 0.25, 0.3, 0.15, 0.1, 0.0307, 0.0491219, 0.0308079, 0.0199778, 0.0204399):
        ks = rng.integers(0, 3000, size=n // 10)
        a = ks * b
        nudges = rng.integers(-3, 4, size=a.shape)
        a = np.array([np.nextafter(x, np.inf if s > 0 else -np.inf) if s else x
                      for x, s in zip(a, nudges)])
        for _ in range(2):
            a = np.where(nudges > 1, np.nextafter(a, np.inf), a)
            a = np.where(nudges < -1, np.nextafter(a, -np.inf), a)
        out.append((a, np.full_like(a, b)))
    a = np.concatenate([p[0] for p in out])
    b = np.concatenate([p[1] for p in out])
    return a, b


def test_known_case():
    assert floor_div(1.0, 0.2) == 4.0
    assert np.floor(1.0 / 0.2) == 5.0


def test_floor_div_is_exact_floor():
    rng = np.random.default_rng(1)
    a, b = adversarial_pairs(rng)
    got = floor_div(a, b)
    want = np.array([exact_floor(x, y) for x, y in zip(a, b)], dtype=np.float64)
    assert np.array_equal(got, want)
    # the naive form disagrees on some of these inputs (so the check has teeth)
    naive = np.floor(a / b)
    assert np.any(naive != want)


def test_floor_div_random_positive():
    rng = np.random.default_rng(2)
    a = rng.uniform(0, 600.0, 20000)
    b = np.full_like(a, 0.2)
    got = floor_div(a, b)
    want = np.array([exact_floor(x, y) for x, y in zip(a, b)], dtype=np.float64)
    assert np.array_equal(got, want)


def test_floor_div_negative_and_zero():
    a = np.array([-0.0, 0.0, -1e-300, -0.2, -0.19999999999999998, -1.0])
    b = np.full_like(a, 0.2)
    want = np.array([exact_floor(x, y) for x, y in zip(a, b)], dtype=np.float64)
    assert np.array_equal(floor_div(a, b), want)


@pytest.mark.parametrize("start,stop,num", [(-2.0, 2.0, 500), (-3.5367, 3.5367, 145),
                                            (0.3, -0.3, 1024), (-15.70796, 15.70796, 2048),
                                            (0.0, 179.6, 453), (-0.0, 0.0, 512)])
def test_linspace_params_reproduce_numpy(start, stop, num):
    """The device evaluates linspace as j*step + start (last = stop)."""
    s, step, e = engine._linspace_params(start, stop, num)
    j = np.arange(num, dtype=np.float64)
    y = j * step + s
    y[-1] = e
    assert np.array_equal(y, np.linspace(start, stop, num))


@pytest.mark.parametrize("N,r", [(256, 0.8), (512, 0.25), (1024, 0.2), (2048, 0.15), (1048, 0.3)])
def test_fftfreq_closed_form(N, r):
    """fftshift(fftfreq(N, r) * 2 * pi)[i] = ((kint * (1/(N r))) * 2) * pi with
    k = (i - N//2) mod N, kint = k or k - N; the device uses this form."""
    q = np.fft.fftshift(np.fft.fftfreq(N, d=r) * 2 * np.pi)
    i = np.arange(N)
    k = (i - N // 2) % N
    kint = np.where(k < (N - 1) // 2 + 1, k, k - N)
    val = 1.0 / (N * r)
    dev = ((kint.astype(np.float64) * val) * 2.0) * np.pi
    assert np.array_equal(q, dev)
    # and it is the textbook frequency grid
    assert np.allclose(q, 2 * np.pi * kint / (N * r), rtol=1e-15, atol=0)


def test_qgrid_linspace_params():
    g = engine.qgrid(0.0307, 2.0)
    start, step, stop = g.linspace_params
    i = np.arange(g.q_num, dtype=np.float64)
    y = i * step + start
    y[-1] = stop
    assert np.array_equal(y, g.q_axis)
