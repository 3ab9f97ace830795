"""The FDBM oracle (oracle/fdbm_ref.py) against closed forms.

* Poisson: a single Fourier mode rho = cos(k.r) on a periodic box has V = 4 pi C cos(k.r) / |k|^2.
* Correlation / convolution: delta functions shift exactly; a delta tip at index 0 reproduces
  rho * dV (correlation) and its mirror (convolution with the reversed tip lands at -i).
* Dispersion: one atom gives -w / (r^2 + RA^2)^3 at every voxel.
* Periodic central difference of sin(2 pi x / L): -cos(2 pi x / L) sin(h k)/h, exactly.
* pad_roll puts the source peak at index (0,0,0) for the reference's roll.
"""
import numpy as np
import pytest

from oracle import fdbm_ref as F


def test_poisson_single_mode():
    n, step = (16, 12, 10), 0.3
    L = np.array(n) * step
    ix = np.arange(n[0])[:, None, None] * step
    iy = np.arange(n[1])[None, :, None] * step
    kx, ky = 2 * np.pi * 2 / L[0], 2 * np.pi * 1 / L[1]
    rho = np.cos(kx * ix + ky * iy) * np.ones(n)
    V = F.fft_poisson(rho, step)
    expect = 4 * np.pi * F.COULOMB_CONST * rho / (kx * kx + ky * ky)
    assert np.allclose(V, expect, atol=1e-12 * np.abs(expect).max())
    # a constant (k = 0) density has no potential (neutralising background)
    assert np.allclose(F.fft_poisson(np.ones(n), step), 0.0, atol=1e-12)


def test_correlation_and_convolution_of_deltas():
    n, step = (8, 6, 10), 0.5
    rho = np.zeros(n); rho[3, 2, 7] = 2.0
    tip = np.zeros(n); tip[0, 0, 0] = 1.0
    o = F.pauli_overlap(rho, tip, step, clip=-np.inf)
    assert np.allclose(o, rho * step ** 3, atol=1e-14)
    tip2 = np.zeros(n); tip2[1, 1, 2] = 1.0              # tip offset by (1,1,2): overlap shifts by -(1,1,2)
    o2 = F.pauli_overlap(rho, tip2, step, clip=-np.inf)
    assert abs(o2[2, 1, 5] - 2.0 * step ** 3) < 1e-14 and abs(o2.sum() - 2.0 * step ** 3) < 1e-13
    # reversed-tip convolution: flip3 of a delta at index 0 sits at (n-1); E = dV V(i - (n-1)) = dV V(i + 1)
    V = np.random.default_rng(1).normal(size=n)
    E = F.es_conv(V, tip, step, tip_rolled=True)
    assert np.allclose(E, step ** 3 * np.roll(V, (-1, -1, -1), axis=(0, 1, 2)), atol=1e-12)


def test_pauli_scale_clip():
    o = np.array([1e-40, 0.5, 2.0]).reshape(1, 1, 3)
    assert np.allclose(F.pauli_scale(o, 3.0, 0.5), 3.0 * np.sqrt([1e-30, 0.5, 2.0]).reshape(1, 1, 3))


def test_dispersion_one_atom():
    n, step, RA, w = (5, 4, 3), 0.7, 1.5, 2.5
    origin = np.array([-1.0, 0.5, 2.0])
    E = F.dispersion([[0.3, -0.2, 0.9, w]], origin, step, n, RA)
    p = origin + step * np.array([4, 1, 2])
    r2 = ((p - [0.3, -0.2, 0.9]) ** 2).sum()
    assert abs(E[4, 1, 2] + w / (r2 + RA * RA) ** 3) < 1e-15


def test_gradient_of_sinusoid():
    n, step = (24, 5, 7), 0.2
    k = 2 * np.pi / (n[0] * step)
    x = np.arange(n[0]) * step
    E = np.sin(k * x)[:, None, None] * np.ones(n)
    G = F.gradient_periodic(E, step)
    expect = -np.cos(k * x) * np.sin(k * step) / step
    assert np.allclose(G[..., 0], expect[:, None, None], atol=1e-13)
    assert np.allclose(G[..., 1], 0.0) and np.allclose(G[..., 2], 0.0) and np.array_equal(G[..., 3], E)


def test_pad_roll_puts_the_peak_at_the_origin():
    raw = np.zeros((5, 4, 6)); raw[2, 1, 4] = 3.0
    shape = (16, 12, 20)
    ox, oy, oz = [(t - c) // 2 for t, c in zip(shape, raw.shape)]
    r = (ox + 2, oy + 1, oz + 4)
    dst = F.pad_roll(raw, shape, (ox, oy, oz), r)
    assert dst[0, 0, 0] == 3.0 and dst.sum() == 3.0
