"""TEST INFRASTRUCTURE ONLY — FP64 NumPy restatement of the FDBM force-field producer.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline / accuracy legs may use this
module, as the checker.  Each function follows the reference's NumPy (CPU) form of the same
stage:
  fft_poisson       spammm/SPM/AFM.py:3280-3298   (fft_poisson_cpu)
  pauli_overlap     spammm/SPM/AFM.py:4085-4089   (compute_pauli_overlap_cpu)
  pauli_scale       spammm/SPM/AFM.py:4106-4128   (scale_pauli_field)
  es_conv           spammm/SPM/AFM.py:4137-4150   (compute_es_conv_field_cpu)
  dispersion        spammm/SPM/AFM.py:4418-4431   (compute_dispersion_grid, NumPy branch)
  gradient_periodic kernels/AFM.cl:1151-1196      (gradient_central_diff, REPEAT addressing)
  pad_roll          kernels/AFM.cl:1202-1221      (fdbm_pad_roll_f32)
Pinned by closed forms in tests/test_oracle_fdbm.py (a single Fourier mode through Poisson,
delta functions through the correlation and convolution, one atom for the dispersion, a
sinusoid for the periodic central difference, the tip peak landing at index 0).
"""
from __future__ import annotations

import numpy as np

COULOMB_CONST = 14.3996448915


def _k2(shape, step):
    ks = [2 * np.pi * np.fft.fftfreq(n, d=step) for n in shape]
    KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
    return KX ** 2 + KY ** 2 + KZ ** 2


def fft_poisson(rho, step):
    rho_k = np.fft.fftn(np.asarray(rho, np.float64))
    k2 = _k2(rho_k.shape, step)
    k2[0, 0, 0] = 1.0
    V_k = 4.0 * np.pi * COULOMB_CONST * rho_k / k2
    V_k[0, 0, 0] = 0.0
    return np.real(np.fft.ifftn(V_k))


def pauli_overlap(rho, tip, step, clip=1e-30):
    o = step ** 3 * np.real(np.fft.ifftn(np.fft.fftn(np.asarray(rho, np.float64)) *
                                         np.conj(np.fft.fftn(np.asarray(tip, np.float64)))))
    return np.clip(o, clip, None)


def pauli_scale(overlap, A, beta, clip=1e-30):
    return A * np.maximum(np.asarray(overlap, np.float64), clip) ** beta


def flip3(a):
    return np.asarray(a)[::-1, ::-1, ::-1]


def es_conv(V, tip_delta, step, tip_rolled=True):
    k = flip3(tip_delta)
    if not tip_rolled:
        nx, ny, nz = k.shape
        k = np.roll(np.roll(np.roll(k, -(nx // 2), 0), -(ny // 2), 1), -(nz // 2), 2)
    return step ** 3 * np.real(np.fft.ifftn(np.fft.fftn(np.asarray(V, np.float64)) *
                                            np.fft.fftn(np.asarray(k, np.float64))))


def es_fused(rho_diff, tip_delta, step):
    """The fused electrostatics of the device path: Poisson, then the convolution with the reversed tip."""
    return es_conv(fft_poisson(rho_diff, step), tip_delta, step, tip_rolled=True)


def dispersion(atoms_w, origin, step, shape, RA=1.5):
    """-sum_a w_a / (|p - r_a|^2 + RA^2)^3, p = origin + step*(ix,iy,iz); atoms_w rows (x,y,z,w)."""
    axes = [origin[a] + step * np.arange(shape[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    E = np.zeros(shape)
    for x, y, z, w in np.asarray(atoms_w, np.float64):
        E -= w / ((X - x) ** 2 + (Y - y) ** 2 + (Z - z) ** 2 + RA * RA) ** 3
    return E


def gradient_periodic(E, step):
    """(nx,ny,nz,4) = (-(E[i+1]-E[i-1])/(2 step) per axis with wrap-around, E)."""
    E = np.asarray(E, np.float64)
    out = np.empty(E.shape + (4,))
    for a in range(3):
        out[..., a] = -(np.roll(E, -1, a) - np.roll(E, 1, a)) / (2.0 * step)
    out[..., 3] = E
    return out


def pad_roll(src, shape, offset, roll):
    """dst[i] = padded[(i + roll) mod n]; padded = zeros with src at offset."""
    pad = np.zeros(shape)
    sx, sy, sz = src.shape
    ox, oy, oz = offset
    pad[ox:ox + sx, oy:oy + sy, oz:oz + sz] = src
    return np.roll(pad, tuple(-int(r) for r in roll), axis=(0, 1, 2))
