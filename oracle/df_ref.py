"""CPU restatement of the reference's Fz -> df functions — TEST INFRASTRUCTURE ONLY.

This module is the checker of the device df operators.  It is used in this build container only:
by the CPU tests (tests/test_oracle_df.py, tests/test_oracle_direct.py) and by
tests/golden/make_df_vectors.py, which writes the input/output vectors the GPU tests read
(tests/golden/df_vectors.npz).  Nothing in ``spammm_amd`` imports it, and it does not travel to
the GPU runs.

Restated definitions (reference file:line):

* ``compute_df``          spammm/SPM/AFM.py:3048-3050  df = -dFz/dz along axis 2, one-sided
  first-order differences at the two end slices, central ones inside.
* ``compute_df_amp``      spammm/SPM/AFM.py:3052-3089  the 9-node mean of -dFz/dz at
  z + u_k A, u_k = cos((2k+1) pi / 18), equal weights 1/9; Fz at fractional z indices by linear
  interpolation with the index clamped to [0, nz-1] (edge replication); result float32.
* ``compute_df_dir``      spammm/SPM/AFM.py:3105-3123  -n . grad(F . n) along every axis with a
  non-zero component of n (differences as compute_df, per axis spacing).
* ``compute_df_amp_dir``  spammm/SPM/AFM.py:3203-3236  the 9-node mean of -n . grad(F . n) sampled
  at (i,j,k) + u_k A n / spacing, trilinear with every index clamped; result float32.
* ``df_loss_seed``        spammm/SPM/AFM.py:3125-3200  central-difference df on the interior
  slices (0 on the two end slices), L = 1/2 sum w (df - t)^2 / sum w, and as dL/dFz (the other
  channels zero) the transpose of the central stencil over all rows, end rows included, applied
  to g = w (df - t) / sum w (the exact transpose when the end slices carry no weight).
* ``compute_dissipation`` spammm/SPM/AFM.py:3238-3277  |trapezoid integral over the scan heights
  within [h - A |n_z|, h + A |n_z|] of (F_fwd - F_bwd) . n| (all heights when fewer than two
  fall inside).

The arithmetic is written out here (index clamping, explicit difference and trapezoid sums)
rather than through the library routines the reference calls.  Pins (tests/test_oracle_df.py):
the closed forms of tests/SPM/test_afm_morse.py:61-114 (Fx = x z, F = n (n . r), vertical equals
directional), the explicit-stencil / FD checks of tests/SPM/test_afm_morse.py:1376-1457, and the
edge handling against the library composition the reference uses (numpy.gradient with
edge_order 1, scipy.ndimage.map_coordinates with order 1 and mode 'nearest') on random volumes.
"""
from __future__ import annotations

import numpy as np

NODES9 = np.cos((2.0 * np.arange(9) + 1.0) * np.pi / 18.0)     # Gauss-Chebyshev nodes, weight 1/9 each


def _deriv(f, h, axis):
    """d f / d(axis) with spacing h: (f[i+1] - f[i-1]) / 2h inside, (f[1] - f[0]) / h and
    (f[n-1] - f[n-2]) / h at the ends."""
    f = np.moveaxis(np.asarray(f, np.float64), axis, -1)
    n = f.shape[-1]
    if n < 2:
        raise ValueError("a derivative needs at least two samples along its axis")
    d = np.empty_like(f)
    d[..., 0] = (f[..., 1] - f[..., 0]) / h
    d[..., n - 1] = (f[..., n - 1] - f[..., n - 2]) / h
    if n > 2:
        d[..., 1:n - 1] = (f[..., 2:] - f[..., :n - 2]) / (2.0 * h)
    return np.moveaxis(d, -1, axis)


def _shift_axis(f, s, axis):
    """Samples of f at fractional index i + s along `axis` (every i), linear between neighbours,
    with the sample index clamped into [0, n-1]."""
    f = np.moveaxis(np.asarray(f, np.float64), axis, -1)
    n = f.shape[-1]
    x = np.clip(np.arange(n, dtype=np.float64) + s, 0.0, n - 1.0)
    i0 = np.minimum(np.floor(x).astype(np.int64), n - 1)
    i1 = np.minimum(i0 + 1, n - 1)
    t = x - i0
    out = f[..., i0] * (1.0 - t) + f[..., i1] * t
    return np.moveaxis(out, -1, axis)


def _unit_and_spacing(spacing, osc_dir):
    n = np.asarray(osc_dir, dtype=np.float64).reshape(-1)
    if n.shape != (3,) or not np.isfinite(n).all():
        raise ValueError("osc_dir must contain three finite components")
    length = float(np.sqrt(n @ n))
    if length <= 0.0:
        raise ValueError("osc_dir must be non-zero")
    dr = np.abs(np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,))).copy()
    if not np.isfinite(dr).all() or (dr <= 0.0).any():
        raise ValueError("spacing must be a positive scalar or (dx,dy,dz)")
    return n / length, dr


def _project(FEs, n):
    FEs = np.asarray(FEs, np.float64)
    return FEs[..., 0] * n[0] + FEs[..., 1] * n[1] + FEs[..., 2] * n[2]


def compute_df(Fz, dz):
    return -_deriv(Fz, abs(float(dz)), 2)


def compute_df_amp(Fz, dz, amp=1.0):
    h = abs(float(dz))
    acc = np.zeros(np.shape(Fz), np.float64)
    for u in NODES9:
        acc -= _deriv(_shift_axis(Fz, u * amp / h, 2), h, 2) / 9.0
    return acc.astype(np.float32)


def compute_df_dir(FEs, spacing, osc_dir=(0., 0., 1.)):
    n, dr = _unit_and_spacing(spacing, osc_dir)
    Fn = _project(FEs, n)
    out = np.zeros_like(Fn)
    for a in range(3):
        if abs(n[a]) > 1e-15:
            out -= n[a] * _deriv(Fn, dr[a], a)
    return out


def compute_df_amp_dir(FEs, spacing, osc_dir=(0., 0., 1.), amp=1.0):
    n, dr = _unit_and_spacing(spacing, osc_dir)
    Fn = _project(FEs, n)
    out = np.zeros_like(Fn)
    for u in NODES9:
        Fs = Fn
        for a in range(3):
            Fs = _shift_axis(Fs, u * float(amp) * n[a] / dr[a], a)
        for a in range(3):
            if abs(n[a]) > 1e-15:
                out -= n[a] * _deriv(Fs, dr[a], a) / 9.0
    return out.astype(np.float32)


def df_loss_seed(FEs, target_df, dtip, weights=None):
    FEs = np.asarray(FEs)
    nx, ny, nz = FEs.shape[:3]
    if float(dtip) >= 0.0:
        raise ValueError("dtip must be <0")
    if nz < 3:
        raise ValueError("need nz>=3")
    h = abs(float(dtip))
    if weights is None:
        w = np.zeros((nx, ny, nz))
        w[:, :, 1:nz - 1] = 1.0
    else:
        w = np.asarray(weights, dtype=np.float64)
    Fz = np.asarray(FEs[..., 2], np.float64)
    df = np.zeros((nx, ny, nz))
    for j in range(1, nz - 1):
        df[:, :, j] = (Fz[:, :, j - 1] - Fz[:, :, j + 1]) / (2.0 * h)
    r = df - np.asarray(target_df, np.float64)
    W = float(w.sum())
    loss = 0.5 * float((w * r * r).sum()) / W
    g = w * r / W                      # dL/d df
    # the seed is the transpose of the stencil (Fz[j-1] - Fz[j+1]) / 2h taken on EVERY row j, the two
    # boundary rows included: with zero weight there (the default) that is the exact transpose of the
    # interior operator; with weights there it also carries g[0] into Fz[1] and g[nz-1] into Fz[nz-2]
    dFz = np.zeros((nx, ny, nz))
    for j in range(nz):
        if j >= 1:
            dFz[:, :, j - 1] += g[:, :, j] / (2.0 * h)
        if j + 1 < nz:
            dFz[:, :, j + 1] -= g[:, :, j] / (2.0 * h)
    dL = np.zeros((nx, ny, nz, 4))
    dL[..., 2] = dFz
    return loss, df, dL


def _trapezoid(f, x):
    """Trapezoid rule over the last axis of f at abscissae x (any order; signed as x runs)."""
    return (0.5 * (f[..., 1:] + f[..., :-1]) * np.diff(x)).sum(-1)


def compute_dissipation(FEs_fwd, FEs_bwd, h_scan, h_df, amp, osc_dir=(0., 0., 1.)):
    n = np.asarray(osc_dir, dtype=np.float64)
    n = n / np.sqrt(n @ n)
    az = float(amp) * abs(n[2])
    h_scan = np.asarray(h_scan, np.float64)
    dF = _project(FEs_fwd, n) - _project(FEs_bwd, n)
    out = np.zeros(dF.shape[:2] + (len(h_df),), np.float32)
    for i, h in enumerate(np.asarray(h_df, np.float64)):
        inside = np.flatnonzero((h_scan >= h - az) & (h_scan <= h + az))
        if inside.size < 2:
            inside = np.arange(h_scan.size)
        out[:, :, i] = np.abs(_trapezoid(dF[:, :, inside], h_scan[inside]))
    return out
