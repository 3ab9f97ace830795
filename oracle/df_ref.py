"""NumPy/SciPy restatement of the reference's Fz -> df functions — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg use this module, as the checker.  The product computes df on the GPU
(``spammm_amd.df``); nothing in ``spammm_amd`` imports this file.

Restated functions (reference file:line):

* ``compute_df``          spammm/SPM/AFM.py:3048-3050  (-np.gradient along z)
* ``compute_df_amp``      spammm/SPM/AFM.py:3052-3089  (9-node Gauss-Chebyshev
  equal-weight mean of -dFz/dz, linear interpolation mode='nearest')
* ``compute_df_dir``      spammm/SPM/AFM.py:3105-3123
* ``compute_df_amp_dir``  spammm/SPM/AFM.py:3203-3236
* ``df_loss_seed``        spammm/SPM/AFM.py:3125-3200
* ``compute_dissipation`` spammm/SPM/AFM.py:3238-3277

Pins: the closed forms of tests/SPM/test_afm_morse.py:61-114 (Fx=x*z, F=n(n.r),
vertical-vs-directional equality) and the explicit-stencil/FD checks of
tests/SPM/test_afm_morse.py:1376-1457 are restated in tests/test_oracle_df.py.
The edge handling of ``compute_df_amp*`` (map_coordinates(order=1,
mode='nearest') + np.gradient edge_order=1) follows the source text with the
same SciPy/NumPy routines; the reference pins only the interior
(tests/SPM/test_afm_morse.py:75-76), so edge values are restated, not
cross-checked.

The Giessibl frequency shift (not present in the reference; named by the build
target) has a continuous definition here, ``giessibl_df_continuous``,
evaluated by adaptive quadrature with the algebraic end-point weight, and a
discrete definition, ``giessibl_df_piecewise``: the same integral applied to the
piecewise-linear interpolant of sampled Fz (clamped outside the samples).
"""
from __future__ import annotations

import numpy as np

EV_A2_TO_N_M = 16.02176634  # spammm/SPM/AFM.py:67-72


def compute_df(Fz, dz):
    return -np.gradient(Fz, abs(dz), axis=2)


def _gc_nodes(n_quad=9):
    k = np.arange(n_quad)
    return np.cos((2 * k + 1) / (2 * n_quad) * np.pi), np.full(n_quad, 1.0 / n_quad)


def compute_df_amp(Fz, dz, amp=1.0):
    from scipy.ndimage import map_coordinates
    nz = Fz.shape[2]
    dz_abs = abs(dz)
    u_k, w_k = _gc_nodes()
    iz_grid = np.arange(nz, dtype=np.float64)
    df = np.zeros_like(Fz, dtype=np.float64)
    nx, ny = Fz.shape[0], Fz.shape[1]
    for uk, wk in zip(u_k, w_k):
        iz_query = iz_grid[None, None, :] + uk * amp / dz_abs
        coords = np.array([np.broadcast_to(np.arange(nx)[:, None, None], (nx, ny, nz)),
                           np.broadcast_to(np.arange(ny)[None, :, None], (nx, ny, nz)),
                           np.broadcast_to(iz_query, (nx, ny, nz))])
        Fz_s = map_coordinates(Fz.astype(np.float64), coords, order=1, mode='nearest')
        df -= wk * np.gradient(Fz_s, dz_abs, axis=2)
    return df.astype(np.float32)


def _dir_spacing(spacing, osc_dir):
    n = np.asarray(osc_dir, dtype=np.float64)
    if n.shape != (3,) or not np.isfinite(n).all():
        raise ValueError("osc_dir must contain three finite components")
    norm = float(np.linalg.norm(n))
    if norm <= 0.0:
        raise ValueError("osc_dir must be non-zero")
    n = n / norm
    dr = (np.full(3, abs(float(spacing))) if np.ndim(spacing) == 0
          else np.abs(np.asarray(spacing, dtype=np.float64)))
    if dr.shape != (3,) or not np.isfinite(dr).all() or np.any(dr <= 0.0):
        raise ValueError("spacing must be a positive scalar or (dx,dy,dz)")
    return n, dr


def compute_df_dir(FEs, spacing, osc_dir=(0., 0., 1.)):
    FEs = np.asarray(FEs)
    n, dr = _dir_spacing(spacing, osc_dir)
    F_n = FEs[..., 0] * n[0] + FEs[..., 1] * n[1] + FEs[..., 2] * n[2]
    df = np.zeros_like(F_n, dtype=np.result_type(F_n.dtype, np.float64))
    for ia in np.flatnonzero(np.abs(n) > 1e-15):
        df -= n[ia] * np.gradient(F_n, dr[ia], axis=ia)
    return df


def compute_df_amp_dir(FEs, spacing, osc_dir=(0., 0., 1.), amp=1.0):
    from scipy.ndimage import map_coordinates
    FEs = np.asarray(FEs)
    n, dr = _dir_spacing(spacing, osc_dir)
    F_n = FEs[..., 0] * n[0] + FEs[..., 1] * n[1] + FEs[..., 2] * n[2]
    active = np.flatnonzero(np.abs(n) > 1e-15)
    u_k, w_k = _gc_nodes()
    df = np.zeros_like(F_n, dtype=np.float64)
    coords0 = np.indices(F_n.shape, dtype=np.float64)
    shift = (float(amp) * n / dr).reshape(3, 1, 1, 1)
    for uk, wk in zip(u_k, w_k):
        F_s = map_coordinates(F_n.astype(np.float64), coords0 + uk * shift, order=1, mode='nearest')
        for ia in active:
            df -= wk * n[ia] * np.gradient(F_s, dr[ia], axis=ia)
    return df.astype(np.float32)


def df_loss_seed(FEs, target_df, dtip, weights=None):
    FEs = np.asarray(FEs)
    target_df = np.asarray(target_df, dtype=np.float64)
    nx, ny, nz = FEs.shape[:3]
    dtip = float(dtip)
    if dtip >= 0.0:
        raise ValueError("dtip must be <0")
    if nz < 3:
        raise ValueError("need nz>=3")
    if weights is None:
        w = np.ones((nx, ny, nz))
        w[:, :, 0] = 0.0
        w[:, :, -1] = 0.0
    else:
        w = np.asarray(weights, dtype=np.float64)
    Fz = FEs[..., 2].astype(np.float64)
    dz = abs(dtip)
    df = np.zeros((nx, ny, nz))
    df[:, :, 1:nz - 1] = -(Fz[:, :, 2:] - Fz[:, :, :-2]) / (2.0 * dz)
    diff = df - target_df
    wsum = float(w.sum())
    loss = 0.5 * float(np.sum(w * diff ** 2)) / wsum
    g = w * diff / wsum
    c = 1.0 / (2.0 * dz)
    dFz = np.zeros((nx, ny, nz))
    dFz[:, :, 1:nz - 1] = (g[:, :, 2:] - g[:, :, :-2]) * c
    dFz[:, :, 0] = g[:, :, 1] * c
    dFz[:, :, -1] = -g[:, :, nz - 2] * c
    dL = np.zeros((nx, ny, nz, 4))
    dL[..., 2] = dFz
    return loss, df, dL


def compute_dissipation(FEs_fwd, FEs_bwd, h_scan, h_df, amp, osc_dir=(0., 0., 1.)):
    FEs_fwd = np.asarray(FEs_fwd, dtype=np.float64)
    FEs_bwd = np.asarray(FEs_bwd, dtype=np.float64)
    h_scan = np.asarray(h_scan, dtype=np.float64)
    h_df = np.asarray(h_df, dtype=np.float64)
    n = np.asarray(osc_dir, dtype=np.float64)
    n = n / np.linalg.norm(n)
    amp_z = float(amp) * abs(n[2])
    dF = (FEs_fwd[..., :3] @ n) - (FEs_bwd[..., :3] @ n)
    nx, ny, nzs = dF.shape
    E = np.zeros((nx, ny, len(h_df)), dtype=np.float32)
    trapz = getattr(np, "trapezoid", None) or np.trapz
    for i, h in enumerate(h_df):
        mask = (h_scan >= h - amp_z) & (h_scan <= h + amp_z)
        if mask.sum() < 2:
            mask = np.ones(nzs, dtype=bool)
        E[:, :, i] = np.abs(trapz(dF[:, :, mask], x=h_scan[mask], axis=2)).astype(np.float32)
    return E


# ---------------------------------------------------------------------------
# Giessibl frequency shift
#   df(z_c) = -(f0 / (pi k A)) * int_{-1}^{1} Fz(z_c + A u) u / sqrt(1-u^2) du
# (F. J. Giessibl, Phys. Rev. B 56, 16010 (1997); Appl. Phys. Lett. 78, 123 (2001)),
# z_c = oscillation centre, A = amplitude (peak), k cantilever stiffness [N/m],
# Fz in eV/A, A in A  ->  factor EV_A2_TO_N_M converts eV/A^2 to N/m.
# ---------------------------------------------------------------------------

def giessibl_df_continuous(fz_func, zc, A, f0, k, knots=None):
    """Continuous Giessibl integral by adaptive quadrature after u = sin(theta):
    int_{-1}^{1} g(u) u / sqrt(1-u^2) du = int_{-pi/2}^{pi/2} g(sin t) sin t dt (no end-point
    singularity).  ``knots``: z positions where fz_func has kinks, passed as break points."""
    from scipy.integrate import quad
    zc = np.atleast_1d(np.asarray(zc, dtype=np.float64))
    out = np.zeros_like(zc)
    for i, z in enumerate(zc):
        pts = None
        if knots is not None:
            u = (np.asarray(knots, dtype=np.float64) - z) / A
            u = u[(u > -1.0) & (u < 1.0)]
            pts = np.arcsin(u) if u.size else None
        val, _ = quad(lambda t: fz_func(z + A * np.sin(t)) * np.sin(t), -np.pi / 2, np.pi / 2, points=pts,
                      epsabs=1e-14, epsrel=1e-12, limit=400)
        out[i] = val
    return -(f0 / (np.pi * k * A)) * EV_A2_TO_N_M * out


def giessibl_df_piecewise(Fz, z, A, f0, k):
    """Giessibl integral of the piecewise-linear interpolant of samples Fz(z) (last axis),
    clamped to the end values outside [min z, max z], evaluated by QUADPACK per output point.

    ``z`` may be ascending or descending; output index i has oscillation centre z[i]."""
    Fz = np.asarray(Fz, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    order = np.argsort(z)
    zs = z[order]
    flat = Fz.reshape(-1, Fz.shape[-1])[:, order]
    out = np.zeros((flat.shape[0], z.size))
    for p in range(flat.shape[0]):
        fz = flat[p]
        out[p] = giessibl_df_continuous(lambda zz: np.interp(zz, zs, fz), z, A, f0, k, knots=zs)
    return out.reshape(Fz.shape[:-1] + (z.size,))
