"""Giessibl frequency shift by quadrature — TEST INFRASTRUCTURE ONLY (the checker of the
device Giessibl operator; nothing in ``spammm_amd`` imports it).

The reference has no Giessibl-weighted df (its compute_df_amp takes an equal-weight mean of
-dFz/dz, spammm/SPM/AFM.py:3052-3089); this is the textbook definition

    df(z_c) = -(f0 / (pi k A)) * int_{-1}^{1} Fz(z_c + A u) u / sqrt(1 - u^2) du

(F. J. Giessibl, Phys. Rev. B 56, 16010 (1997); Appl. Phys. Lett. 78, 123 (2001)), z_c the
oscillation centre, A the amplitude [A], k the cantilever stiffness [N/m], Fz in eV/A; the
factor EV_A2_TO_N_M converts eV/A^2 to N/m (spammm/SPM/AFM.py:67-72).

``giessibl_df_continuous`` evaluates it by adaptive quadrature after u = sin(t);
``giessibl_df_piecewise`` applies it to the piecewise-linear interpolant of sampled Fz (clamped
to the end values outside the samples).  Pinned in tests/test_oracle_df.py by closed forms: a
linear force gives -(f0/2k) dFz/dz for every A, the small-A limit, and an exponential force the
modified Bessel function I1.
"""
from __future__ import annotations

import numpy as np

EV_A2_TO_N_M = 16.02176634  # spammm/SPM/AFM.py:67-72


def giessibl_df_continuous(fz_func, zc, A, f0, k, knots=None):
    """The Giessibl integral of fz_func at centres zc.  With u = sin(t) the weight's end-point
    singularity disappears: int_{-1}^{1} g(u) u / sqrt(1-u^2) du = int_{-pi/2}^{pi/2} g(sin t) sin t dt.
    ``knots``: z positions where fz_func has kinks, passed to the quadrature as break points."""
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
    """Giessibl integral of the piecewise-linear interpolant of samples Fz(z) (last axis), clamped to
    the end values outside [min z, max z].  ``z`` may be ascending or descending; output index i has
    oscillation centre z[i]."""
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
