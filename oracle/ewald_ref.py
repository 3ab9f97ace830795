"""TEST INFRASTRUCTURE ONLY — FP64 NumPy restatement of the 2D Ewald sum for periodic slabs.

Only tests/ and bench.py's checker legs may use this module.  It follows
spammm/surfaces/Ewald2D.py (make_reciprocal_2d :39-60, generate_G_vectors :63-84, compute_C_G
:87-100, compute_w_per_ion :103-116, eval_potential_vacuum :119-133, eval_potential_full_1d
:158-173, eval_potential_brute :176-204) and the OpenCL kernels of kernels/surface.cl:1036-1318:
  phi(rho, z) = C [ -(2 pi / A) sum_i q_i |z - z_i|
                    + sum_{G != 0} sum_i (2 pi / (A |G|)) q_i Re e^{i G.(rho - rho_i)} e^{-|G| |z - z_i|} ]
(the vacuum side z > max z_i has the equivalent form sum_G C_G e^{i G.rho} e^{-|G| z}).
Pinned in tests/test_oracle_ewald.py by: vacuum form = full form above the slab, full form =
brute-force image sum over growing circles, Laplace's equation in the vacuum, periodicity,
and the field = -grad of the potential.
"""
from __future__ import annotations

import numpy as np

COULOMB_CONST = 14.3996448915


def reciprocal_2d(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    area = abs(a[0] * b[1] - a[1] * b[0])
    b1 = 2 * np.pi / area * np.array([b[1], -b[0]])
    b2 = 2 * np.pi / area * np.array([-a[1], a[0]])
    return area, b1, b2


def g_vectors(b1, b2, n_harm):
    hk = [(h, k) for h in range(-n_harm, n_harm + 1) for k in range(-n_harm, n_harm + 1) if (h, k) != (0, 0)]
    hk = np.array(hk)
    Gx = hk[:, 0] * b1[0] + hk[:, 1] * b2[0]
    Gy = hk[:, 0] * b1[1] + hk[:, 1] * b2[1]
    return hk, Gx, Gy, np.hypot(Gx, Gy)


class Slab:
    """Ions (n,4) = (x, y, z, q) periodic along a, b (2-vectors)."""

    def __init__(self, ions, a, b, n_harm):
        self.ions = np.asarray(ions, float)
        self.area, self.b1, self.b2 = reciprocal_2d(a, b)
        self.a, self.b = np.asarray(a, float), np.asarray(b, float)
        self.hk, self.Gx, self.Gy, self.Gn = g_vectors(self.b1, self.b2, n_harm)
        x, y, z, q = self.ions.T
        self.w = (2 * np.pi / (self.area * self.Gn))[:, None] * q[None, :] * \
            np.exp(-1j * (self.Gx[:, None] * x[None, :] + self.Gy[:, None] * y[None, :]))
        self.C = (self.w * np.exp(self.Gn[:, None] * z[None, :])).sum(1)

    def phi_vacuum(self, P):
        P = np.atleast_2d(P)
        ph = np.exp(1j * (P[:, 0:1] * self.Gx[None, :] + P[:, 1:2] * self.Gy[None, :]))
        return COULOMB_CONST * np.real((self.C[None, :] * ph * np.exp(-self.Gn[None, :] * P[:, 2:3])).sum(1))

    def phi_full(self, P):
        P = np.atleast_2d(P)
        if len(P) > 1024:
            return np.concatenate([self.phi_full(P[i:i + 1024]) for i in range(0, len(P), 1024)])
        x, y, z, q = self.ions.T
        dz = np.abs(P[:, 2:3] - z[None, :])                                      # (np, ni)
        phi0 = -(2 * np.pi / self.area) * (q[None, :] * dz).sum(1)
        ph = np.exp(1j * (P[:, 0:1] * self.Gx[None, :] + P[:, 1:2] * self.Gy[None, :]))   # (np, nG)
        dec = np.exp(-self.Gn[None, :, None] * dz[:, None, :])                    # (np, nG, ni)
        s = np.real((ph[:, :, None] * self.w[None, :, :] * dec).sum((1, 2)))
        return COULOMB_CONST * (phi0 + s)

    def field_full(self, P):
        """(E_x, E_y, E_z) = -grad phi, analytic."""
        P = np.atleast_2d(P)
        if len(P) > 1024:
            return np.concatenate([self.field_full(P[i:i + 1024]) for i in range(0, len(P), 1024)])
        x, y, z, q = self.ions.T
        d = P[:, 2:3] - z[None, :]
        sg = np.sign(d)
        dz = np.abs(d)
        ph = np.exp(1j * (P[:, 0:1] * self.Gx[None, :] + P[:, 1:2] * self.Gy[None, :]))
        dec = np.exp(-self.Gn[None, :, None] * dz[:, None, :])
        t = ph[:, :, None] * self.w[None, :, :] * dec                              # (np, nG, ni)
        dphidx = np.real((1j * self.Gx[None, :, None] * t).sum((1, 2)))
        dphidy = np.real((1j * self.Gy[None, :, None] * t).sum((1, 2)))
        dphidz = (-(2 * np.pi / self.area) * (q[None, :] * sg).sum(1)
                  + np.real((-self.Gn[None, :, None] * sg[:, None, :] * t).sum((1, 2))))
        return -COULOMB_CONST * np.stack([dphidx, dphidy, dphidz], 1)

    def phi_brute(self, P, n_rep, square=False):
        """Direct sum over the images n a + m b with n^2 + m^2 <= n_rep^2 (eval_potential_brute,
        kernels/surface.cl:1260-1305), or every |n|,|m| <= n_rep if square (eval_potential_brute,
        spammm/surfaces/Ewald2D.py:176-204)."""
        P = np.atleast_2d(P)
        R = [(n * self.a + m * self.b) for n in range(-n_rep, n_rep + 1) for m in range(-n_rep, n_rep + 1)
             if square or n * n + m * m <= n_rep * n_rep]
        out = np.zeros(len(P))
        for Rv in R:
            d = P[:, None, :] - (self.ions[None, :, :3] + np.array([Rv[0], Rv[1], 0.0]))
            out += (self.ions[None, :, 3] / np.sqrt((d * d).sum(2))).sum(1)
        return COULOMB_CONST * out


def phi_cluster(P, ions):
    """Coulomb potential of a finite set of ions (n,4), no images (eval_potential_cluster,
    kernels/surface.cl:1321-1381)."""
    P = np.atleast_2d(np.asarray(P, float))
    ions = np.asarray(ions, float)
    d = P[:, None, :3] - ions[None, :, :3]
    return COULOMB_CONST * (ions[None, :, 3] / np.sqrt((d * d).sum(2))).sum(1)


def tip_grid(slab, spec_p0, step, shape, Qs, QZs):
    """(nz,ny,nx,4) raw grid of the tip charges in the slab's potential: E = sum_k Qs_k phi(r + QZs_k z),
    F = -grad E (axis-aligned grid at p0 + step*(ix,iy,iz))."""
    nx, ny, nz = shape
    ax = [spec_p0[a] + step[a] * np.arange(shape[a]) for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    out = np.zeros((P.shape[0], 4))
    for q, dzq in zip(Qs, QZs):
        if q == 0:
            continue
        Pk = P + np.array([0, 0, dzq])
        out[:, 3] += q * slab.phi_full(Pk)
        out[:, :3] += q * slab.field_full(Pk)
    return out.reshape(nz, ny, nx, 4)
