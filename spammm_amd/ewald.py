"""2D Ewald electrostatics of periodic slabs on the device.

``SurfaceEwald`` mirrors SurfaceEwaldCL (spammm/surfaces/SurfaceEwald.py:44-449): prepare_system,
eval_vacuum, eval_full, eval_brute, eval_cluster, taking and returning NumPy arrays like the
reference.  ``Ewald2D`` mirrors the container of spammm/surfaces/Ewald2D.py:223-331 (phi_vacuum_xy,
phi_full_2d, phi_full_1d, phi_brute_1d, charge_density_xy, grid helpers) with every evaluation
done by the same device kernels.  ``SurfaceEwald.tip_grid`` writes the slab's electrostatics as a
probe force-field grid (E and F = -grad E for tip charges Qs at offsets QZs), the periodic
Coulomb part of a substrate force field.

Differences from the reference, all in accuracy: structure factors and every point evaluation
accumulate in FP64 (the reference's kernels run in FP32, kernels/surface.cl:71); ions sharing
a height are summed into one layer term per G; G whose structure factor is below drop_tol of the
largest in every layer (default 1e-12: those that vanish up to FP64 rounding) are skipped.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import device as D

COULOMB_CONST = 14.3996448915
DEFAULT_DROP_TOL = 1e-12


def make_reciprocal_2d(a, b):
    """(area, b1, b2) with a_i . b_j = 2 pi delta_ij (spammm/surfaces/Ewald2D.py:39-60)."""
    a = np.asarray(a, dtype=np.float64).reshape(-1)[:2]
    b = np.asarray(b, dtype=np.float64).reshape(-1)[:2]
    area = abs(a[0] * b[1] - a[1] * b[0])
    if not area > 1e-14:
        raise ValueError(f"degenerate lattice: area={area}")
    b1 = 2 * np.pi / area * np.array([b[1], -b[0]])
    b2 = 2 * np.pi / area * np.array([-a[1], a[0]])
    return area, b1, b2


def generate_G_vectors(b1, b2, n_harm):
    """(hs, ks, Gx, Gy, Gn) for |h|,|k| <= n_harm, G != 0, h outer (spammm/surfaces/Ewald2D.py:63-84)."""
    r = np.arange(-n_harm, n_harm + 1)
    H, K = np.meshgrid(r, r, indexing="ij")
    keep = ~((H == 0) & (K == 0))
    hs, ks = H[keep], K[keep]
    Gx = hs * b1[0] + ks * b2[0]
    Gy = hs * b1[1] + ks * b2[1]
    return hs, ks, Gx, Gy, np.hypot(Gx, Gy)


def group_layers(ions):
    """Ions (n,4) float32 reordered by layer (one exact float32 z per layer, ascending) and the
    layer offsets (n_layers+1,) int32."""
    ions = np.ascontiguousarray(np.asarray(ions, dtype=np.float32).reshape(-1, 4))
    if len(ions) == 0:
        raise ValueError("no ions")
    order = np.argsort(ions[:, 2], kind="stable")
    srt = ions[order]
    z = srt[:, 2]
    starts = np.flatnonzero(np.r_[True, z[1:] != z[:-1]])
    off = np.r_[starts, len(srt)].astype(np.int32)
    return np.ascontiguousarray(srt), off


def _d2(v):
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    return (ctypes.c_double * 2)(float(v[0]), float(v[1]))


def _check(rc, what):
    if rc != _lib.PPM_OK:
        raise RuntimeError(f"{what} failed ({rc}): {_lib.load().ppm_last_error().decode()}")


def _points(X, Y, Z, dev):
    import torch
    X = np.asarray(X, dtype=np.float32)
    shape = X.shape
    P = np.zeros((X.size, 4), dtype=np.float32)
    P[:, 0] = X.ravel()
    P[:, 1] = np.broadcast_to(np.asarray(Y, dtype=np.float32), shape).ravel()
    P[:, 2] = np.broadcast_to(np.asarray(Z, dtype=np.float32), shape).ravel()
    return torch.from_numpy(P).to(dev), shape


class SurfaceEwald:
    """Device 2D Ewald evaluator with the interface of SurfaceEwaldCL
    (spammm/surfaces/SurfaceEwald.py:44-449)."""

    def __init__(self, device=None):
        self.dev = D.require_gpu(device)
        self.ion_data = None
        self.a_vec = self.b_vec = None
        self.b1 = self.b2 = None
        self.area = None
        self.n_harm = None
        self.n_layers = None
        self.system = None
        self.ions_d = None

    def make_reciprocal_2d(self, a_vec, b_vec):
        return make_reciprocal_2d(a_vec, b_vec)

    def generate_G_vectors(self, b1, b2, n_harm):
        """G_data (N_G,4) = (h, k, |G|, 0) float32 as SurfaceEwaldCL.generate_G_vectors
        (spammm/surfaces/SurfaceEwald.py:146-163)."""
        hs, ks, _, _, Gn = generate_G_vectors(b1, b2, n_harm)
        return np.stack([hs, ks, Gn, np.zeros_like(Gn)], 1).astype(np.float32)

    @property
    def N_G(self):
        return _lib.load().ppm_ewald_num_g(int(self.n_harm))

    def prepare_system(self, ion_data, a_vec, b_vec, n_harm=3, drop_tol=DEFAULT_DROP_TOL):
        """Layer structure factors and the kept G list (SurfaceEwaldCL.prepare_system,
        spammm/surfaces/SurfaceEwald.py:165-234).  ion_data (n,4) = (x, y, z, q)."""
        import torch
        L = _lib.load()
        self.ion_data = np.ascontiguousarray(np.asarray(ion_data, dtype=np.float32).reshape(-1, 4))
        self.a_vec = np.asarray(a_vec, dtype=np.float64).reshape(-1)[:2].copy()
        self.b_vec = np.asarray(b_vec, dtype=np.float64).reshape(-1)[:2].copy()
        self.n_harm = int(n_harm)
        self.area, self.b1, self.b2 = make_reciprocal_2d(self.a_vec, self.b_vec)
        srt, off = group_layers(self.ion_data)
        self.n_layers = len(off) - 1
        nbytes = L.ppm_ewald_system_bytes(self.n_layers, self.n_harm)
        if nbytes == 0:
            raise ValueError(f"ppm_ewald_system_bytes: {L.ppm_last_error().decode()}")
        self.ions_d = torch.from_numpy(self.ion_data).to(self.dev)
        self._sorted_d = torch.from_numpy(srt).to(self.dev)
        self._off_d = torch.from_numpy(off).to(self.dev)
        self.system = torch.empty(int(nbytes), dtype=torch.uint8, device=self.dev)
        _check(L.ppm_ewald_prepare(self._sorted_d.data_ptr(), self._off_d.data_ptr(), self.n_layers,
                                   ctypes.byref(_d2(self.a_vec)), ctypes.byref(_d2(self.b_vec)), self.n_harm,
                                   float(drop_tol), self.system.data_ptr(), int(nbytes),
                                   D._stream(self.system)), "ppm_ewald_prepare")
        return self

    def _need(self):
        if self.system is None:
            raise ValueError("System not prepared. Call prepare_system() first.")

    def n_kept(self) -> int:
        """Number of kept G, one per pair +-G (reads the count from the device): the series sums
        2 * n_kept() G vectors."""
        self._need()
        import torch
        off = int(_lib.load().ppm_ewald_kept_count_offset(self.n_layers, self.n_harm))
        return int(self.system[off:off + 4].view(torch.int32).item())

    def eval_points(self, P, mode=_lib.PPM_EWALD_FULL):
        """phi at device points (n,4) float32; returns an (n,) float32 device tensor."""
        import torch
        self._need()
        P = D._f32dev(P, "points", 4)
        phi = torch.empty(P.shape[0], dtype=torch.float32, device=P.device)
        _check(_lib.load().ppm_ewald_eval(int(mode), P.data_ptr(), int(P.shape[0]), self.system.data_ptr(),
                                          int(self.system.numel()), self.n_layers, self.n_harm, ctypes.byref(_d2(self.a_vec)),
                                          ctypes.byref(_d2(self.b_vec)), phi.data_ptr(), D._stream(P)),
               "ppm_ewald_eval")
        return phi

    def eval_vacuum(self, X, Y, z, use_fast=False):
        """Potential on an XY grid at height z above every ion (SurfaceEwaldCL.eval_vacuum,
        spammm/surfaces/SurfaceEwald.py:236-296); use_fast is accepted and ignored."""
        P, shape = _points(X, Y, np.full(np.shape(X), z, dtype=np.float32), self.dev)
        return self.eval_points(P, _lib.PPM_EWALD_VACUUM).cpu().numpy().reshape(shape)

    def eval_full(self, X, Y, Z):
        """Potential at arbitrary positions (SurfaceEwaldCL.eval_full, SurfaceEwald.py:298-347)."""
        P, shape = _points(X, Y, Z, self.dev)
        return self.eval_points(P, _lib.PPM_EWALD_FULL).cpu().numpy().reshape(shape)

    def eval_brute(self, X, Y, Z, N_rep=20, square=False):
        """Direct image sum over n^2 + m^2 <= N_rep^2 (SurfaceEwaldCL.eval_brute,
        SurfaceEwald.py:349-405); square=True sums every |n|,|m| <= N_rep (Ewald2D.py:176-204)."""
        import torch
        if self.ion_data is None:
            raise ValueError("System not prepared. Call prepare_system() first.")
        P, shape = _points(X, Y, Z, self.dev)
        phi = torch.empty(P.shape[0], dtype=torch.float32, device=self.dev)
        _check(_lib.load().ppm_ewald_brute(P.data_ptr(), int(P.shape[0]), self.ions_d.data_ptr(),
                                           int(self.ions_d.shape[0]), ctypes.byref(_d2(self.a_vec)),
                                           ctypes.byref(_d2(self.b_vec)), int(N_rep), int(bool(square)),
                                           phi.data_ptr(), D._stream(phi)), "ppm_ewald_brute")
        return phi.cpu().numpy().reshape(shape)

    def eval_cluster(self, X, Y, Z, ion_data):
        """Finite-cluster Coulomb sum, no images (SurfaceEwaldCL.eval_cluster, SurfaceEwald.py:407-449)."""
        import torch
        P, shape = _points(X, Y, Z, self.dev)
        ions = torch.from_numpy(np.ascontiguousarray(np.asarray(ion_data, dtype=np.float32).reshape(-1, 4))).to(self.dev)
        phi = torch.empty(P.shape[0], dtype=torch.float32, device=self.dev)
        _check(_lib.load().ppm_ewald_cluster(P.data_ptr(), int(P.shape[0]), ions.data_ptr(), int(ions.shape[0]),
                                             phi.data_ptr(), D._stream(phi)), "ppm_ewald_cluster")
        return phi.cpu().numpy().reshape(shape)

    def tip_grid(self, spec: D.GridSpec, Qs, QZs, out=None, accumulate=False):
        """(nz,ny,nx,4) float32 device grid (F, E) of the tip charges Qs at z offsets QZs in the slab's
        potential: E = sum_k Qs_k phi(r + QZs_k z), F = -grad E, on a grid with dA, dB in the xy
        plane and dC along z.  accumulate adds onto `out`."""
        import torch
        self._need()
        L = _lib.load()
        nx, ny, nz = (int(v) for v in spec.n)
        if out is None:
            if accumulate:
                raise ValueError("accumulate needs an output grid")
            out = torch.empty((nz, ny, nx, 4), dtype=torch.float32, device=self.dev)
        D._f32dev(out, "out", 4)
        if out.numel() != 4 * spec.nvox:
            raise ValueError(f"out has {out.numel()} floats, the grid needs {4 * spec.nvox}")
        wsb = L.ppm_ewald_grid_workspace_size(self.n_harm, nz)
        ws = torch.empty(int(wsb), dtype=torch.uint8, device=self.dev)
        cs = spec.c()
        _check(L.ppm_ewald_tip_grid(self.system.data_ptr(), int(self.system.numel()), self.n_layers, self.n_harm,
                                    ctypes.byref(_d2(self.a_vec)), ctypes.byref(_d2(self.b_vec)), ctypes.byref(cs),
                                    _lib.f4(Qs), _lib.f4(QZs), int(bool(accumulate)), ws.data_ptr(), int(wsb),
                                    out.data_ptr(), D._stream(out)), "ppm_ewald_tip_grid")
        return out


class Ewald2D:
    """The container of spammm/surfaces/Ewald2D.py:223-331 evaluated by the device kernels."""

    def __init__(self, a_vec, b_vec, rx, ry, rz, q, n_harm=3, device=None, drop_tol=DEFAULT_DROP_TOL):
        self.a_vec = np.asarray(a_vec, dtype=float).reshape(-1)[:2]
        self.b_vec = np.asarray(b_vec, dtype=float).reshape(-1)[:2]
        self.rx, self.ry, self.rz, self.q = (np.asarray(v, dtype=float).reshape(-1) for v in (rx, ry, rz, q))
        self.n_harm = int(n_harm)
        Q_tot = float(np.sum(self.q))
        if abs(Q_tot) > 1e-6:
            print(f"WARNING Ewald2D: cell not neutral, Q_tot={Q_tot:.6f}")
        self.area, self.b1, self.b2 = make_reciprocal_2d(self.a_vec, self.b_vec)
        self.hs, self.ks, self.Gx, self.Gy, self.Gn = generate_G_vectors(self.b1, self.b2, self.n_harm)
        ions = np.stack([self.rx, self.ry, self.rz, self.q], 1)
        self.ew = SurfaceEwald(device).prepare_system(ions, self.a_vec, self.b_vec, self.n_harm, drop_tol)

    @classmethod
    def from_AtomicSystem(cls, sys, n_harm=3, device=None):
        """From an object with apos (N,3), qs (N,) and lvec (3,3) (Ewald2D.from_AtomicSystem,
        spammm/surfaces/Ewald2D.py:250-266)."""
        if getattr(sys, "lvec", None) is None:
            raise ValueError("AtomicSystem must have lattice vectors (lvec)")
        if getattr(sys, "qs", None) is None:
            raise ValueError("AtomicSystem must have charges (qs)")
        lvec = np.asarray(sys.lvec, dtype=float)
        apos = np.asarray(sys.apos, dtype=float)
        return cls(lvec[0, :2], lvec[1, :2], apos[:, 0], apos[:, 1], apos[:, 2], np.asarray(sys.qs), n_harm, device)

    @property
    def N_G(self):
        return len(self.Gx)

    @property
    def z_max(self):
        return float(np.max(self.rz))

    @property
    def z_min(self):
        return float(np.min(self.rz))

    @property
    def Lx(self):
        return float(np.linalg.norm(self.a_vec))

    @property
    def Ly(self):
        return float(np.linalg.norm(self.b_vec))

    def phi_vacuum_xy(self, X, Y, z):
        return self.ew.eval_vacuum(X, Y, z).astype(np.float64)

    def phi_full_2d(self, X, Y, Z):
        return self.ew.eval_full(X, Y, Z).astype(np.float64)

    def phi_full_1d(self, x0, y0, z_arr):
        z = np.asarray(z_arr, dtype=float).reshape(-1)
        return self.ew.eval_full(np.full(z.shape, x0), np.full(z.shape, y0), z).astype(np.float64)

    def phi_brute_1d(self, x0, y0, z_arr, N_rep=20):
        """Every image |n|,|m| <= N_rep, as Ewald2D.eval_potential_brute (Ewald2D.py:176-204)."""
        z = np.asarray(z_arr, dtype=float).reshape(-1)
        return self.ew.eval_brute(np.full(z.shape, x0), np.full(z.shape, y0), z, N_rep, square=True).astype(np.float64)

    def charge_density_xy(self, X, Y):
        """Truncated Fourier series of the in-plane charge sum_G S_G cos(G.rho),
        S_G = (1/A) sum_i q_i cos(G.rho_i) (reconstruct_charge_xy, Ewald2D.py:207-220)."""
        import torch
        dev = self.ew.dev
        Gx = torch.tensor(self.Gx, dtype=torch.float64, device=dev)
        Gy = torch.tensor(self.Gy, dtype=torch.float64, device=dev)
        rx = torch.tensor(self.rx, device=dev)
        ry = torch.tensor(self.ry, device=dev)
        q = torch.tensor(self.q, device=dev)
        S = (q[None, :] * torch.cos(Gx[:, None] * rx[None, :] + Gy[:, None] * ry[None, :])).sum(1) / self.area
        Xt = torch.tensor(np.asarray(X, dtype=float), device=dev).reshape(-1)
        Yt = torch.tensor(np.asarray(Y, dtype=float), device=dev).reshape(-1)
        rho = (S[:, None] * torch.cos(Gx[:, None] * Xt[None, :] + Gy[:, None] * Yt[None, :])).sum(0)
        return rho.cpu().numpy().reshape(np.shape(X))

    def make_xy_grid(self, Ng=200):
        xv = np.linspace(0, self.Lx, Ng)
        yv = np.linspace(0, self.Ly, Ng)
        return np.meshgrid(xv, yv)

    def make_xz_grid(self, Ng=200, z_lo=None, z_hi=None, y_fixed=0.0):
        if z_lo is None:
            z_lo = self.z_min - 1.0
        if z_hi is None:
            z_hi = self.z_max + 4.0
        xv = np.linspace(0, self.Lx, Ng)
        zv = np.linspace(z_lo, z_hi, Ng)
        X, Z = np.meshgrid(xv, zv)
        return X, np.full_like(X, y_fixed), Z
