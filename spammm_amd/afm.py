"""AFMulator — drop-in for the probe-particle AFM path of spammm/SPM/AFM.py on one MI355X.

Keeps the call surface of ``class AFMulator`` (spammm/SPM/AFM.py:265) for the
PPM path — ``assign_params``, ``setup_grid`` / ``setup_grid_world`` /
``setup_grid_lvec``, ``make_forcefield``, ``run_scan``, ``get_raw_FE``,
``setup_fdbm_grid`` / ``scan_fdbm``, ``get_df`` — and the module functions
``compute_df``, ``compute_df_amp``, ``compute_df_dir``, ``compute_df_amp_dir``,
``df_loss_seed`` and ``compute_dissipation`` (spammm/SPM/AFM.py:3048-3277), with
the same argument meaning, return shapes and exceptions.

Differences in mechanism (not in results):
  * the force field lives in a device tensor (nz,ny,nx,4) float32 instead of an
    OpenCL image; sampling is the same REPEAT/linear-filter rule in software;
  * every method still returns host NumPy arrays (the reference blocks on
    ``queue.finish()`` too); the device results stay available as attributes
    (``FF``, ``scan_FEs_dev``, ``scan_PPs_dev``) for callers that chain GPU work;
  * ``run_scan`` additionally records probe positions and a per-state iteration
    count (+n converged, -n not) — the reference's grid path has no telemetry;
  * ``set_pbc`` adds periodic x,y images to ``make_forcefield``.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib, df as _df
from . import device as D
from .params import (ELEMENTS_REFERENCE, EV_A2_PER_N_M, N_M_PER_EV_A2, Z_TO_NAME, mix_lj, mix_morse,
                     read_element_types, stiffness_eVA2_to_Nm, stiffness_Nm_to_eVA2)

COULOMB_CONST = 14.3996448915  # [eV*Ang/e^2]  kernels/common.cl:29

__all__ = [
    "AFMulator", "COULOMB_CONST", "EV_A2_PER_N_M", "N_M_PER_EV_A2", "stiffness_Nm_to_eVA2",
    "stiffness_eVA2_to_Nm", "build_scan_xy_points_vectorized", "build_scan_points_vectorized",
    "compute_df", "compute_df_amp", "compute_df_dir", "compute_df_amp_dir", "df_loss_seed", "df_giessibl",
    "compute_dissipation",
]


def build_scan_xy_points_vectorized(scan_p0, scan_da, scan_db, nx, ny):
    """(n_scan,4) float32 ix-major scan starts (spammm/SPM/AFM.py:82-96)."""
    return D.scan_points(scan_p0, scan_da, scan_db, nx, ny)


def build_scan_points_vectorized(scan_xs, scan_ys, h_scan):
    """(nx*ny*nz,4) float32 (x,y,z,0) points, ij order (spammm/SPM/AFM.py:99-123)."""
    scan_xs = np.asarray(scan_xs, dtype=np.float32)
    scan_ys = np.asarray(scan_ys, dtype=np.float32)
    h_scan = np.asarray(h_scan, dtype=np.float32)
    gx, gy, gz = np.meshgrid(scan_xs, scan_ys, h_scan, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel(), gz.ravel(), np.zeros(gx.size, dtype=np.float32)], axis=1)
    return np.ascontiguousarray(pts, dtype=np.float32)


def _read_xyz(path):
    """Minimal .xyz reader: element names, positions and an optional 5th charge column."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    en, pos, qs = [], [], []
    for ln in lines[2:2 + n]:
        p = ln.split()
        name = p[0]
        if name.isdigit():
            name = Z_TO_NAME.get(int(name), name)
        en.append(name)
        pos.append([float(p[1]), float(p[2]), float(p[3])])
        qs.append(float(p[4]) if len(p) > 4 else 0.0)
    return en, np.array(pos, dtype=np.float64), np.array(qs, dtype=np.float64)


class Molecule:
    """The attributes of AtomicSystem the AFM path reads: natoms, enames, apos, qs, lvec."""

    def __init__(self, enames, apos, qs=None, lvec=None):
        self.enames = list(enames)
        self.apos = np.asarray(apos, dtype=np.float64).reshape(-1, 3)
        self.natoms = len(self.enames)
        self.qs = None if qs is None else np.asarray(qs, dtype=np.float64)
        self.lvec = None if lvec is None else np.asarray(lvec, dtype=np.float64)


class AFMulator:
    """Probe-particle AFM simulator on one MI355X (LJ / Morse + point charges)."""

    @staticmethod
    def _make_dinv_axis_aligned(L, origin, n):
        return D.dinv_axis_aligned(L, origin, n)

    @staticmethod
    def _make_dinv_lvec(invMT, n):
        return D.dinv_lvec(invMT, n)

    DEFAULT_tipA = np.array([1., 0., 0., 0.], dtype=np.float32)
    DEFAULT_tipB = np.array([0., 1., 0., 0.], dtype=np.float32)
    DEFAULT_tipC = np.array([0., 0., 1., -0.1], dtype=np.float32)
    DEFAULT_stiffness = np.array([-0.03, -0.03, -0.03, -1.0], dtype=np.float32)
    DEFAULT_dpos0 = np.array([0., 0., -4.0, 4.0], dtype=np.float32)
    DEFAULT_relax_pars = np.array([0.5, 0.1, 0.02, 0.5], dtype=np.float32)
    DEFAULT_surfFF = np.array([0., 0., 0., 0.], dtype=np.float32)
    DEFAULT_tipQs = np.array([0., -0.1, 0.1, 0.], dtype=np.float32)
    DEFAULT_tipQZs = np.array([0., 1.8, 3.6, 0.], dtype=np.float32)

    # FIRE loop bounds of kernels/AFM.cl:121-122, runtime here.
    N_RELAX_STEP_MAX = 128
    F2CONV = 1e-8

    def __init__(self, cl_src_dir=None, use_morse=False, nloc=32, use_fire=True, device=None):
        # use_fire picks the integrator of every relaxation this instance runs (run_scan, scan_fdbm,
        # run_scan_morse_direct), as the reference compiles its program with -D OPT_FIRE=<use_fire>
        # (spammm/SPM/AFM.py:320-321): FIRE, or the damped velocity v *= (1 - damp).
        self.device = D.require_gpu(device)
        self.use_fire = bool(use_fire)
        self.use_morse = use_morse
        self.nloc = nloc
        self.mol = self.elem_types = None
        self.atoms_arr = self.cLJs_arr = None
        self.FF = None
        self.n = self.p0 = self.L = None
        self.dA = self.dB = self.dC = None
        self.dinvA = self.dinvB = self.dinvC = None
        self.mol_shift = np.zeros(3, dtype=np.float32)
        self.npbc = (0, 0)
        self.lvA = np.zeros(3, dtype=np.float32)
        self.lvB = np.zeros(3, dtype=np.float32)
        self.cap = True
        self.max_iter = self.N_RELAX_STEP_MAX
        self.f2conv = self.F2CONV
        self.tipA = self.DEFAULT_tipA.copy()
        self.tipB = self.DEFAULT_tipB.copy()
        self.tipC = self.DEFAULT_tipC.copy()
        self.stiffness = self.DEFAULT_stiffness.copy()
        self.dpos0 = self.DEFAULT_dpos0.copy()
        self.relax_pars = self.DEFAULT_relax_pars.copy()
        self.surfFF = self.DEFAULT_surfFF.copy()
        self.tipQs = self.DEFAULT_tipQs.copy()
        self.tipQZs = self.DEFAULT_tipQZs.copy()
        self.img_FF_fdbm = None
        self.fdbm_spec = None
        self.fdbm_origin = self.fdbm_step = self.fdbm_shape = None
        self.scan_FEs_dev = self.scan_PPs_dev = None

    # ── molecule and parameters ─────────────────────────────────────────────

    def load_molecule(self, xyz_path):
        en, pos, qs = _read_xyz(xyz_path)
        self.mol = Molecule(en, pos, qs)
        return self.mol

    def set_molecule(self, enames, apos, qs=None, lvec=None):
        self.mol = Molecule(enames, apos, qs, lvec)
        return self.mol

    def assign_params(self, params_path=None, tip_R=1.452, tip_E=0.0006808, tip_alpha=-1.8, elements=None):
        """LJ (c6,c12) or Morse (R0,E0,alpha) rows by the mixing rules of spammm/SPM/AFM.py:730-768."""
        assert self.mol is not None, "call load_molecule() first"
        if elements is None:
            elements = read_element_types(params_path) if params_path is not None else ELEMENTS_REFERENCE
        self.elem_types = elements
        na = self.mol.natoms
        atoms = np.zeros((na, 4), dtype=np.float32)
        atoms[:, :3] = self.mol.apos[:na, :3].astype(np.float32)
        if self.mol.qs is not None:
            atoms[:, 3] = self.mol.qs[:na].astype(np.float32)
        for e in self.mol.enames[:na]:
            assert e in elements, f"Element '{e}' not in element table"
        if self.use_morse:
            self.cLJs_arr = mix_morse(self.mol.enames[:na], elements, tip_R, tip_E, tip_alpha)
        else:
            self.cLJs_arr = mix_lj(self.mol.enames[:na], elements, tip_R, tip_E)
        self.atoms_arr = atoms
        return atoms, self.cLJs_arr

    # ── grid set-up (spammm/SPM/AFM.py:772-893) ─────────────────────────────

    def _set_spec(self, n, p0, dA, dB, dC, dinv):
        self.n = np.array(n, dtype=np.int32)
        self.p0 = np.array([p0[0], p0[1], p0[2], 0.], dtype=np.float32)
        self.dA = np.array([dA[0], dA[1], dA[2], 0.], dtype=np.float32)
        self.dB = np.array([dB[0], dB[1], dB[2], 0.], dtype=np.float32)
        self.dC = np.array([dC[0], dC[1], dC[2], 0.], dtype=np.float32)
        self.dinvA, self.dinvB, self.dinvC = dinv

    @property
    def spec(self) -> D.GridSpec:
        return D.GridSpec(tuple(int(x) for x in self.n), self.p0[:3], self.dA[:3], self.dB[:3], self.dC[:3],
                          self.dinvA, self.dinvB, self.dinvC)

    def setup_grid(self, n=(100, 100, 60), L=None, margin=3.0, z_top=12.0, shift_atoms=True):
        assert self.atoms_arr is not None, "call assign_params() first"
        apos = self.atoms_arr[:, :3]
        mn, mx = apos.min(axis=0), apos.max(axis=0)
        if L is None:
            L = np.array([(mx[0] - mn[0]) + 2 * margin, (mx[1] - mn[1]) + 2 * margin,
                          (mx[2] - mn[2]) + margin + z_top], dtype=np.float32)
        else:
            L = np.array(L, dtype=np.float32)
        nx, ny, nz = n
        p0_raw = np.array([mn[0] - margin, mn[1] - margin, mn[2] - margin / 2], dtype=np.float32)
        if shift_atoms:
            self.mol_shift = -p0_raw
            self.atoms_arr[:, :3] += self.mol_shift
            origin = np.zeros(3, dtype=np.float32)
        else:
            self.mol_shift = np.zeros(3, dtype=np.float32)
            origin = p0_raw
        self._set_spec(n, origin, [L[0] / nx, 0, 0], [0, L[1] / ny, 0], [0, 0, L[2] / nz],
                       D.dinv_axis_aligned(L, origin, (nx, ny, nz)))
        self.L = L
        return L

    def setup_grid_world(self, p0_xyz, L_xyz, n):
        assert self.atoms_arr is not None, "call assign_params() first"
        nx, ny, nz = [int(v) for v in n]
        p0 = np.asarray(p0_xyz, dtype=np.float32).reshape(3)
        L = np.asarray(L_xyz, dtype=np.float32).reshape(3)
        if np.any(L <= 0) or min(nx, ny, nz) < 2:
            raise ValueError(f"setup_grid_world: invalid p0={p0} L={L} n={n}")
        self.mol_shift = np.zeros(3, dtype=np.float32)
        self._set_spec((nx, ny, nz), p0, [L[0] / nx, 0, 0], [0, L[1] / ny, 0], [0, 0, L[2] / nz],
                       D.dinv_axis_aligned(L, p0, (nx, ny, nz)))
        self.L = L
        return L

    def setup_grid_lvec(self, n=(100, 100, 60), margin_z=2.0, z_top=12.0):
        assert self.atoms_arr is not None, "call assign_params() first"
        assert self.mol is not None and getattr(self.mol, "lvec", None) is not None, \
            "setup_grid_lvec requires mol.lvec (periodic system)"
        nx, ny, nz = [int(x) for x in n]
        if nx <= 1 or ny <= 1 or nz <= 1:
            raise ValueError(f"setup_grid_lvec: invalid n={n}")
        a = self.mol.lvec[0, :3].astype(np.float32)
        b = self.mol.lvec[1, :3].astype(np.float32)
        apos = self.atoms_arr[:, :3]
        zmin, zmax = float(apos[:, 2].min()), float(apos[:, 2].max())
        Lz = (zmax - zmin) + float(margin_z) + float(z_top)
        if Lz <= 0:
            raise ValueError(f"setup_grid_lvec: invalid Lz={Lz}")
        p0z = zmin - 0.5 * float(margin_z)
        self.mol_shift = np.array([0., 0., -p0z], dtype=np.float32)
        self.atoms_arr[:, :3] += self.mol_shift[None, :]
        M = np.stack([a, b, np.array([0., 0., Lz], dtype=np.float32)], axis=1).astype(np.float64)
        det = float(np.linalg.det(M))
        if abs(det) < 1e-12:
            raise ValueError(f"setup_grid_lvec: singular cell matrix det={det}")
        invMT = np.linalg.inv(M).T.astype(np.float32)
        self._set_spec((nx, ny, nz), np.zeros(3), a / nx, b / ny, [0., 0., Lz / nz], D.dinv_lvec(invMT, (nx, ny, nz)))
        self.L = np.array([np.linalg.norm(a), np.linalg.norm(b), Lz], dtype=np.float32)
        return self.L

    def set_pbc(self, npbc=(1, 1), lvA=None, lvB=None):
        """Periodic x,y images for make_forcefield: atoms + ix*lvA + iy*lvB, |ix|<=npbc[0], |iy|<=npbc[1].
        Default lattice vectors: the grid cell edges (n*dA, n*dB)."""
        self.npbc = (int(npbc[0]), int(npbc[1]))
        self.lvA = np.asarray(lvA if lvA is not None else self.dA[:3] * self.n[0], dtype=np.float32)
        self.lvB = np.asarray(lvB if lvB is not None else self.dB[:3] * self.n[1], dtype=np.float32)

    # ── force field (spammm/SPM/AFM.py:897-946) ─────────────────────────────

    def make_forcefield(self, bAlloc=True):
        import torch
        assert self.n is not None, "call setup_grid() first"
        nx, ny, nz = (int(v) for v in self.n)
        free = torch.cuda.mem_get_info(self.device)[0]
        need = nx * ny * nz * 16
        if need > free and (self.FF is None or self.FF.numel() * 4 != need):
            raise MemoryError(f"FF grid needs {need} bytes > free device memory {free}")
        atoms = torch.as_tensor(self.atoms_arr, dtype=torch.float32, device=self.device).contiguous()
        cpar = torch.as_tensor(self.cLJs_arr, dtype=torch.float32, device=self.device).contiguous()
        if self.FF is None or tuple(self.FF.shape) != (nz, ny, nx, 4):
            self.FF = torch.empty((nz, ny, nx, 4), dtype=torch.float32, device=self.device)
        model = _lib.PPM_MODEL_MORSE if self.use_morse else _lib.PPM_MODEL_LJ
        D.grid_project(atoms, cpar, model, self.spec, self.tipQs, self.tipQZs, self.npbc, self.lvA, self.lvB,
                       cap=self.cap, out=self.FF)
        torch.cuda.synchronize(self.device)

    def get_FF(self):
        """Host copy of the force field as (nx,ny,nz,4)."""
        assert self.FF is not None, "call make_forcefield() first"
        return self.FF.permute(2, 1, 0, 3).cpu().numpy()

    # ── scans ────────────────────────────────────────────────────────────────

    def _auto_scan(self, nx_s, ny_s):
        apos = self.atoms_arr[:, :3]
        mn, mx = apos.min(axis=0), apos.max(axis=0)
        x0 = mn[0] + (mx[0] - mn[0]) * 0.05
        y0 = mn[1] + (mx[1] - mn[1]) * 0.05
        z0 = float(mx[2]) + 5.0 + abs(float(self.dpos0[2]))
        scan_p0 = np.array([x0, y0, z0], dtype=np.float32)
        scan_da = np.array([(mx[0] - mn[0]) * 0.9 / max(nx_s - 1, 1), 0., 0.], dtype=np.float32)
        scan_db = np.array([0., (mx[1] - mn[1]) * 0.9 / max(ny_s - 1, 1), 0.], dtype=np.float32)
        return scan_p0, scan_da, scan_db

    @staticmethod
    def _check_osc(osc_dir):
        osc_n = np.asarray(osc_dir, dtype=np.float32)
        nrm = float(np.linalg.norm(osc_n))
        if osc_n.shape != (3,) or not np.isfinite(nrm) or nrm <= 0.0:
            raise ValueError(f"osc_dir must be a finite non-zero vector, got {osc_dir!r}")

    def _tip(self, dtip):
        tipC = self.tipC.copy()
        tipC[3] = np.float32(dtip)
        return D.Tip(self.tipA, self.tipB, tipC, self.stiffness, self.dpos0, self.relax_pars, self.surfFF)

    def run_scan(self, nxy=(50, 50), nz=60, dtip=-0.1, scan_p0=None, scan_da=None, scan_db=None, bAlloc=True,
                 osc_dir=(0., 0., 1.), check_before=False):
        """relaxStrokesTilted over an (nx,ny) raster; returns FEs (nx,ny,nz,4) and pts (nx,ny,3).
        iz=0 is the first (highest) tip position, as in the reference."""
        import torch
        assert self.FF is not None, "call make_forcefield() first"
        nx_s, ny_s = nxy
        if scan_p0 is None:
            scan_p0, scan_da, scan_db = self._auto_scan(nx_s, ny_s)
        self._check_osc(osc_dir)
        pts = D.scan_points(scan_p0, scan_da, scan_db, nx_s, ny_s)
        pts_d = torch.as_tensor(pts, device=self.device)
        FEs, PPs = D.relax_tilted(self.FF, self.spec, pts_d, int(nz), self._tip(dtip), self.max_iter, self.f2conv,
                                  check_before, scan_ny=ny_s, damped=not self.use_fire)
        self.scan_FEs_dev, self.scan_PPs_dev = FEs, PPs
        FEs_h = FEs.cpu().numpy().reshape(nx_s, ny_s, nz, 4)
        return FEs_h, pts[:, :3].reshape(nx_s, ny_s, 3)

    def get_PPs(self):
        """(nx*ny, nz, 4) probe positions and signed iteration counts of the last run_scan."""
        assert self.scan_PPs_dev is not None, "call run_scan() first"
        return self.scan_PPs_dev.cpu().numpy()

    def get_raw_FE(self, nxy=(60, 60), nz=21, dtip=-0.2, scan_p0=None, scan_da=None, scan_db=None, bAlloc=True,
                   osc_dir=(0., 0., 1.)):
        """getFEinStrokesTilted: unrelaxed FEs (nx,ny,nz,4) and pts (nx,ny,3) (spammm/SPM/AFM.py:2523-2580)."""
        import torch
        assert self.FF is not None, "call make_forcefield() first"
        nx_s, ny_s = nxy
        if scan_p0 is None:
            scan_p0, scan_da, scan_db = self._auto_scan(nx_s, ny_s)
        self._check_osc(osc_dir)
        pts = D.scan_points(scan_p0, scan_da, scan_db, nx_s, ny_s)
        pts_d = torch.as_tensor(pts, device=self.device)
        dTip = np.array([0., 0., dtip, 0.], dtype=np.float32)
        tip = D.Tip(self.tipA, self.tipB, self.tipC)
        FEs = D.fe_in_strokes_tilted(self.FF, self.spec, pts_d, int(nz), tip, dTip, self.dpos0)
        return FEs.cpu().numpy().reshape(nx_s, ny_s, nz, 4), pts[:, :3].reshape(nx_s, ny_s, 3)

    # ── gridless direct Morse path (spammm/SPM/AFM.py:2063-2519) ────────────

    def _direct_inputs(self, fn):
        assert self.use_morse, f"{fn} requires use_morse=True (call AFMulator(use_morse=True))"
        assert self.atoms_arr is not None and self.cLJs_arr is not None, f"{fn}: call assign_params() first"
        na = int(len(self.atoms_arr))
        if not (1 <= na <= _lib.PPM_DIRECT_MAX_ATOMS):
            raise ValueError(f"{fn}: nAtoms={na} out of range [1,{_lib.PPM_DIRECT_MAX_ATOMS}]; limit is "
                             f"{_lib.PPM_DIRECT_MAX_ATOMS} (no silent grid fallback)")
        assert self.atoms_arr.shape == (na, 4) and self.atoms_arr.dtype == np.float32, \
            f"atoms_arr must be (nAtoms,4) float32, got {self.atoms_arr.shape} {self.atoms_arr.dtype}"
        assert self.cLJs_arr.shape == (na, 4) and self.cLJs_arr.dtype == np.float32, \
            f"cLJs_arr (Morse) must be (nAtoms,4) float32, got {self.cLJs_arr.shape} {self.cLJs_arr.dtype}"
        return na

    def run_scan_morse_direct(self, nxy=(50, 50), nz=60, dtip=-0.1, scan_p0=None, scan_da=None, scan_db=None, *,
                              workgroup_size=64, bAlloc=True, return_pp=True):
        """Direct Morse + Coulomb relaxed scan without a grid (relaxStrokesTiltedMorseDirect,
        kernels/AFM.cl:1419-1511; spammm/SPM/AFM.py:2063-2233).  Returns FEs (nx,ny,nz,4) float32
        (iz=0 = highest tip, tip-rotated sample FE, uncapped), points_xyz (nx,ny,3) and, if
        return_pp, PPs (nx,ny,nz,4) = (position, +iterations / -iterations if not converged).
        Raises RuntimeError on any non-converged or non-finite state, naming the first (ix,iy,iz).
        ``workgroup_size`` is validated for interface parity; results do not depend on it."""
        import torch
        fn = "run_scan_morse_direct"
        na = self._direct_inputs(fn)
        if scan_p0 is None or scan_da is None or scan_db is None:
            raise ValueError(f"{fn}: scan_p0, scan_da, scan_db are MANDATORY (no bbox auto-derive during fitting)")
        scan_p0, scan_da, scan_db = (np.asarray(v, dtype=np.float32).reshape(3) for v in (scan_p0, scan_da, scan_db))
        assert np.isfinite(scan_p0).all() and np.isfinite(scan_da).all() and np.isfinite(scan_db).all(), \
            f"{fn}: scan_p0/da/db must be finite"
        if not float(dtip) < 0.0:
            raise ValueError(f"{fn}: dtip must be <0, got dtip={dtip}")
        try:
            nx_s, ny_s = int(nxy[0]), int(nxy[1])
        except Exception as e:
            raise ValueError(f"{fn}: nxy must be a 2-tuple of positive ints, got {nxy!r}") from e
        nz = int(nz)
        assert nx_s > 0 and ny_s > 0 and nz > 0, f"{fn}: nx,ny,nz must be positive, got {nxy} nz={nz}"
        assert int(workgroup_size) > 0, f"{fn}: workgroup_size must be positive, got {workgroup_size}"
        pts = D.scan_points(scan_p0, scan_da, scan_db, nx_s, ny_s)
        dev = self.device
        atoms = torch.as_tensor(self.atoms_arr, device=dev)
        cMs = torch.as_tensor(self.cLJs_arr, device=dev)
        FEs, PPs = D.relax_direct_morse(atoms, cMs, torch.as_tensor(pts, device=dev), nz, self._tip(dtip), self.tipQs,
                                        self.tipQZs, self.max_iter, self.f2conv, scan_ny=ny_s,
                                        damped=not self.use_fire)
        FEs_h = FEs.cpu().numpy().reshape(nx_s, ny_s, nz, 4)
        PPs_h = PPs.cpu().numpy().reshape(nx_s, ny_s, nz, 4)
        points_xyz = pts[:, :3].reshape(nx_s, ny_s, 3)
        bad = ~np.isfinite(PPs_h).all(axis=-1) | (PPs_h[..., 3] < 0.0)
        if bad.any():
            ix, iy, iz = np.unravel_index(int(np.argmax(bad)), bad.shape)
            pp = PPs_h[ix, iy, iz]
            reason = ("nonfinite" if not np.isfinite(pp).all() else f"non-converged (iters={int(abs(pp[3]))})")
            raise RuntimeError(f"{fn}: PP telemetry failed at (ix={ix},iy={iy},iz={iz}): {reason}  "
                               f"PP=({pp[0]:.6f},{pp[1]:.6f},{pp[2]:.6f}) PP.w={pp[3]:.1f}")
        self._morse_direct_fwd_state = {
            "nAtoms": na, "nx": nx_s, "ny": ny_s, "nz": nz, "dtip": float(dtip),
            "points": points_xyz.copy(), "atoms": self.atoms_arr.tobytes(), "cLJs": self.cLJs_arr.tobytes(),
            **{nm: getattr(self, nm).copy() for nm in ("tipA", "tipB", "stiffness", "dpos0", "tipQs", "tipQZs",
                                                        "surfFF")},
        }
        if return_pp:
            return FEs_h, points_xyz, PPs_h
        return FEs_h, points_xyz

    def vjp_scan_morse_direct(self, PPs, points_xyz, dtip, dL_dFEs, *, workgroup_size=64, bAlloc=True):
        """Implicit-equilibrium VJP of run_scan_morse_direct (kernels/AFM.cl:1711-2029;
        spammm/SPM/AFM.py:2241-2519): grad_theta (nAtoms,6) float64 in theta order
        (x,y,z,R0,E0,Q) and diagnostics {residual_norms, lambda_mins, condition_estimates,
        status_codes} (nx,ny,nz).  Raises ValueError on malformed input or non-converged PPs,
        RuntimeError on a stale forward state or any failed adjoint state."""
        import torch
        fn = "vjp_scan_morse_direct"
        na = self._direct_inputs(fn)
        PPs, points_xyz, dL_dFEs = np.asarray(PPs), np.asarray(points_xyz), np.asarray(dL_dFEs)
        if PPs.ndim != 4 or PPs.shape[3] != 4:
            raise ValueError(f"{fn}: PPs must be (nx,ny,nz,4), got {PPs.shape}")
        if PPs.dtype != np.float32:
            raise ValueError(f"{fn}: PPs must be float32, got {PPs.dtype}")
        if points_xyz.ndim != 3 or points_xyz.shape[2] != 3:
            raise ValueError(f"{fn}: points_xyz must be (nx,ny,3), got {points_xyz.shape}")
        if points_xyz.dtype != np.float32:
            raise ValueError(f"{fn}: points_xyz must be float32, got {points_xyz.dtype}")
        if dL_dFEs.shape != PPs.shape:
            raise ValueError(f"{fn}: dL_dFEs shape {dL_dFEs.shape} != PPs shape {PPs.shape}")
        if dL_dFEs.dtype != np.float32:
            raise ValueError(f"{fn}: dL_dFEs must be float32, got {dL_dFEs.dtype}")
        nx, ny, nz = PPs.shape[:3]
        if points_xyz.shape[:2] != (nx, ny):
            raise ValueError(f"{fn}: points_xyz ({points_xyz.shape[:2]}) != PPs ({nx},{ny})")
        if not np.isfinite(PPs).all():
            raise ValueError(f"{fn}: PPs contains nonfinite values")
        if not np.isfinite(dL_dFEs).all():
            raise ValueError(f"{fn}: dL_dFEs contains nonfinite values")
        if (PPs[..., 3] <= 0.0).any():
            ix, iy, iz = np.unravel_index(int(np.argmax(PPs[..., 3] <= 0.0)), (nx, ny, nz))
            raise ValueError(f"{fn}: PPs[{ix},{iy},{iz}].w={PPs[ix, iy, iz, 3]:.1f} <= 0 (non-converged); "
                             "VJP requires all converged states")
        if not float(dtip) < 0.0:
            raise ValueError(f"{fn}: dtip must be <0, got dtip={dtip}")
        assert int(workgroup_size) > 0, f"{fn}: workgroup_size must be positive, got {workgroup_size}"
        fwd = getattr(self, "_morse_direct_fwd_state", None)
        if fwd is None:
            raise RuntimeError(f"{fn}: no prior run_scan_morse_direct forward state recorded "
                               "(call run_scan_morse_direct first); cannot validate VJP inputs")
        bad = []
        if fwd["nAtoms"] != na:
            bad.append(f"nAtoms fwd={fwd['nAtoms']} vs vjp={na}")
        if (fwd["nx"], fwd["ny"], fwd["nz"]) != (nx, ny, nz):
            bad.append(f"scan shape fwd=({fwd['nx']},{fwd['ny']},{fwd['nz']}) vs vjp=({nx},{ny},{nz})")
        if abs(fwd["dtip"] - float(dtip)) > 1e-12:
            bad.append(f"dtip fwd={fwd['dtip']!r} vs vjp={float(dtip)!r}")
        if fwd["points"].shape != points_xyz.shape or not np.array_equal(fwd["points"], points_xyz):
            bad.append("points_xyz does not match the forward scan raster")
        if self.atoms_arr.tobytes() != fwd["atoms"]:
            bad.append("atoms_arr content changed since forward")
        if self.cLJs_arr.tobytes() != fwd["cLJs"]:
            bad.append("cLJs_arr content changed since forward")
        for nm in ("tipA", "tipB", "stiffness", "dpos0", "tipQs", "tipQZs", "surfFF"):
            if not np.array_equal(getattr(self, nm), fwd[nm]):
                bad.append(f"{nm} changed since forward")
        if bad:
            raise RuntimeError(f"{fn}: stale or mismatched forward state:\n  - " + "\n  - ".join(bad))
        dev = self.device
        pts = np.zeros((nx * ny, 4), dtype=np.float32)
        pts[:, :3] = points_xyz.reshape(-1, 3)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)  # noqa: E731
        grad, lam, diag = D.direct_morse_vjp(t(self.atoms_arr), t(self.cLJs_arr), t(pts), nz, self._tip(dtip),
                                             self.tipQs, self.tipQZs, t(PPs.reshape(-1, nz, 4)),
                                             t(dL_dFEs.reshape(-1, nz, 4)))
        diag_h = diag.cpu().numpy().reshape(nx, ny, nz, 4)
        if not np.isfinite(diag_h).all():
            ix, iy, iz = np.unravel_index(int(np.argmax(~np.isfinite(diag_h).all(axis=-1))), (nx, ny, nz))
            raise RuntimeError(f"{fn}: adjoint_diag nonfinite at (ix={ix},iy={iy},iz={iz}): diag={diag_h[ix, iy, iz]}")
        status = diag_h[..., 3].astype(np.int64)
        if (status != 0).any():
            ix, iy, iz = np.unravel_index(int(np.argmax(status != 0)), (nx, ny, nz))
            d = diag_h[ix, iy, iz]
            raise RuntimeError(f"{fn}: adjoint failed at (ix={ix},iy={iy},iz={iz}): status_code={int(d[3])} "
                               f"residual_norm={d[0]:.6e} lambda_min={d[1]:.6e} condition_estimate={d[2]:.6e}")
        g = grad.cpu().numpy()
        if not np.isfinite(g).all():
            raise RuntimeError(f"{fn}: grad_theta nonfinite at atom {int(np.argmax(~np.isfinite(g).all(-1)))}")
        diagnostics = {"residual_norms": diag_h[..., 0].astype(np.float64),
                       "lambda_mins": diag_h[..., 1].astype(np.float64),
                       "condition_estimates": diag_h[..., 2].astype(np.float64), "status_codes": status}
        return g, diagnostics

    @property
    def mol_z(self):
        if self.atoms_arr is None:
            return 0.0
        return float(self.atoms_arr[:, 2].max())

    # ── externally supplied force field (spammm/SPM/AFM.py:387-642) ─────────

    def setup_fdbm_grid(self, F_total, origin, step):
        """Upload an (nx,ny,nz,4) force field (Fx,Fy,Fz,E) for scan_fdbm."""
        import torch
        F_total = np.asarray(F_total, dtype=np.float32)
        nx, ny, nz = F_total.shape[:3]
        L = np.array([nx * step, ny * step, nz * step], dtype=np.float32)
        origin = np.asarray(origin, dtype=np.float32)
        self.img_FF_fdbm = torch.as_tensor(np.ascontiguousarray(F_total.transpose(2, 1, 0, 3)),
                                           device=self.device)
        self.fdbm_spec = D.axis_aligned_spec(origin, L, (nx, ny, nz))
        self.fdbm_origin, self.fdbm_step, self.fdbm_shape = origin, float(step), (nx, ny, nz)

    def install_fdbm_grid(self, grid, origin, step, shape):
        """Use a device force grid (nz,ny,nx,4) (e.g. from fdbm.stage3_fdbm_fields_fast) as the FDBM
        grid without a host round trip (setup_fdbm_grid_from_img, spammm/SPM/AFM.py:413-429)."""
        nx, ny, nz = (int(v) for v in shape[:3])
        D._f32dev(grid, "grid", 4)
        if grid.numel() != 4 * nx * ny * nz:
            raise ValueError("grid does not match shape")
        L = np.array([nx * step, ny * step, nz * step], dtype=np.float32)
        origin = np.asarray(origin, dtype=np.float32)
        self.img_FF_fdbm = grid
        self.fdbm_spec = D.axis_aligned_spec(origin, L, (nx, ny, nz))
        self.fdbm_origin, self.fdbm_step, self.fdbm_shape = origin, float(step), (nx, ny, nz)

    def setup_fdbm_grid_from_img(self, img_F, shape, origin, step):
        """Device-side copy of a force grid (nz,ny,nx,4) (e.g. the make_forcefield grid) into the FDBM
        grid for scan_fdbm (setup_fdbm_grid_from_img, spammm/SPM/AFM.py:413-429)."""
        D._f32dev(img_F, "img_F", 4)
        self.install_fdbm_grid(img_F.clone(), origin, step, shape)

    def compute_gradient_cl(self, E_field, step, bAlloc=True):
        """(nx,ny,nz,4) = (-grad E, E) of a host (nx,ny,nz) energy by periodic central differences
        on the GPU (spammm/SPM/AFM.py:2740-2790, gradient_central_diff kernels/AFM.cl:1151-1196)."""
        import torch
        from . import fdbm
        E = torch.as_tensor(np.ascontiguousarray(E_field, dtype=np.float32), device=self.device)
        grid = fdbm.compose_gradient((E,), float(step))
        return np.ascontiguousarray(grid.cpu().numpy().transpose(2, 1, 0, 3))

    def scan_fdbm(self, scan_xs, scan_ys, probe_heights, mol_z=0.0, ppm_mode=True, use_fire=True, K_LAT=0.03,
                  K_RAD=20.0, bond_length=3.0, stiffness=None, dpos0=None, relax_pars=None, osc_dir=(0., 0., 1.),
                  base_pos=(0., 0., 0.), reverse=False):
        """relaxStrokes over scan_xs x scan_ys; FEs (nx,ny,nz,4) with iz=0 the lowest height, and tip_disp.
        ``use_fire`` chooses the default relax_pars only, as in the reference (spammm/SPM/AFM.py:572-580):
        the instance's (FIRE) or (dt, damp) = (0.3, 0.2) of the damped CPU twin; the integrator itself is
        the one the instance was built with (AFMulator(use_fire=...))."""
        import torch
        assert self.img_FF_fdbm is not None, "call setup_fdbm_grid() first"
        nx_s, ny_s, nz_s = len(scan_xs), len(scan_ys), len(probe_heights)
        self._check_osc(osc_dir)
        base = np.asarray(base_pos, dtype=np.float32)
        if base.shape != (3,) or not np.isfinite(base).all():
            raise ValueError(f"base_pos must contain three finite coordinates, got {base_pos!r}")
        stiffness_v = (np.array([-K_LAT, -K_LAT, -K_LAT, -K_RAD], dtype=np.float32) if stiffness is None
                       else np.array(stiffness, dtype=np.float32))
        dpos0_v = (np.array([0., 0., -bond_length, bond_length], dtype=np.float32) if dpos0 is None
                   else np.array(dpos0, dtype=np.float32))
        if relax_pars is not None:
            relax_v = np.array(relax_pars, dtype=np.float32)
        elif use_fire:
            relax_v = self.relax_pars.copy()
        else:
            relax_v = np.array([0.3, 0.2, 0.03, 0.3], dtype=np.float32)
        hs = np.sort(probe_heights)
        dh = float(hs[-1] - hs[0]) / max(nz_s - 1, 1) if nz_s > 1 else 0.
        if reverse:
            dTip = np.array([0., 0., +dh, 0.], dtype=np.float32)
            z_start = float(hs[0]) + mol_z + bond_length + float(base[2])
        else:
            dTip = np.array([0., 0., -dh, 0.], dtype=np.float32)
            z_start = float(hs[-1]) + mol_z + bond_length + float(base[2])
        gx, gy = np.meshgrid(scan_xs, scan_ys, indexing="ij")
        pts = np.zeros((nx_s * ny_s, 4), dtype=np.float32)
        pts[:, 0] = gx.ravel() + base[0]
        pts[:, 1] = gy.ravel() + base[1]
        pts[:, 2] = z_start
        pts_d = torch.as_tensor(pts, device=self.device)
        FEs, disps, _ = D.relax_strokes(self.img_FF_fdbm, self.fdbm_spec, pts_d, nz_s, dTip, stiffness_v, dpos0_v,
                                        relax_v, self.max_iter, self.f2conv, scan_ny=ny_s, want_pps=False,
                                        damped=not self.use_fire)
        FEs = FEs.cpu().numpy().reshape(nx_s, ny_s, nz_s, 4)
        disps = disps.cpu().numpy().reshape(nx_s, ny_s, nz_s, 4)
        if not reverse:
            FEs, disps = FEs[:, :, ::-1, :], disps[:, :, ::-1, :]
        return FEs, {"dx": disps[..., 0], "dy": disps[..., 1], "dz": disps[..., 2]}

    def scan_fdbm_2d(self, scan_xs, scan_ys, probe_heights, mol_z=0.0, K_LAT=0.03, dt=0.3, damp=0.2):
        """relaxStrokes2D over scan_xs x scan_ys (spammm/SPM/AFM.py:644-701): lateral-only damped
        relaxation at fixed probe heights; FEs (nx,ny,nz,4) with iz=0 the lowest height, and tip_disp
        (dz = 0)."""
        import torch
        assert self.img_FF_fdbm is not None, "call setup_fdbm_grid() first"
        nx_s, ny_s, nz_s = len(scan_xs), len(scan_ys), len(probe_heights)
        hd = np.sort(probe_heights)[::-1]
        dh = float(hd[0] - hd[-1]) / max(nz_s - 1, 1) if nz_s > 1 else 0.
        gx, gy = np.meshgrid(scan_xs, scan_ys, indexing="ij")
        pts = np.zeros((nx_s * ny_s, 4), dtype=np.float32)
        pts[:, 0] = gx.ravel()
        pts[:, 1] = gy.ravel()
        pts[:, 2] = float(hd[0]) + mol_z
        pts_d = torch.as_tensor(pts, device=self.device)
        FEs, disps, _ = D.relax_strokes_2d(self.img_FF_fdbm, self.fdbm_spec, pts_d, nz_s, K_LAT, dh, dt, damp,
                                           self.max_iter, self.f2conv, scan_ny=ny_s)
        FEs = FEs.cpu().numpy().reshape(nx_s, ny_s, nz_s, 4)[:, :, ::-1, :]
        disps = disps.cpu().numpy().reshape(nx_s, ny_s, nz_s, 4)[:, :, ::-1, :]
        return FEs, {"dx": disps[..., 0], "dy": disps[..., 1], "dz": disps[..., 2]}

    # ── frequency shift ──────────────────────────────────────────────────────

    def get_df(self, FEs, k_tip=1800.0):
        """-mean_z(np.gradient(Fz)) / (2 k_tip), shape (nx,ny) (spammm/SPM/AFM.py:2591-2599)."""
        Fz = np.asarray(FEs)[..., 2]
        g = _apply_np("gradient", Fz, 1.0)       # -np.gradient(Fz) with unit spacing
        return g.mean(axis=2) / (2.0 * k_tip)


# ──────────────────────────────────────────────────────────────────────────────
# module-level df functions (spammm/SPM/AFM.py:3048-3277), computed on the GPU
# ──────────────────────────────────────────────────────────────────────────────

def _dev():
    return D.require_gpu()


def _apply_np(kind, Fz, dz, amp=1.0, f0=30000.0, k=1800.0):
    """Apply a z operator to a host (…, nz) array on the GPU; returns float64 host array."""
    import torch
    Fz = np.asarray(Fz)
    t = torch.as_tensor(np.ascontiguousarray(Fz, dtype=np.float32), device=_dev())
    out = _df.apply_z(kind, t, dz, amp, f0, k, channel=None)
    return out.cpu().numpy().astype(np.float64)


def compute_df(Fz, dz):
    """df = -dFz/dz (np.gradient semantics along axis 2, spacing |dz|)."""
    return _apply_np("gradient", Fz, dz)


def compute_df_amp(Fz, dz, amp=1.0):
    """Finite-amplitude mean of -dFz/dz with 9 equal-weight Gauss-Chebyshev nodes (arcsine weight)."""
    return _apply_np("amp9", Fz, dz, amp).astype(np.float32)


def df_giessibl(Fz, dz, amp=1.0, f0=30000.0, k=1800.0):
    """Giessibl frequency shift [Hz] along axis 2; ``dz`` signed (negative for a descending scan)."""
    return _apply_np("giessibl", Fz, dz, amp, f0, k).astype(np.float32)


def _dir_spacing(spacing, osc_dir):
    n = np.asarray(osc_dir, dtype=np.float64)
    if n.shape != (3,) or not np.isfinite(n).all():
        raise ValueError(f"osc_dir must contain three finite components, got {osc_dir!r}")
    norm = float(np.linalg.norm(n))
    if norm <= 0.0:
        raise ValueError(f"osc_dir must be non-zero, got {osc_dir!r}")
    n = n / norm
    dr = (np.full(3, abs(float(spacing)), dtype=np.float64) if np.ndim(spacing) == 0
          else np.abs(np.asarray(spacing, dtype=np.float64)))
    if dr.shape != (3,) or not np.isfinite(dr).all() or np.any(dr <= 0.0):
        raise ValueError(f"spacing must be a positive scalar or (dx,dy,dz), got {spacing!r}")
    return n, dr


def _axis_layout(shape, axis):
    """(n_outer, n_axis, n_inner) and strides (in elements) of a C-ordered (nx,ny,nz) volume."""
    nx, ny, nz = shape
    if axis == 0:
        return 1, nx, ny * nz
    if axis == 1:
        return nx, ny, nz
    return nx * ny, nz, 1


def _volume_op(FEs_d, chw, terms, shape):
    """sum over terms of alpha * (ops applied along z, then y, then x) to v = sum_c chw[c]*FEs[..., c].

    ``terms``: list of ([op along x or None, op along y or None, op along z or None], alpha).
    The first operator of a term reads the (nx,ny,nz,4) volume with channel weights ``chw``;
    later operators read the contiguous (nx,ny,nz) intermediate; the last one accumulates."""
    import torch
    nx, ny, nz = shape
    acc = torch.zeros((nx, ny, nz), dtype=torch.float32, device=FEs_d.device)
    for ops, alpha in terms:
        order = [a for a in (2, 1, 0) if ops[a] is not None]
        cur, cur_chw, cur_nch, c = FEs_d, chw, 4, 4
        for k, axis in enumerate(order):
            outer, nax, inner = _axis_layout(shape, axis)
            last = k == len(order) - 1
            res = ops[axis].apply(cur, (nax * inner * c, inner * c, c), outer, inner, chw=cur_chw, nch=cur_nch,
                                  out=acc.view(outer, nax, inner) if last else None,
                                  alpha=alpha if last else 1.0, accumulate=last)
            if not last:
                cur, cur_chw, cur_nch, c = res.view(nx, ny, nz), (1.0, 0, 0, 0), 1, 1
    return acc


def compute_df_dir(FEs, spacing, osc_dir=(0., 0., 1.)):
    """df = -n.grad(F.n) on an xyz force volume (spammm/SPM/AFM.py:3105-3123); z stays the slice axis."""
    import torch
    FEs = np.asarray(FEs)
    if FEs.ndim != 4 or FEs.shape[-1] < 3:
        raise ValueError(f"FEs must have shape (nx,ny,nz,>=3), got {FEs.shape}")
    n, dr = _dir_spacing(spacing, osc_dir)
    shape = FEs.shape[:3]
    for ia in np.flatnonzero(np.abs(n) > 1e-15):
        if shape[ia] < 2:
            raise ValueError(f"FEs axis {ia} needs at least 2 samples for osc_dir={tuple(n)}")
    F4 = np.zeros(shape + (4,), dtype=np.float32)
    F4[..., :FEs.shape[-1]] = FEs[..., :4]
    dev = _dev()
    Fd = torch.as_tensor(F4, device=dev)
    chw = (float(n[0]), float(n[1]), float(n[2]), 0.0)
    terms = []
    for ia in np.flatnonzero(np.abs(n) > 1e-15):
        ops = [None, None, None]
        ops[ia] = _df.AxisOperator(_df.gradient_matrix(shape[ia], dr[ia]), dev)
        terms.append((ops, -float(n[ia])))
    out = _volume_op(Fd, chw, terms, shape)
    return out.cpu().numpy().astype(np.float64)


def compute_df_amp_dir(FEs, spacing, osc_dir=(0., 0., 1.), amp=1.0):
    """Finite-amplitude directional df (spammm/SPM/AFM.py:3203-3236): 9 equal-weight Gauss-Chebyshev
    nodes of -n.grad(F.n) sampled at r + u_k*amp*n (trilinear, clamped), computed as separable
    per-axis shift/gradient operators on the GPU."""
    import torch
    FEs = np.asarray(FEs)
    if FEs.ndim != 4 or FEs.shape[-1] < 3:
        raise ValueError(f"FEs must have shape (nx,ny,nz,>=3), got {FEs.shape}")
    n, dr = _dir_spacing(spacing, osc_dir)
    shape = FEs.shape[:3]
    active = np.flatnonzero(np.abs(n) > 1e-15)
    for ia in active:
        if shape[ia] < 2:
            raise ValueError(f"FEs axis {ia} needs at least 2 samples for osc_dir={tuple(n)}")
    F4 = np.zeros(shape + (4,), dtype=np.float32)
    F4[..., :FEs.shape[-1]] = FEs[..., :4]
    dev = _dev()
    Fd = torch.as_tensor(F4, device=dev)
    chw = (float(n[0]), float(n[1]), float(n[2]), 0.0)
    u, w = _df.gc9()
    terms = []
    for uk, wk in zip(u, w):
        shifts = [uk * amp * n[a] / dr[a] for a in range(3)]
        for ia in active:
            ops = [None, None, None]
            for a in range(3):
                S = _df.shift_matrix(shape[a], shifts[a]) if abs(shifts[a]) > 0 else None
                if a == ia:
                    G = _df.gradient_matrix(shape[a], dr[a])
                    ops[a] = _df.AxisOperator(G @ S if S is not None else G, dev)
                elif S is not None:
                    ops[a] = _df.AxisOperator(S, dev)
            terms.append((ops, -float(wk * n[ia])))
    out = _volume_op(Fd, chw, terms, shape)
    return out.cpu().numpy().astype(np.float32)


def df_loss_seed(FEs, target_df, dtip, weights=None):
    """Central-difference df loss and its exact adjoint (spammm/SPM/AFM.py:3125-3200).
    Returns (loss, df (nx,ny,nz) float64, dL_dFEs (nx,ny,nz,4) float64, only .z nonzero)."""
    import torch
    FEs = np.asarray(FEs)
    target_df = np.asarray(target_df, dtype=np.float64)
    nx, ny, nz = FEs.shape[:3]
    if FEs.ndim != 4 or FEs.shape[3] < 3:
        raise ValueError(f"df_loss_seed: FEs must be (nx,ny,nz,>=4), got {FEs.shape}")
    if target_df.shape != (nx, ny, nz):
        raise ValueError(f"df_loss_seed: target_df must be ({nx},{ny},{nz}), got {target_df.shape}")
    dtip = float(dtip)
    if dtip >= 0.0:
        raise ValueError(f"df_loss_seed: dtip must be <0, got dtip={dtip}")
    if nz < 3:
        raise ValueError(f"df_loss_seed: need nz>=3 for central difference, got nz={nz}")
    if weights is None:
        w = np.ones((nx, ny, nz))
        w[:, :, 0] = 0.0
        w[:, :, -1] = 0.0
    else:
        w = np.asarray(weights, dtype=np.float64)
        if w.shape != (nx, ny, nz):
            raise ValueError(f"df_loss_seed: weights must be ({nx},{ny},{nz}), got {w.shape}")
    wsum = float(w.sum())
    if wsum <= 0.0:
        raise ValueError("df_loss_seed: sum(weights) must be > 0")
    dev = _dev()
    Fz = torch.as_tensor(np.ascontiguousarray(FEs[..., 2], dtype=np.float32), device=dev)
    df = _df.apply_z("central", Fz, dtip, channel=None)
    wt = torch.as_tensor(w, dtype=torch.float64, device=dev)
    diff = df.double() - torch.as_tensor(target_df, device=dev)
    loss = 0.5 * float((wt * diff * diff).sum()) / wsum
    g = (wt * diff / wsum).float().contiguous()
    opT = _df.z_operator("central_seed", nz, dtip, dev)
    dFz = opT.apply(g, (nz, 1, 0), nx * ny, 1).reshape(nx, ny, nz)
    dL = np.zeros((nx, ny, nz, 4), dtype=np.float64)
    dL[..., 2] = dFz.cpu().numpy()
    return loss, df.cpu().numpy().astype(np.float64), dL


def compute_dissipation(FEs_fwd, FEs_bwd, h_scan, h_df, amp, osc_dir=(0., 0., 1.)):
    """|trapezoid integral of (F_fwd - F_bwd).n over [h-A_z, h+A_z]| per display height
    (spammm/SPM/AFM.py:3238-3277), evaluated as a banded z operator on the GPU."""
    import torch
    FEs_fwd = np.asarray(FEs_fwd, dtype=np.float32)
    FEs_bwd = np.asarray(FEs_bwd, dtype=np.float32)
    h_scan = np.asarray(h_scan, dtype=np.float64)
    h_df = np.asarray(h_df, dtype=np.float64)
    n = np.asarray(osc_dir, dtype=np.float64)
    n = n / np.linalg.norm(n)
    amp_z = float(amp) * abs(n[2])
    nx, ny, nzs = FEs_fwd.shape[:3]
    M = np.zeros((len(h_df), nzs))
    for i, h in enumerate(h_df):
        mask = (h_scan >= h - amp_z) & (h_scan <= h + amp_z)
        if mask.sum() < 2:
            mask = np.ones(nzs, dtype=bool)
        idx = np.flatnonzero(mask)
        xs = h_scan[idx]
        for a in range(len(idx) - 1):                        # trapezoid rule weights
            hseg = xs[a + 1] - xs[a]
            M[i, idx[a]] += 0.5 * hseg
            M[i, idx[a + 1]] += 0.5 * hseg
    dev = _dev()
    dF = torch.as_tensor(np.ascontiguousarray(FEs_fwd[..., :4] - FEs_bwd[..., :4]), device=dev)
    op = _df.AxisOperator(M, dev)
    chw = (float(n[0]), float(n[1]), float(n[2]), 0.0)
    out = op.apply(dF, (nzs * 4, 4, 0), nx * ny, 1, chw=chw, nch=3)
    return out.abs().reshape(nx, ny, len(h_df)).cpu().numpy().astype(np.float32)
