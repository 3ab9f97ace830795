"""Synthetic workloads of the benchmark configurations (seeded, deterministic).

* ``naphthalene`` — planar molecule of two fused hexagons (C-C 1.40 A), 10 C +
  8 H (C-H 1.08 A), z=0, charges U(-0.15,0.15) e shifted to net 0; periodic
  20x20x20 A cell on a 0.1 A grid (200^3); scan 100x100 at 0.2 A; tip z 10 -> 6.1 A
  in 40 steps of 0.1 A (probe 4 A below the tip); CO-like probe R0=1.66 A,
  E0=0.0091 eV, q=-0.05 e; k_lat 0.25 N/m, k_rad 20 N/m; FIRE |F|<1e-4 eV/A,
  <=100 iterations; df with A=1 A, f0=30 kHz, k=1800 N/m.
* ``flake`` — flat polycyclic sp2 flake: honeycomb lattice (1.42 A) grown by
  breadth-first search from the centre, 10% of the heavy atoms N or O, H on edge
  atoms (1.08 A), 200 atoms in total, charges U(-0.3,0.3) e neutralised, z jitter
  N(0,0.05 A); periodic 30x30x24 A cell on a 0.1 A grid (300x300x240); scan
  256x256 over the cell, 60 z-steps of 0.1 A.
* ``batch`` — molecules with U{10..150} atoms, square cells U(16,32) A x 20 A,
  0.1 A grids, 128x128 pixels, 40 z-steps.
* ``substrate_scene`` — a 60-atom flat molecule 3 A above a 4-layer rock-salt slab
  (8x8 cells, a = 5.64 A, +-0.8 e alternating, Na/Cl parameters, periodic in x,y) on a
  0.05 A grid (902x902x300), 64 placements (shift U(+-2 A), rotation U(0, 2 pi) about z),
  192x192 pixels x 60 z-steps.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .params import ELEMENTS_BENCH, mix_lj, stiffness_Nm_to_eVA2

PROBE_R0 = 1.66
PROBE_E0 = 0.0091
PROBE_Q = -0.05


@dataclass
class Workload:
    name: str
    enames: list
    atoms: np.ndarray          # (na,4) float32 x,y,z,q
    cLJ: np.ndarray            # (na,2) float32 c6,c12
    n: tuple                   # grid (nx,ny,nz)
    p0: np.ndarray             # grid origin (3,)
    L: np.ndarray              # cell lengths (3,)
    npbc: tuple                # periodic image range along x,y
    scan_p0: np.ndarray
    scan_da: np.ndarray
    scan_db: np.ndarray
    nxy: tuple
    nz: int
    dtip: float = -0.1
    Qs: np.ndarray = field(default_factory=lambda: np.array([PROBE_Q, 0, 0, 0], np.float32))
    QZs: np.ndarray = field(default_factory=lambda: np.zeros(4, np.float32))
    k_lat_Nm: float = 0.25
    k_rad_Nm: float = 20.0
    bond: float = 4.0
    max_iter: int = 100
    f2conv: float = 1e-8
    amp: float = 1.0
    f0: float = 30000.0
    k_cant: float = 1800.0

    @property
    def stiffness(self):
        kl, kr = stiffness_Nm_to_eVA2(self.k_lat_Nm), stiffness_Nm_to_eVA2(self.k_rad_Nm)
        return np.array([-kl, -kl, -kl, -kr], dtype=np.float32)

    @property
    def dpos0(self):
        return np.array([0., 0., -self.bond, self.bond], dtype=np.float32)

    @property
    def lvA(self):
        return np.array([self.L[0], 0., 0.], dtype=np.float32)

    @property
    def lvB(self):
        return np.array([0., self.L[1], 0.], dtype=np.float32)

    @property
    def n_scan(self):
        return int(self.nxy[0]) * int(self.nxy[1])

    @property
    def nvox(self):
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    @property
    def pair_evals(self):
        return self.nvox * self.atoms.shape[0] * (2 * self.npbc[0] + 1) * (2 * self.npbc[1] + 1)


def _lj_params(enames):
    return mix_lj(enames, ELEMENTS_BENCH, PROBE_R0, PROBE_E0)


def naphthalene_geometry():
    """10 C + 8 H of naphthalene in the z=0 plane, centred on the origin."""
    b = 1.40
    s3 = math.sqrt(3.0)
    # two hexagons sharing the vertical edge at x=0
    C = [(0.0, b / 2), (0.0, -b / 2)]
    for sgn in (-1.0, 1.0):
        cx = sgn * b * s3 / 2
        C += [(cx + sgn * b * s3 / 2, b / 2), (cx + sgn * b * s3 / 2, -b / 2), (cx, b), (cx, -b)]
    C = np.array(C)
    H = []
    for x, y in C:
        if abs(x) < 1e-9:
            continue   # fused carbons carry no H
        ctr = np.array([math.copysign(b * s3 / 2, x), 0.0])
        d = np.array([x, y]) - ctr
        d /= np.linalg.norm(d)
        H.append(np.array([x, y]) + 1.08 * d)
    pos = np.zeros((18, 3))
    pos[:10, :2] = C
    pos[10:, :2] = np.array(H)
    return ["C"] * 10 + ["H"] * 8, pos


def naphthalene(seed=0) -> Workload:
    rng = np.random.default_rng(seed)
    en, pos = naphthalene_geometry()
    q = rng.uniform(-0.15, 0.15, len(en))
    q -= q.mean()
    atoms = np.zeros((len(en), 4), np.float32)
    atoms[:, :3] = pos
    atoms[:, 3] = q
    return Workload("naphthalene", en, atoms, _lj_params(en), (200, 200, 200),
                    np.array([-10., -10., -10.], np.float32), np.array([20., 20., 20.], np.float32), (1, 1),
                    scan_p0=np.array([-9.9, -9.9, 10.0], np.float32), scan_da=np.array([0.2, 0., 0.], np.float32),
                    scan_db=np.array([0., 0.2, 0.], np.float32), nxy=(100, 100), nz=40)


def flake_geometry(n_total, rng, sub_frac=0.1, jitter=0.05, a=1.42):
    """Honeycomb flake grown by BFS, edge H, exactly ``n_total`` atoms; heavy atoms C/N/O."""
    s3 = math.sqrt(3.0)
    a1 = np.array([1.5 * a, s3 / 2 * a])
    a2 = np.array([1.5 * a, -s3 / 2 * a])
    nbr_A = [(0, 0, 1), (-1, 0, 1), (0, -1, 1)]     # A(i,j) -> B sites
    nbr_B = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]       # B(i,j) -> A sites

    def xy(i, j, s):
        return i * a1 + j * a2 + (np.array([a, 0.0]) if s else np.zeros(2))

    def neigh(site):
        i, j, s = site
        tab = nbr_A if s == 0 else nbr_B
        return [(i + di, j + dj, t) for di, dj, t in tab]

    order, seen, frontier = [], {(0, 0, 0)}, [(0, 0, 0)]
    while frontier and len(order) < 4 * n_total:
        nxt = []
        for st in frontier:
            order.append(st)
            for nb in neigh(st):
                if nb not in seen:
                    seen.add(nb)
                    nxt.append(nb)
        nxt.sort(key=lambda t: np.linalg.norm(xy(*t)))
        frontier = nxt
    order.sort(key=lambda t: (round(float(np.linalg.norm(xy(*t))), 6), t))

    def count(nh):
        heavy = set(order[:nh])
        nH = sum(3 - sum(nb in heavy for nb in neigh(st)) for st in heavy)
        return nh + nH, heavy

    best = None
    for nh in range(1, len(order)):
        tot, heavy = count(nh)
        if tot >= n_total:
            best = (nh, heavy)
            break
    nh, heavy = best
    heavy_list = order[:nh]
    pos, en = [], []
    for st in heavy_list:
        pos.append(xy(*st))
        en.append("C")
    Hs = []
    for st in heavy_list:
        p = xy(*st)
        for nb in neigh(st):
            if nb not in heavy:
                d = xy(*nb) - p
                d /= np.linalg.norm(d)
                Hs.append(p + 1.08 * d)
    Hs = Hs[: max(0, n_total - nh)]
    pos += Hs
    en += ["H"] * len(Hs)
    pos = np.array(pos[:n_total])
    en = en[:n_total]
    heavy_idx = np.arange(min(nh, n_total))
    nsub = int(round(sub_frac * len(heavy_idx)))
    sub = rng.choice(heavy_idx, size=nsub, replace=False)
    for k, i in enumerate(sub):
        en[i] = "N" if k % 2 == 0 else "O"
    P = np.zeros((len(en), 3))
    P[:, :2] = pos - pos.mean(axis=0)
    P[:, 2] = rng.normal(0.0, jitter, len(en))
    return en, P


def flake(seed=1, n_atoms=200) -> Workload:
    rng = np.random.default_rng(seed)
    en, pos = flake_geometry(n_atoms, rng)
    q = rng.uniform(-0.3, 0.3, len(en))
    q -= q.mean()
    atoms = np.zeros((len(en), 4), np.float32)
    atoms[:, :3] = pos
    atoms[:, 3] = q
    nxy = (256, 256)
    step = 30.0 / 256
    return Workload("flake", en, atoms, _lj_params(en), (300, 300, 240),
                    np.array([-15., -15., -8.], np.float32), np.array([30., 30., 24.], np.float32), (1, 1),
                    scan_p0=np.array([-15. + step / 2, -15. + step / 2, 12.0], np.float32),
                    scan_da=np.array([step, 0., 0.], np.float32), scan_db=np.array([0., step, 0.], np.float32),
                    nxy=nxy, nz=60)


def flake_differentiable(tip_z0=10.0, seed=1) -> Workload:
    """The flake on its 300x300x240 grid scanned at 128x128 pixels x 40 z-steps from tip height
    tip_z0 (10 A: probe 6 -> 2.1 A), the differentiable-image workload."""
    w = flake(seed)
    step = 30.0 / 128
    w.name = "flake_differentiable"
    w.nxy, w.nz = (128, 128), 40
    w.scan_p0 = np.array([-15. + step / 2, -15. + step / 2, tip_z0], np.float32)
    w.scan_da = np.array([step, 0., 0.], np.float32)
    w.scan_db = np.array([0., step, 0.], np.float32)
    return w


def perturbed_positions(w: Workload, sigma=0.1, seed=3) -> np.ndarray:
    """(na, 3) float64 atom positions displaced by N(0, sigma) per coordinate (the target of the
    differentiable workload)."""
    rng = np.random.default_rng(seed)
    return w.atoms[:, :3].astype(np.float64) + rng.normal(0, sigma, w.atoms[:, :3].shape)


def batch_molecule(i: int, base_seed=100) -> Workload:
    rng = np.random.default_rng(base_seed + i)
    na = int(rng.integers(10, 151))
    Lxy = float(rng.uniform(16.0, 32.0))
    en, pos = flake_geometry(na, rng)
    q = rng.uniform(-0.3, 0.3, len(en))
    q -= q.mean()
    atoms = np.zeros((len(en), 4), np.float32)
    atoms[:, :3] = pos
    atoms[:, 3] = q
    n = int(round(Lxy / 0.1))
    L = np.array([n * 0.1, n * 0.1, 20.0], np.float32)
    step = float(L[0]) / 128
    return Workload(f"batch{i}", en, atoms, _lj_params(en), (n, n, 200),
                    np.array([-L[0] / 2, -L[1] / 2, -6.0], np.float32), L, (1, 1),
                    scan_p0=np.array([-L[0] / 2 + step / 2, -L[1] / 2 + step / 2, 10.0], np.float32),
                    scan_da=np.array([step, 0., 0.], np.float32), scan_db=np.array([0., step, 0.], np.float32),
                    nxy=(128, 128), nz=40)


def batch(n_mol=512, base_seed=100):
    """The batched configuration: ``n_mol`` molecules (see ``batch_molecule``)."""
    return [batch_molecule(i, base_seed) for i in range(n_mol)]


def small_batch_molecule(i: int, base_seed=500) -> Workload:
    """A reduced batch member (coarse 0.25 A grid, small scan) that the oracle checks in seconds."""
    rng = np.random.default_rng(base_seed + i)
    na = int(rng.integers(6, 41))
    Lxy = float(rng.uniform(8.0, 12.0))
    en, pos = flake_geometry(na, rng)
    q = rng.uniform(-0.3, 0.3, len(en))
    q -= q.mean()
    atoms = np.zeros((len(en), 4), np.float32)
    atoms[:, :3] = pos
    atoms[:, 3] = q
    n = int(round(Lxy / 0.25))
    L = np.array([n * 0.25, n * 0.25, 12.0], np.float32)
    nxs, nys = int(rng.integers(9, 25)), int(rng.integers(7, 21))
    return Workload(f"sbatch{i}", en, atoms, _lj_params(en), (n, n, 48),
                    np.array([-L[0] / 2, -L[1] / 2, -3.0], np.float32), L, (1, 1),
                    scan_p0=np.array([-L[0] / 2 + 0.3, -L[1] / 2 + 0.2, 10.5], np.float32),
                    scan_da=np.array([(L[0] - 0.6) / nxs, 0., 0.], np.float32),
                    scan_db=np.array([0., (L[1] - 0.4) / nys, 0.], np.float32), nxy=(nxs, nys), nz=14)


def small(seed=7, n=(48, 40, 40), nxy=(12, 10), nz=12) -> Workload:
    """A reduced naphthalene case (coarser grid) that the FP64 oracle finishes in seconds."""
    w = naphthalene(seed)
    L = np.array([12.0, 10.0, 10.0], np.float32)
    w.name = "small"
    w.n = n
    w.L = L
    w.p0 = np.array([-6.0, -5.0, -3.0], np.float32)
    w.scan_p0 = np.array([-3.3, -2.7, 9.5], np.float32)
    w.scan_da = np.array([0.55, 0., 0.], np.float32)
    w.scan_db = np.array([0., 0.55, 0.], np.float32)
    w.nxy = nxy
    w.nz = nz
    return w


def rocksalt_slab(cells=8, a=5.64, layers=4, q=0.8):
    """Rock-salt slab: `layers` (001) layers of (2 cells)^2 ions at a/2 spacing, top layer at
    z = 0, charges +q / -q alternating in x, y and z (Na+ / Cl-).  Returns (enames, atoms)."""
    m = 2 * cells
    en, rows = [], []
    for l in range(layers):
        for i in range(m):
            for j in range(m):
                pos = (i + j + l) % 2 == 0
                en.append("Na" if pos else "Cl")
                rows.append((i * a / 2, j * a / 2, -l * a / 2, q if pos else -q))
    return en, np.array(rows, np.float32)


def substrate_scene(seed=4, n_mol=60, n_poses=64, cells=8, step=0.05, nz_grid=300, nxy=(192, 192), nz=60,
                    scan_size=20.0, tip_top=16.0):
    """The substrate configuration (see the module docstring); returns a dict of host arrays."""
    rng = np.random.default_rng(seed)
    a = 5.64
    sen, sub = rocksalt_slab(cells, a)
    Lxy = cells * a
    men, pos = flake_geometry(n_mol, rng)
    mol = np.zeros((n_mol, 4), np.float32)
    mol[:, 0] = pos[:, 0] + Lxy / 2
    mol[:, 1] = pos[:, 1] + Lxy / 2
    mol[:, 2] = pos[:, 2] + 3.0
    q = rng.uniform(-0.3, 0.3, n_mol)
    mol[:, 3] = q - q.mean()
    poses = np.zeros((n_poses, 4))
    poses[:, :2] = rng.uniform(-2.0, 2.0, (n_poses, 2))
    poses[:, 3] = rng.uniform(0.0, 2 * np.pi, n_poses)
    n_xy = int(round(Lxy / step))
    L = np.array([Lxy, Lxy, nz_grid * step], np.float32)
    p0 = np.array([0.0, 0.0, -2.5], np.float32)
    sstep = scan_size / nxy[0]
    scan_p0 = np.array([Lxy / 2 - scan_size / 2 + sstep / 2, Lxy / 2 - scan_size / 2 + sstep / 2, tip_top],
                       np.float32)
    return {
        "sub_enames": sen, "sub_atoms": sub, "sub_cLJ": mix_lj(sen, ELEMENTS_BENCH, PROBE_R0, PROBE_E0),
        "mol_enames": men, "mol_atoms": mol, "mol_cLJ": mix_lj(men, ELEMENTS_BENCH, PROBE_R0, PROBE_E0),
        "poses": poses, "n": (n_xy, n_xy, nz_grid), "p0": p0, "L": L,
        "lvA": np.array([Lxy, 0, 0], np.float32), "lvB": np.array([0, Lxy, 0], np.float32), "npbc": (1, 1),
        "scan_p0": scan_p0, "scan_da": np.array([sstep, 0, 0], np.float32),
        "scan_db": np.array([0, sstep, 0], np.float32), "nxy": tuple(nxy), "nz": nz, "dtip": -0.1,
        "Qs": np.array([PROBE_Q, 0, 0, 0], np.float32), "QZs": np.zeros(4, np.float32),
        "k_lat_Nm": 0.25, "k_rad_Nm": 20.0, "bond": 4.0,
    }


def small_substrate_scene(seed=9):
    """A reduced substrate configuration that the FP64 oracle finishes in seconds: 2x2-cell slab
    of 2 layers, a 6-atom molecule, 0.25 A grid (45x45x52), 10x10 pixels x 8 z-steps."""
    d = substrate_scene(seed=seed, n_mol=6, n_poses=4, cells=2, step=0.25, nz_grid=52, nxy=(10, 10), nz=8,
                        scan_size=4.0, tip_top=11.0)
    en, sub = rocksalt_slab(2, 5.64, layers=2)
    d["sub_enames"], d["sub_atoms"] = en, sub
    d["sub_cLJ"] = mix_lj(en, ELEMENTS_BENCH, PROBE_R0, PROBE_E0)
    d["poses"][:, :2] *= 0.25
    return d
