"""Slab inputs of the 2D Ewald tests.

The NaCl(001) slabs under tests/golden/substrates are data files of the reference
(data/substrates/NaCl_1x1_L3.xyz, NaCl_1x1_L2_checker.xyz, NaCl_8x8_L3.xyz): an .xyz whose
comment line is `lvs ax ay az bx by bz cx cy cz` and whose 5th column is the ion charge.
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SUBSTRATES = os.path.join(HERE, "golden", "substrates")


def load_substrate(name):
    """(rx, ry, rz, q, a_vec, b_vec) of a fixture (the fields tests/test_surface.py:10-18 of the
    reference reads from AtomicSystem)."""
    with open(os.path.join(SUBSTRATES, name)) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    lv = np.array([float(v) for v in lines[1].split()[1:10]]).reshape(3, 3)
    rows = np.array([[float(v) for v in ln.split()[1:5]] for ln in lines[2:2 + n]])
    return rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3], lv[0, :2].copy(), lv[1, :2].copy()


def ions_of(rx, ry, rz, q):
    return np.stack([rx, ry, rz, q], 1).astype(np.float32)


def evjen_cluster(rx, ry, rz, q, a_vec, b_vec, K=20):
    """Finite checkerboard cluster of the NaCl(001) slab on the half-cell sublattice, sites
    -K..K along x and y in every layer, boundary weights 1/2 per edge (Evjen), the top layer
    positive at the origin and the sign alternating with depth (the cluster the reference's
    tests compare eval_cluster with, tests/test_surface.py:20-72)."""
    hx, hy = float(a_vec[0]) / 2.0, float(b_vec[1]) / 2.0
    q0 = float(abs(q[0]))
    zs = np.unique(np.asarray(rz, dtype=np.float64))          # ascending: last = top layer
    i = np.arange(-K, K + 1)
    I, J = np.meshgrid(i, i, indexing="ij")
    w = np.where(np.abs(I) == K, 0.5, 1.0) * np.where(np.abs(J) == K, 0.5, 1.0)
    chk = np.where((I + J) % 2 == 0, 1.0, -1.0)
    out = []
    for li, z in enumerate(zs):
        sign = 1.0 if (len(zs) - 1 - li) % 2 == 0 else -1.0
        out.append(np.stack([I.ravel() * hx, J.ravel() * hy, np.full(I.size, z),
                             (sign * q0 * chk * w).ravel()], 1))
    return np.concatenate(out).astype(np.float32)


def rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)))


def correlation(a, b):
    return float(np.corrcoef(np.asarray(a, float), np.asarray(b, float))[0, 1])
