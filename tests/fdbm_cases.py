"""Synthetic FDBM inputs: Gaussian sample densities around a few atoms, a Gaussian tip density and
a neutral (difference-of-Gaussians) tip charge density, on FFT-friendly grids."""
import numpy as np


def gaussian(shape, origin, step, centers, sigmas, weights):
    axes = [origin[a] + step * np.arange(shape[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    out = np.zeros(shape)
    for c, s, w in zip(centers, sigmas, weights):
        r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2
        out += w * np.exp(-r2 / (2 * s * s)) / (2 * np.pi * s * s) ** 1.5
    return out


def case(shape=(40, 36, 48), step=0.2, seed=3, tip_shape=(14, 14, 16)):
    rng = np.random.default_rng(seed)
    origin = -0.5 * step * np.array(shape, float)
    na = 6
    pos = rng.uniform(-1.5, 1.5, size=(na, 3)) * np.array([1.0, 1.0, 0.2])
    Z = rng.choice([1, 6, 7, 8], size=na)
    rho_scf = gaussian(shape, origin, step, pos, [0.55] * na, Z.astype(float))
    q = rng.uniform(-0.3, 0.3, size=na)
    q -= q.mean()
    rho_diff = (gaussian(shape, origin, step, pos, [0.5] * na, q)
                - gaussian(shape, origin, step, pos + [0, 0, 0.3], [0.7] * na, q))
    t0 = -0.5 * step * (np.array(tip_shape, float) - 1)
    tip_tot = gaussian(tip_shape, t0, step, [[0, 0, 0]], [0.6], [6.0])
    tip_del = (gaussian(tip_shape, t0, step, [[0, 0, 0.2]], [0.5], [0.4])
               - gaussian(tip_shape, t0, step, [[0, 0, -0.2]], [0.5], [0.4]))
    return dict(shape=shape, step=step, origin=origin, pos=pos, Z=Z, rho_scf=rho_scf.astype(np.float32),
                rho_diff=rho_diff.astype(np.float32), tip_tot=tip_tot.astype(np.float32),
                tip_del=tip_del.astype(np.float32))
