"""Shared cases and checkers of the direct Morse path tests (CPU oracle and GPU)."""
import numpy as np

from oracle import oracle as O

TIP = dict(tipA=np.array([1., 0., 0., 0.]), tipB=np.array([0., 1., 0., 0.]),
           stiffness=np.array([-0.03, -0.03, -0.03, -1.0]), dpos0=np.array([0., 0., -4.0, 4.0]),
           surfFF=np.zeros(4))
RELAX = np.array([0.5, 0.1, 0.02, 0.5])


def toy_system(n_atoms=2, seed=7):
    """Small random molecule in a 2 A box, Morse (R0,E0,K) per atom, two tip charge sites
    (the shape of the reference's toy system, tests/SPM/test_afm_morse.py:626-642)."""
    rng = np.random.default_rng(seed)
    atoms = np.zeros((n_atoms, 4))
    atoms[:, :3] = rng.uniform(-1.0, 1.0, (n_atoms, 3))
    atoms[:, 3] = rng.uniform(-0.3, 0.3, n_atoms)
    cMs = np.zeros((n_atoms, 4))
    cMs[:, 0] = rng.uniform(2.5, 3.5, n_atoms)
    cMs[:, 1] = rng.uniform(1e-3, 5e-3, n_atoms)
    cMs[:, 2] = -1.8
    return atoms, cMs, np.array([0., -0.1, 0.1, 0.]), np.array([0., 1.8, 3.6, 0.])


def toy_scan(atoms, nxy=(3, 3), clearance=5.0, dpos0_z=4.0):
    """Scan raster over 90% of the molecule's x,y extent, starting clearance + |dpos0_z| above it."""
    mn, mx = atoms[:, :3].min(0), atoms[:, :3].max(0)
    p0 = np.array([mn[0] + 0.05 * (mx[0] - mn[0]), mn[1] + 0.05 * (mx[1] - mn[1]), mx[2] + clearance + dpos0_z],
                  np.float32)
    da = np.array([(mx[0] - mn[0]) * 0.9 / max(nxy[0] - 1, 1), 0, 0], np.float32)
    db = np.array([0, (mx[1] - mn[1]) * 0.9 / max(nxy[1] - 1, 1), 0], np.float32)
    pts = np.zeros((nxy[0] * nxy[1], 4))
    for ix in range(nxy[0]):
        for iy in range(nxy[1]):
            pts[ix * nxy[1] + iy, :3] = p0 + da * ix + db * iy
    return pts


def tipC(dtip):
    return np.array([0., 0., 1., dtip])


def total_force(atoms, cMs, Qs, QZs, q, tip, tc):
    """G = F_sample + R^T tipForce(R (q - tip)) + tipC surfFF, in double, written out independently."""
    A, B, C = TIP["tipA"][:3], TIP["tipB"][:3], tc[:3]
    fe = O.pair_fe_d(1, atoms, cMs, np.asarray(q)[None], Qs, QZs)[0]
    dw = np.asarray(q) - tip
    d = np.array([dw @ A, dw @ B, dw @ C])
    k, d0 = TIP["stiffness"], TIP["dpos0"]
    r = max(np.linalg.norm(d), 1e-10)
    g = (d - d0[:3]) * k[:3] + d * (k[3] * (r - d0[3]) / r)
    return fe[:3] + A * g[0] + B * g[1] + C * g[2] + C * TIP["surfFF"][0]


def relax64(atoms, cMs, Qs, QZs, pts, nz, dtip, f2conv=1e-20, max_iter=5000):
    return O.relax_direct_d(atoms, cMs, Qs, QZs, pts, nz, TIP["tipA"], TIP["tipB"], tipC(dtip), TIP["stiffness"],
                            TIP["dpos0"], RELAX, TIP["surfFF"], max_iter, f2conv)


def vjp64(atoms, cMs, Qs, QZs, pts, nz, dtip, PPs, U):
    return O.direct_vjp(atoms, cMs, Qs, QZs, pts, nz, TIP["tipA"], TIP["tipB"], tipC(dtip), TIP["stiffness"],
                        TIP["dpos0"], TIP["surfFF"], PPs, U)


def fd_component_vjp(atoms, cMs, Qs, QZs, pts, nz, dtip, PPs, U, hq=1e-5, hp=1e-6):
    """The reference's CPU64 VJP construction: every derivative by central differences of the
    FP64 pair law, the 3x3 adjoint solved with numpy (tests/SPM/test_afm_morse.py:1240-1309)."""
    na = atoms.shape[0]
    A, B, C = TIP["tipA"][:3], TIP["tipB"][:3], tipC(dtip)[:3]
    Rm = np.stack([A, B, C])
    grad = np.zeros((na, 6))

    def rot_fe(at, cm, q):
        fe = O.pair_fe_d(1, at, cm, np.asarray(q)[None], Qs, QZs)[0]
        return np.concatenate([Rm @ fe[:3], fe[3:]])

    def perturb(at, cm, i, p, h):
        at, cm = at.copy(), cm.copy()
        if p < 3:
            at[i, p] += h
        elif p == 3:
            cm[i, 0] += h
        elif p == 4:
            cm[i, 1] += h
        else:
            at[i, 3] += h
        return at, cm

    for ip in range(pts.shape[0]):
        for iz in range(nz):
            q = PPs[ip, iz, :3]
            u = U[ip, iz]
            tip = pts[ip, :3] + C * dtip * iz
            J = np.zeros((3, 3))
            dOdq = np.zeros((4, 3))
            for j in range(3):
                e = np.zeros(3)
                e[j] = hq
                J[:, j] = (total_force(atoms, cMs, Qs, QZs, q + e, tip, tipC(dtip)) -
                           total_force(atoms, cMs, Qs, QZs, q - e, tip, tipC(dtip))) / (2 * hq)
                dOdq[:, j] = (rot_fe(atoms, cMs, q + e) - rot_fe(atoms, cMs, q - e)) / (2 * hq)
            lam = np.linalg.solve(J.T, dOdq.T @ u)
            for i in range(na):
                for p in range(6):
                    ap, cp = perturb(atoms, cMs, i, p, hp)
                    am, cm = perturb(atoms, cMs, i, p, -hp)
                    fp = O.pair_fe_d(1, ap, cp, q[None], Qs, QZs)[0]
                    fm = O.pair_fe_d(1, am, cm, q[None], Qs, QZs)[0]
                    dG = (fp[:3] - fm[:3]) / (2 * hp)
                    dO = (rot_fe(ap, cp, q) - rot_fe(am, cm, q)) / (2 * hp)
                    grad[i, p] += u @ dO - lam @ dG
    return grad


def converged_case(n_atoms=2, seed=31, nxy=(3, 3), nz=5, dtip=-0.2):
    atoms, cMs, Qs, QZs = toy_system(n_atoms, seed)
    pts = toy_scan(atoms, nxy)
    FEs, PPs = relax64(atoms, cMs, Qs, QZs, pts, nz, dtip)
    assert (PPs[..., 3] > 0).all()
    return atoms, cMs, Qs, QZs, pts, FEs, PPs
