"""Pins of the oracle's pair law and grid projection (CPU only).

Restates the reference's own checks of the pair physics
(tests/SPM/test_afm_morse.py:694-857: F=-dU/dq by central FD, translation /
permutation / charge-linearity invariants) for the LJ and Morse laws of
kernels/AFM.cl:730-757, adds single-atom closed forms and checks periodic images
and the |E|>100 rescale against independent constructions.
"""
import numpy as np
import pytest

from oracle import oracle as O

COUL = 14.3996448915


def toy(n=3, seed=7, model=1):
    """Deterministic toy system in the style of _make_toy_system (tests/SPM/test_afm_morse.py:626-642)."""
    rng = np.random.default_rng(seed)
    atoms = np.zeros((n, 4))
    atoms[:, :3] = rng.uniform(-1.0, 1.0, (n, 3))
    atoms[:, 3] = rng.uniform(-0.3, 0.3, n)
    if model == 1:
        c = np.zeros((n, 4))
        c[:, 0] = rng.uniform(2.5, 3.5, n)
        c[:, 1] = rng.uniform(1e-3, 5e-3, n)
        c[:, 2] = -1.8
    else:
        R0 = rng.uniform(3.0, 3.8, n)
        E0 = rng.uniform(1e-3, 5e-3, n)
        c = np.stack([2 * E0 * R0 ** 6, E0 * R0 ** 12], axis=1)
    return atoms, c, np.array([0., -0.1, 0.1, 0.]), np.array([0., 1.8, 3.6, 0.])


def fd_force(model, atoms, c, q, Qs, QZs, h=1e-6):
    g = np.zeros_like(q)
    for ax in range(3):
        qp, qm = q.copy(), q.copy()
        qp[:, ax] += h
        qm[:, ax] -= h
        g[:, ax] = (O.pair_fe_d(model, atoms, c, qp, Qs, QZs)[:, 3] - O.pair_fe_d(model, atoms, c, qm, Qs, QZs)[:, 3]) / (2 * h)
    return -g


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("charges", [False, True])
def test_force_is_minus_energy_gradient(model, charges):
    atoms, c, Qs, QZs = toy(4, seed=17 + model, model=model)
    if not charges:
        Qs = np.zeros(4)
    q = np.random.default_rng(23).uniform(-2, 2, (10, 3)) + np.array([0, 0, 3.0])
    fe = O.pair_fe_d(model, atoms, c, q, Qs, QZs)
    F = fd_force(model, atoms, c, q, Qs, QZs)
    scale = max(1.0, np.abs(fe[:, :3]).max())
    assert np.abs(fe[:, :3] - F).max() < 1e-6 * scale


def test_coulomb_only_matches_closed_form():
    atoms = np.array([[0.3, -0.2, 0.1, 0.7]])
    c = np.zeros((1, 2))   # c6=c12=0: no LJ
    Qs = np.array([-0.05, 0.2, 0., 0.])
    QZs = np.array([0., 1.5, 0., 0.])
    q = np.array([[1.0, 2.0, 3.0], [-1.5, 0.5, 2.0]])
    fe = O.pair_fe_d(0, atoms, c, q, Qs, QZs)
    ref = np.zeros((2, 4))
    for k in range(2):
        d = q + np.array([0, 0, QZs[k]]) - atoms[0, :3]
        r2 = (d * d).sum(1) + 1e-4
        E = COUL * Qs[k] * atoms[0, 3] / np.sqrt(r2)
        ref[:, 3] += E
        ref[:, :3] += d * (E / r2)[:, None]
    np.testing.assert_allclose(fe, ref, rtol=1e-12, atol=1e-15)


def test_lj_minimum_at_R0():
    R0, E0 = 3.5, 0.004
    atoms = np.zeros((1, 4))
    c = np.array([[2 * E0 * R0 ** 6, E0 * R0 ** 12]])
    r = np.sqrt(R0 ** 2 - 1e-4)      # r^2 + R2SAFE = R0^2
    fe = O.pair_fe_d(0, atoms, c, np.array([[0, 0, r]]), np.zeros(4), np.zeros(4))
    assert abs(fe[0, 3] + E0) < 1e-15
    assert np.abs(fe[0, :3]).max() < 1e-15


def test_morse_minimum_at_R0():
    R0, E0 = 3.2, 0.003
    atoms = np.zeros((1, 4))
    c = np.array([[R0, E0, -1.8, 0.]])
    r = np.sqrt(R0 ** 2 - 1e-4)
    fe = O.pair_fe_d(1, atoms, c, np.array([[r, 0, 0]]), np.zeros(4), np.zeros(4))
    assert abs(fe[0, 3] + E0) < 1e-15
    assert np.abs(fe[0, :3]).max() < 1e-14
    # repulsive inside, attractive outside
    fin = O.pair_fe_d(1, atoms, c, np.array([[R0 - 0.5, 0, 0]]), np.zeros(4), np.zeros(4))
    fout = O.pair_fe_d(1, atoms, c, np.array([[R0 + 0.5, 0, 0]]), np.zeros(4), np.zeros(4))
    assert fin[0, 0] > 0 and fout[0, 0] < 0


@pytest.mark.parametrize("model", [0, 1])
def test_translation_invariant(model):
    atoms, c, Qs, QZs = toy(3, 41, model)
    q = np.random.default_rng(43).uniform(-1, 1, (6, 3))
    shift = np.array([3.3, -2.1, 5.7])
    a2 = atoms.copy()
    a2[:, :3] += shift
    np.testing.assert_allclose(O.pair_fe_d(model, a2, c, q + shift, Qs, QZs), O.pair_fe_d(model, atoms, c, q, Qs, QZs),
                               rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("model", [0, 1])
def test_permutation_invariant(model):
    atoms, c, Qs, QZs = toy(4, 67, model)
    q = np.random.default_rng(71).uniform(-1, 1, (5, 3)) + [0, 0, 3]
    perm = [2, 0, 3, 1]
    np.testing.assert_allclose(O.pair_fe_d(model, atoms[perm], c[perm], q, Qs, QZs),
                               O.pair_fe_d(model, atoms, c, q, Qs, QZs), rtol=1e-12, atol=1e-15)


def test_charge_linearity():
    atoms, c, Qs, QZs = toy(3, 83, 1)
    q = np.random.default_rng(89).uniform(-1, 1, (6, 3)) + [0, 0, 3]
    fe0 = O.pair_fe_d(1, atoms, c, q, Qs, QZs)
    fm = O.pair_fe_d(1, atoms, c, q, np.zeros(4), QZs)
    a2 = atoms.copy()
    a2[:, 3] *= 2.5
    fe2 = O.pair_fe_d(1, a2, c, q, Qs, QZs)
    np.testing.assert_allclose(fe2 - fm, 2.5 * (fe0 - fm), rtol=1e-10, atol=1e-14)


def test_periodic_images_equal_replicated_atoms():
    atoms, c, Qs, QZs = toy(3, 5, 0)
    lvA, lvB = np.array([7.0, 0.5, 0.0]), np.array([-0.3, 6.0, 0.0])
    q = np.random.default_rng(9).uniform(-3, 3, (7, 3)) + [0, 0, 2.5]
    fe_img = O.pair_fe_d(0, atoms, c, q, Qs, QZs, npbc=(1, 2), lvA=lvA, lvB=lvB)
    rep_a, rep_c = [], []
    for j in range(3):
        for iy in range(-2, 3):
            for ix in range(-1, 2):
                a = atoms[j].copy()
                a[:3] += ix * lvA + iy * lvB
                rep_a.append(a)
                rep_c.append(c[j])
    fe_rep = O.pair_fe_d(0, np.array(rep_a), np.array(rep_c), q, Qs, QZs)
    np.testing.assert_allclose(fe_img, fe_rep, rtol=1e-12, atol=1e-14)


def test_grid_projection_equals_pair_sum_and_cap():
    """Grid voxels = pair_fe at p0+dA*ia+dB*ib+dC*ic, then |E|>100 rescale (kernels/AFM.cl:984-985)."""
    atoms, c, Qs, QZs = toy(3, 3, 0)
    n = (6, 5, 7)
    p0 = np.array([-1.5, -1.2, -0.4])
    dA, dB, dC = np.array([0.5, 0, 0]), np.array([0.1, 0.55, 0]), np.array([0, 0, 0.45])
    G = O.grid_project(0, atoms, c, n, p0, dA, dB, dC, Qs, QZs, cap=True)
    assert G.shape == (7, 5, 6, 4)
    ia, ib, ic = np.meshgrid(np.arange(6), np.arange(5), np.arange(7), indexing="ij")
    f = lambda v: v.astype(np.float32).astype(np.float64)   # the grid geometry is float32
    pos = f(p0) + ia[..., None] * f(dA) + ib[..., None] * f(dB) + ic[..., None] * f(dC)
    fe = O.pair_fe(0, atoms, c, pos.reshape(-1, 3), Qs, QZs)
    s = np.minimum(1.0, 100.0 / np.abs(fe[:, 3]))
    fe = fe * s[:, None]
    ref = fe.reshape(6, 5, 7, 4).transpose(2, 1, 0, 3)
    assert (np.abs(ref[..., 3]) > 99.9).any(), "test grid should reach the capped region"
    np.testing.assert_allclose(G, ref, rtol=2e-7, atol=1e-6)
    # float32 restatement agrees with the float64 one to float32 accuracy, measured against the
    # magnitude of each voxel's (Fx,Fy,Fz,E) vector (terms of opposite sign cancel in the sum)
    G32 = O.grid_project(0, atoms, c, n, p0, dA, dB, dC, Qs, QZs, cap=True, precision="f32")
    mag = np.linalg.norm(G.astype(np.float64), axis=-1, keepdims=True)
    assert (np.abs(G32 - G) <= 1e-5 * mag + 1e-6).all()
