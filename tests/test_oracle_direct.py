"""Oracle of the gridless direct Morse path (relaxStrokesTiltedMorseDirect and its adjoint).

Pins (none retypes the oracle's formulas):
* every converged probe position is an equilibrium: the total force re-evaluated with
  independent NumPy (pair law from ``pair_fe_d`` + tip spring written out here) is below
  1e-4 eV/A (tests/SPM/test_afm_morse.py:1048-1089);
* FP32 and FP64 relaxations agree at tight convergence;
* the analytic adjoint equals, channel by channel, the reference's finite-difference
  component construction (J, dO/dq, dG/dtheta, dO/dtheta by central differences,
  tests/SPM/test_afm_morse.py:1240-1309);
* the adjoint's directional derivative equals re-relaxed finite differences of the FP64
  forward on a plateau of three step sizes (tests/SPM/test_afm_morse.py:1523-1582);
* each gate (non-converged, residual, non-finite, unstable) fires on a constructed state.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.direct_cases import (RELAX, TIP, converged_case, fd_component_vjp, relax64, tipC,
                                 total_force, toy_scan, toy_system, vjp64)

R = pytest.importorskip("oracle.df_ref")   # the restatement stays in the build container

def test_direct_relax_residual_at_every_state():
    atoms, cMs, Qs, QZs = toy_system(2, 7)
    nz, dtip = 5, -0.2
    pts = toy_scan(atoms)
    FEs, PPs = O.relax_direct(atoms, cMs, Qs, QZs, pts, nz, TIP["tipA"], TIP["tipB"], tipC(dtip), TIP["stiffness"],
                              TIP["dpos0"], RELAX, TIP["surfFF"], 100, 1e-8, "f64")
    assert (PPs[..., 3] > 0).all()
    worst = 0.0
    for ip in range(pts.shape[0]):
        for iz in range(nz):
            tip = pts[ip, :3] + np.array([0, 0, dtip]) * iz
            worst = max(worst, np.linalg.norm(total_force(atoms, cMs, Qs, QZs, PPs[ip, iz, :3].astype(np.float64),
                                                          tip, tipC(dtip))))
    assert worst < 1e-4, worst


def test_direct_relax_fp32_matches_fp64_at_tight_convergence():
    atoms, cMs, Qs, QZs = toy_system(3, 19)
    nz, dtip = 6, -0.2
    pts = toy_scan(atoms, (4, 3))
    args = (atoms, cMs, Qs, QZs, pts, nz, TIP["tipA"], TIP["tipB"], tipC(dtip), TIP["stiffness"], TIP["dpos0"],
            RELAX, TIP["surfFF"], 2000)
    F32, P32 = O.relax_direct(*args, 1e-12, "f32")
    F64, P64 = O.relax_direct(*args, 1e-16, "f64")
    assert (P32[..., 3] > 0).all() and (P64[..., 3] > 0).all()
    assert np.abs(P32[..., :3] - P64[..., :3]).max() < 1e-4
    assert np.abs(F32[..., 2] - F64[..., 2]).max() < 1e-4


def test_direct_vjp_matches_fd_component_construction_per_channel():
    nz, dtip = 5, -0.2
    atoms, cMs, Qs, QZs, pts, FEs, PPs = converged_case(nz=nz, dtip=dtip)
    nx, ny = 3, 3
    target = np.random.default_rng(33).standard_normal((nx, ny, nz))
    _, _, dL = R.df_loss_seed(FEs.reshape(nx, ny, nz, 4), target, dtip)
    U = dL.reshape(-1, nz, 4)
    g, lam, diag = vjp64(atoms, cMs, Qs, QZs, pts, nz, dtip, PPs, U)
    assert (diag[:, 3] == 0).all(), diag[:, 3]
    gf = fd_component_vjp(atoms, cMs, Qs, QZs, pts, nz, dtip, PPs, U)
    for p in range(6):
        scale = np.abs(gf[:, p]).max() + 1e-30
        assert np.abs(g[:, p] - gf[:, p]).max() / scale < 1e-5, (p, g[:, p], gf[:, p])


def test_direct_vjp_matches_re_relaxed_fd_plateau():
    """dot(grad, v) vs [L(theta+hv) - L(theta-hv)]/2h with re-relaxation, three decreasing h."""
    nz, dtip, nxy = 5, -0.2, (3, 3)
    atoms, cMs, Qs, QZs, pts, FEs, PPs = converged_case(nz=nz, dtip=dtip, nxy=nxy)
    Fz = FEs[..., 2].reshape(*nxy, nz)
    target = np.zeros((*nxy, nz))
    target[:, :, 1:-1] = -(Fz[:, :, 2:] - Fz[:, :, :-2]) / (2 * abs(dtip))
    target += np.random.default_rng(33).standard_normal(target.shape) * 0.01
    _, _, dL = R.df_loss_seed(FEs.reshape(*nxy, nz, 4), target, dtip)
    g, _, diag = vjp64(atoms, cMs, Qs, QZs, pts, nz, dtip, PPs, dL.reshape(-1, nz, 4))
    assert (diag[:, 3] == 0).all()
    v = np.random.default_rng(55).standard_normal((atoms.shape[0], 6)) * 0.01
    lin = float((g * v).sum())

    def loss(h):
        at, cm = atoms.copy(), cMs.copy()
        at[:, :3] += h * v[:, :3]
        cm[:, 0] += h * v[:, 3]
        cm[:, 1] += h * v[:, 4]
        at[:, 3] += h * v[:, 5]
        F, _ = relax64(at, cm, Qs, QZs, pts, nz, dtip)
        return R.df_loss_seed(F.reshape(*nxy, nz, 4), target, dtip)[0]

    rels = []
    for h in (1e-3, 1e-4, 1e-5):
        fd = (loss(h) - loss(-h)) / (2 * h)
        rels.append(abs(fd - lin) / abs(fd))
    assert max(rels) < 1e-4, rels


def test_direct_vjp_gates_fire_on_constructed_states():
    nz, dtip = 5, -0.2
    atoms, cMs, Qs, QZs, pts, FEs, PPs = converged_case(nz=nz, dtip=dtip)
    U = np.zeros_like(PPs)
    U[..., 2] = 1.0
    bad = PPs.copy()
    bad[0, 0, 3] = -7.0                      # forward not converged -> 1
    bad[0, 1, 0] += 0.1                      # off equilibrium -> residual gate 1
    Ub = U.copy()
    Ub[0, 2, 1] = np.nan                     # non-finite upstream -> 2
    g, lam, diag = vjp64(atoms, cMs, Qs, QZs, pts, nz, dtip, bad, Ub)
    st = diag[:, 3].reshape(-1, nz)
    assert st[0, 0] == 1 and st[0, 1] == 1 and st[0, 2] == 2
    assert diag.reshape(-1, nz, 4)[0, 1, 0] >= 5e-4
    assert (st.ravel()[3:] == 0).all()
    assert (lam.reshape(-1, nz, 4)[0, :3] == 0).all()
    # dropped states contribute nothing: the gradient equals the one over the kept states only
    keep = np.ones(PPs.shape[:2], bool)
    keep[0, :3] = False
    U2 = U * keep[..., None]
    g2, _, _ = vjp64(atoms, cMs, Qs, QZs, pts, nz, dtip, PPs, U2)
    np.testing.assert_allclose(g, g2, rtol=1e-12, atol=1e-15)


def test_direct_vjp_unstable_state_is_gated():
    """Repulsive (positive) springs at their rest length, no atom force: G = 0 but -J has
    negative eigenvalues -> status 3."""
    atoms = np.array([[0., 0., -40., 0.]])
    cMs = np.array([[3.0, 1e-3, -1.8, 0.]])
    Qs, QZs = np.zeros(4), np.zeros(4)
    pts = np.array([[0., 0., 10., 0.]])
    PPs = np.array([[[0., 0., 6., 1.]]])
    U = np.zeros((1, 1, 4))
    U[..., 2] = 1.0
    stiff = TIP["stiffness"].copy()
    old = TIP["stiffness"]
    TIP["stiffness"] = -stiff
    try:
        _, _, diag = vjp64(atoms, cMs, Qs, QZs, pts, 1, -0.1, PPs, U)
    finally:
        TIP["stiffness"] = old
    assert diag[0, 3] == 3 and diag[0, 1] < 0, diag


@pytest.mark.parametrize("seed", [3, 4])
def test_direct_relax_uses_direct_pair_sum_as_sample(seed):
    """The FE output at each state is the rotated direct pair sum at the returned PP."""
    atoms, cMs, Qs, QZs = toy_system(4, seed)
    pts = toy_scan(atoms, (2, 3))
    FEs, PPs = relax64(atoms, cMs, Qs, QZs, pts, 4, -0.2, f2conv=1e-12)
    fe = O.pair_fe_d(1, atoms, cMs, PPs[..., :3].reshape(-1, 3), Qs, QZs).reshape(FEs.shape)
    np.testing.assert_allclose(FEs, fe, rtol=1e-12, atol=1e-14)
