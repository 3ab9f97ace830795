"""Device 2D Ewald (spammm_amd.ewald over ppm_ewald_*) against the FP64 oracle (oracle/ewald_ref.py).

* eval_full / eval_vacuum on a skewed three-layer slab, points between and above the layers;
* the G compaction: an 8x8 supercell (the reference's NaCl_8x8_L3 fixture) at n_harm 32 keeps
  only the reciprocal vectors of its 1x1 cell and equals the 1x1 cell at n_harm 4;
* the image sums (circle and square) and the finite cluster;
* the probe grid (E and F = -grad E for two tip charges at two z offsets) on an axis-aligned and a
  skewed grid, with and without accumulation onto a base;
* the reference's own SurfaceEwaldCL tests (tests/test_surface.py:160-262) on its fixtures.
Tolerances are relative to the largest magnitude of the compared quantity: 2e-6 for the FP64
point evaluators (their outputs are FP32), 2e-5 for the FP32 grid tables and phases.
"""
import numpy as np
import pytest

from oracle import ewald_ref as E
from spammm_amd import workloads as WL
from tests import ewald_cases as EC

pytestmark = pytest.mark.gpu


def _skew_slab(seed=7):
    rng = np.random.default_rng(seed)
    a, b = np.array([6.1, 0.0]), np.array([1.7, 5.3])
    rows = []
    for z, n in ((0.0, 5), (-1.9, 4), (-3.7, 3)):
        f = rng.random((n, 2))
        xy = f[:, :1] * a + f[:, 1:] * b
        rows += [[x, y, z, qq] for (x, y), qq in zip(xy, rng.uniform(-0.8, 0.8, n))]
    ions = np.array(rows, np.float32)
    ions[:, 3] -= ions[:, 3].mean()
    return ions, a, b


def _close(got, want, rel, what):
    err = np.abs(np.asarray(got, np.float64) - want).max()
    scale = np.abs(want).max()
    print(f"{what}: max|err| {err:.3e}, max|ref| {scale:.3e}", flush=True)
    assert err <= rel * scale, f"{what}: {err:.3e} > {rel} * {scale:.3e}"


def test_points_full_and_vacuum_match_oracle():
    from spammm_amd.ewald import SurfaceEwald
    ions, a, b = _skew_slab()
    ew = SurfaceEwald().prepare_system(ions, a, b, n_harm=6)
    o = E.Slab(ions.astype(np.float64), a, b, 6)
    rng = np.random.default_rng(1)
    P = np.stack([rng.uniform(-9, 15, 64), rng.uniform(-7, 12, 64), rng.uniform(-4.4, 4.0, 64)], 1).astype(np.float32)
    got = ew.eval_full(P[:, 0], P[:, 1], P[:, 2])
    _close(got, o.phi_full(P.astype(np.float64)), 2e-6, "eval_full")
    Pv = P.copy()
    Pv[:, 2] = np.abs(Pv[:, 2]) + 0.5
    got = ew.eval_vacuum(Pv[:, 0], Pv[:, 1], Pv[0, 2])
    Pv[:, 2] = Pv[0, 2]
    _close(got, o.phi_vacuum(Pv.astype(np.float64)), 2e-6, "eval_vacuum")


def test_supercell_compaction_equals_unit_cell():
    from spammm_amd.ewald import SurfaceEwald
    r1 = EC.load_substrate("NaCl_1x1_L3.xyz")
    r8 = EC.load_substrate("NaCl_8x8_L3.xyz")
    ew8 = SurfaceEwald().prepare_system(EC.ions_of(*r8[:4]), r8[4], r8[5], n_harm=32)
    ions1 = EC.ions_of(*r1[:4]).astype(np.float64)
    ions1[:, 3] *= np.abs(r8[3]).max() / np.abs(r1[3]).max()       # the fixtures carry +-0.7 e and +-1 e
    o1 = E.Slab(ions1, r1[4], r1[5], 4)
    # kept G of the supercell = the 1x1 cell's G (|h|,|k| <= 4) with a nonzero layer structure factor
    x, y, z, q = o1.ions.T
    Wl = [(o1.w[:, z == zl]).sum(1) for zl in np.unique(z)]
    nonzero = np.max(np.abs(np.stack(Wl)), 0) > 1e-9 * np.abs(np.stack(Wl)).max()
    assert 2 * ew8.n_kept() == int(nonzero.sum()), (ew8.n_kept(), int(nonzero.sum()))
    rng = np.random.default_rng(2)
    P = np.stack([rng.uniform(0, 32, 48), rng.uniform(0, 32, 48), rng.uniform(-8.0, 4.0, 48)], 1).astype(np.float32)
    _close(ew8.eval_full(P[:, 0], P[:, 1], P[:, 2]), o1.phi_full(P.astype(np.float64)), 2e-6, "8x8 vs 1x1")


def test_image_sums_and_cluster_match_oracle():
    from spammm_amd.ewald import SurfaceEwald
    rx, ry, rz, q, a, b = EC.load_substrate("NaCl_1x1_L3.xyz")
    ions = EC.ions_of(rx, ry, rz, q)
    ew = SurfaceEwald().prepare_system(ions, a, b, n_harm=4)
    o = E.Slab(ions.astype(np.float64), a, b, 4)
    z = np.linspace(-2.0, 5.0, 24).astype(np.float32)
    P = np.stack([np.full_like(z, 0.3), np.full_like(z, 1.1), z], 1)
    for square in (False, True):
        got = ew.eval_brute(P[:, 0], P[:, 1], P[:, 2], N_rep=10, square=square)
        _close(got, o.phi_brute(P.astype(np.float64), 10, square=square), 2e-6, f"brute square={square}")
    cl = EC.evjen_cluster(rx, ry, rz, q, a, b, K=12)
    got = ew.eval_cluster(P[:, 0], P[:, 1], P[:, 2], cl)
    _close(got, E.phi_cluster(P.astype(np.float64), cl.astype(np.float64)), 2e-6, "cluster")


def _tip_case():
    en, ions = WL.rocksalt_slab(cells=1, layers=2)
    a, b = np.array([5.64, 0.0]), np.array([0.0, 5.64])
    return ions, a, b, np.array([-0.1, 0.05, 0.0, 0.0]), np.array([0.0, 0.3, 0.0, 0.0])


def test_tip_grid_matches_oracle_axis_aligned():
    import torch
    from spammm_amd import device as D
    from spammm_amd.ewald import SurfaceEwald
    ions, a, b, Qs, QZs = _tip_case()
    ew = SurfaceEwald().prepare_system(ions, a, b, n_harm=8)
    n = (24, 20, 22)
    spec = D.axis_aligned_spec((0.1, -0.2, -3.3), (5.64, 5.0, 22 * 0.41), n)
    g = ew.tip_grid(spec, Qs, QZs).cpu().numpy().astype(np.float64)
    o = E.Slab(ions.astype(np.float64), a, b, 8)
    ref = E.tip_grid(o, spec.p0.astype(np.float64), np.array([spec.dA[0], spec.dB[1], spec.dC[2]], np.float64),
                     n, Qs, QZs)
    for c, name in enumerate(("Fx", "Fy", "Fz", "E")):
        _close(g[..., c], ref[..., c], 2e-5, f"tip grid {name}")
    base = torch.randn((n[2], n[1], n[0], 4), dtype=torch.float32, device="cuda")
    acc = base.clone()
    ew.tip_grid(spec, Qs, QZs, out=acc, accumulate=True)
    assert torch.allclose(acc - base, torch.from_numpy(g.astype(np.float32)).cuda(), atol=1e-5, rtol=1e-5)


def test_tip_grid_skewed_grid_matches_point_forms():
    from spammm_amd import device as D
    from spammm_amd.ewald import SurfaceEwald
    ions, a, b = _skew_slab(3)
    Qs, QZs = np.array([0.07, 0.0, -0.02, 0.0]), np.array([0.0, 0.0, -0.4, 0.0])
    ew = SurfaceEwald().prepare_system(ions, a, b, n_harm=7)
    n = (16, 14, 12)
    p0 = np.array([-1.0, 0.5, 0.6], np.float32)
    dA = np.array([0.38, 0.0, 0.0], np.float32)
    dB = np.array([0.12, 0.37, 0.0], np.float32)
    dC = np.array([0.0, 0.0, 0.29], np.float32)
    L = np.array([dA * n[0], dB * n[1], dC * n[2]], np.float64)
    spec = D.GridSpec(n, p0, dA, dB, dC, *D.dinv_lvec(np.linalg.inv(L).T, n))
    g = ew.tip_grid(spec, Qs, QZs).cpu().numpy().astype(np.float64)
    ia, ib, ic = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    P = (p0[None, :] + ia.ravel()[:, None] * dA + ib.ravel()[:, None] * dB + ic.ravel()[:, None] * dC).astype(np.float64)
    o = E.Slab(ions.astype(np.float64), a, b, 7)
    Eref = np.zeros(len(P))
    Fref = np.zeros((len(P), 3))
    for qq, dz in zip(Qs, QZs):
        if qq != 0:
            Eref += qq * o.phi_full(P + [0, 0, dz])
            Fref += qq * o.field_full(P + [0, 0, dz])
    got = g.transpose(2, 1, 0, 3)[ia.ravel(), ib.ravel(), ic.ravel()]     # (ic,ib,ia) storage
    _close(got[:, 3], Eref, 2e-5, "skew E")
    for c in range(3):
        _close(got[:, c], Fref[:, c], 2e-5, f"skew F{c}")


def test_invalid_geometry_raises():
    from spammm_amd import device as D
    from spammm_amd.ewald import SurfaceEwald
    ions, a, b, Qs, QZs = _tip_case()
    ew = SurfaceEwald().prepare_system(ions, a, b, n_harm=3)
    spec = D.axis_aligned_spec((0, 0, 0), (5, 5, 5), (8, 8, 8))
    spec.dC = np.array([0.1, 0.0, 0.625], np.float32)
    with pytest.raises(RuntimeError, match="dC along z"):
        ew.tip_grid(spec, Qs, QZs)
    with pytest.raises(ValueError, match="n_harm"):
        SurfaceEwald().prepare_system(ions, a, b, n_harm=0)


# ---- the reference's SurfaceEwaldCL tests (tests/test_surface.py:160-262) ----

def _ref_pair(name, n_harm):
    from spammm_amd.ewald import Ewald2D, SurfaceEwald
    rx, ry, rz, q, a, b = EC.load_substrate(name)
    cl = SurfaceEwald().prepare_system(EC.ions_of(rx, ry, rz, q), a, b, n_harm=n_harm)
    return cl, Ewald2D(a, b, rx, ry, rz, q, n_harm=n_harm), E.Slab(EC.ions_of(rx, ry, rz, q).astype(float), a, b, n_harm)


def test_reference_vacuum_and_full_parity():
    cl, py, o = _ref_pair("NaCl_1x1_L3.xyz", 3)
    xv = np.linspace(0, 4, 20)
    X, Y = np.meshgrid(xv, xv)
    phi = cl.eval_vacuum(X.astype(np.float32), Y.astype(np.float32), 2.0)
    P = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 2.0)], 1)
    _close(phi.ravel(), o.phi_vacuum(P), 2e-6, "vacuum z=2")
    assert np.allclose(py.phi_vacuum_xy(X, Y, 2.0), phi, rtol=0, atol=0)
    z = np.linspace(1.0, 5.0, 50)
    full = cl.eval_full(np.full((1, 50), 0.05, np.float32), np.full((1, 50), 0.05, np.float32),
                        z.reshape(1, -1).astype(np.float32))[0]
    _close(full, o.phi_full(np.stack([np.full(50, 0.05), np.full(50, 0.05), z.astype(np.float32)], 1)), 2e-6, "full 1d")
    assert np.allclose(py.phi_full_1d(0.05, 0.05, z.astype(np.float32)), full, rtol=0, atol=0)


@pytest.mark.parametrize("name,n_harm", [("NaCl_1x1_L3.xyz", 4), ("NaCl_8x8_L3.xyz", 32)])
def test_reference_zscan_against_image_sum(name, n_harm):
    """The 8x8 supercell runs at n_harm 32, the G set of the 1x1 cell at n_harm 4: at n_harm 4 its
    first G with a nonzero structure factor (h = 8) is cut off and the potential above the slab is
    flat, so a correlation is undefined."""
    cl, _, _ = _ref_pair(name, n_harm)
    z = np.linspace(1.0, 5.0, 50).reshape(1, -1).astype(np.float32)
    X = np.full((1, 50), 0.05, np.float32)
    ew = cl.eval_full(X, X, z)[0]
    br = cl.eval_brute(X, X, z, N_rep=20)[0]
    assert EC.correlation(ew, br) > 0.999
    assert EC.rmse(ew - ew.mean(), br - br.mean()) < 0.01


def test_reference_lateral_scans_against_evjen_cluster():
    """x, y and diagonal scans at 2 A over the checkerboard slab against the K=20 Evjen cluster:
    equal up to a constant (the reference reports a 1e-6..3e-6 RMSE floor in FP32)."""
    from spammm_amd.ewald import SurfaceEwald
    rx, ry, rz, q, a, b = EC.load_substrate("NaCl_1x1_L2_checker.xyz")
    cl = SurfaceEwald().prepare_system(EC.ions_of(rx, ry, rz, q), a, b, n_harm=8)
    cluster = EC.evjen_cluster(rx, ry, rz, q, a, b, K=20)
    t = np.linspace(-2.0, 2.0, 80, endpoint=False).astype(np.float32)
    z = np.full_like(t, rz.max() + 2.0)
    zero = np.zeros_like(t)
    for X, Y in ((t, zero), (zero, t), (t, t)):
        ew = cl.eval_full(X, Y, z)
        br = cl.eval_cluster(X, Y, z, cluster)
        d = ew.astype(np.float64) - br
        assert np.abs(d - d.mean()).max() < 2e-6 * np.abs(ew).max() + 1e-7
    ex = cl.eval_full(t, zero, z)
    ey = cl.eval_full(zero, t, z)
    assert np.allclose(ex, ey, atol=1e-6 * np.abs(ex).max())
