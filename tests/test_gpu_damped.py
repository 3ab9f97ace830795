"""GPU parity of the damped-velocity relaxations (the reference's OPT_FIRE=0 build,
AFMulator(use_fire=False), kernels/AFM.cl:378-382, :536-540, :1480-1484) and of relaxStrokes2D
(kernels/AFM.cl:396-450, AFMulator.scan_fdbm_2d spammm/SPM/AFM.py:644-701) against the FP32
op-order-faithful oracle on the same grid: bit-exact (positions, iteration counts, FEs, disps).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O          # noqa: E402
from spammm_amd import afm as A         # noqa: E402
from spammm_amd import device as D      # noqa: E402
from spammm_amd import workloads as W   # noqa: E402
from tests.test_gpu_parity import dev, grid_gpu, tip_of   # noqa: E402

DAMPED_PARS = np.array([0.3, 0.2, 0.03, 0.3], np.float32)   # scan_fdbm(use_fire=False) defaults


def test_damped_strokes_bitexact():
    w = W.small()
    G, sp = grid_gpu(w)
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    dTip = np.array([0., 0., -0.1, 0.], np.float32)
    st, d0 = w.stiffness, np.array([0., 0., -3.0, 3.0], np.float32)
    FEs, disps, PPs = D.relax_strokes(G, sp, torch.as_tensor(pts, device=dev()), w.nz, dTip, st, d0, DAMPED_PARS,
                                      128, 1e-8, damped=True)
    FEo, dio, PPo = O.relax_strokes(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, dTip, st, d0,
                                    DAMPED_PARS, 128, 1e-8, "f32", damped=True)
    np.testing.assert_array_equal(PPs.cpu().numpy(), PPo)
    np.testing.assert_array_equal(FEs.cpu().numpy(), FEo)
    np.testing.assert_array_equal(disps.cpu().numpy(), dio)
    conv = (PPo[..., 3] > 0).mean()
    print(f"[damped strokes] converged {conv:.3f}, mean iterations {np.abs(PPo[..., 3]).mean():.1f}")
    assert conv > 0.5


def test_damped_tilted_bitexact():
    w = W.small()
    G, sp = grid_gpu(w)
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    tip = tip_of(w)
    tip.relax = DAMPED_PARS.copy()
    FEs, PPs = D.relax_tilted(G, sp, torch.as_tensor(pts, device=dev()), w.nz, tip, 128, 1e-8,
                              scan_ny=w.nxy[1], damped=True)
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, tip.tipA, tip.tipB,
                              tip.tipC, tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, 128, 1e-8, False, "f32",
                              damped=True)
    np.testing.assert_array_equal(PPs.cpu().numpy(), PPo)
    np.testing.assert_array_equal(FEs.cpu().numpy(), FEo)


@pytest.mark.parametrize("n_pix", [(12, 10), (48, 40)])
def test_relax_2d_bitexact(n_pix):
    """Both lane-group widths (12x10 pixels: 4 lanes per point; 48x40: still 4) and a raster with a
    partial last tile; K_LAT / dt / damp of scan_fdbm_2d's defaults with a stiffer spring so that most
    states converge inside N_RELAX_STEP_MAX."""
    w = W.small()
    G, sp = grid_gpu(w)
    nx, ny = n_pix
    xs = np.linspace(w.scan_p0[0], w.scan_p0[0] + (w.nxy[0] - 1) * w.scan_da[0], nx).astype(np.float32)
    ys = np.linspace(w.scan_p0[1], w.scan_p0[1] + (w.nxy[1] - 1) * w.scan_db[1], ny).astype(np.float32)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    pts = np.zeros((nx * ny, 4), np.float32)
    pts[:, 0], pts[:, 1], pts[:, 2] = gx.ravel(), gy.ravel(), w.scan_p0[2] - 4.0
    for k_lat in (0.03, 0.5):
        FEs, disps, PPs = D.relax_strokes_2d(G, sp, torch.as_tensor(pts, device=dev()), w.nz, k_lat, 0.1, 0.3, 0.2,
                                             128, 1e-8, scan_ny=ny, want_pps=True)
        FEo, dio, PPo = O.relax_2d(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, k_lat, 0.1, 0.3,
                                   0.2, 128, 1e-8, "f32")
        np.testing.assert_array_equal(PPs.cpu().numpy(), PPo)
        np.testing.assert_array_equal(FEs.cpu().numpy(), FEo)
        np.testing.assert_array_equal(disps.cpu().numpy(), dio)
        print(f"[2d] k_lat {k_lat}: converged {(PPo[..., 3] > 0).mean():.3f}")


def _fdbm_afm(w, use_fire):
    """An AFMulator whose FDBM grid is the small workload's force grid (installed on the device)."""
    afm = A.AFMulator(use_fire=use_fire)
    G, sp = grid_gpu(w)
    step = float(w.L[0] / w.n[0])
    afm.install_fdbm_grid(G, w.p0, step, tuple(int(v) for v in w.n))
    return afm, G, sp


def test_facade_use_fire_false_and_scan_fdbm_2d():
    """AFMulator(use_fire=False).scan_fdbm = the damped strokes kernel (with the damped defaults of
    relax_pars), iz flipped to ascending height; scan_fdbm_2d = relaxStrokes2D from the highest height
    down, flipped; the FIRE instance gives other results."""
    w = W.small()
    afm, G, sp = _fdbm_afm(w, use_fire=False)
    xs = np.linspace(-1.0, 1.0, 6).astype(np.float32)
    ys = np.linspace(-0.8, 0.8, 5).astype(np.float32)
    hs = np.linspace(3.0, 4.0, 6).astype(np.float32)
    FEs, disp = afm.scan_fdbm(xs, ys, hs, mol_z=0.0, use_fire=False, K_LAT=0.03, K_RAD=20.0, bond_length=3.0)
    pts = np.zeros((30, 4), np.float32)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    pts[:, 0], pts[:, 1], pts[:, 2] = gx.ravel(), gy.ravel(), hs.max() + 3.0
    dh = float(hs.max() - hs.min()) / 5
    FEo, dio, _ = O.relax_strokes(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, 6,
                                  np.array([0, 0, -dh, 0], np.float32), np.array([-0.03, -0.03, -0.03, -20.0], np.float32),
                                  np.array([0, 0, -3.0, 3.0], np.float32), DAMPED_PARS, afm.max_iter, afm.f2conv, "f32",
                                  damped=True)
    np.testing.assert_array_equal(FEs, FEo.reshape(6, 5, 6, 4)[:, :, ::-1])
    np.testing.assert_array_equal(disp["dx"], dio.reshape(6, 5, 6, 4)[:, :, ::-1, 0])
    F2, d2 = afm.scan_fdbm_2d(xs, ys, hs, mol_z=0.0, K_LAT=0.03)
    pts[:, 2] = hs.max()
    F2o, d2o, _ = O.relax_2d(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, 6, 0.03, dh, 0.3, 0.2,
                             afm.max_iter, afm.f2conv, "f32")
    np.testing.assert_array_equal(F2, F2o.reshape(6, 5, 6, 4)[:, :, ::-1])
    np.testing.assert_array_equal(d2["dy"], d2o.reshape(6, 5, 6, 4)[:, :, ::-1, 1])
    assert np.all(d2["dz"] == 0)
    afm_f, _, _ = _fdbm_afm(w, use_fire=True)
    FEf, _ = afm_f.scan_fdbm(xs, ys, hs, mol_z=0.0, use_fire=False, K_LAT=0.03, K_RAD=20.0, bond_length=3.0)
    assert not np.array_equal(FEf, FEs)
