"""FP64 reference-precision forward: grid projection and relaxed scan in double on the device.

Bars: the FP64 grid equals the FP64 oracle grid (oracle/ppm_oracle.cpp, every operation in double)
per voxel within 1e-11 ||(F,E)|| + 1e-12 (summation order only); the FP64 relaxation on the
oracle's FP64 grid has the oracle's iteration counts and convergence flags at every state and
probe positions / FEs within 1e-9 (same operation sequence; device vs host libm only).
Closed forms pin the oracle itself (tests/test_oracle_pair.py, tests/test_oracle_relax.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O            # noqa: E402
from spammm_amd import device as D        # noqa: E402
from spammm_amd import workloads as W     # noqa: E402
from spammm_amd import params as P        # noqa: E402


def _grid_case(model, tip_sites):
    w = W.small()
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    atoms = w.atoms.astype(np.float64)
    if model == 0:
        cpar = w.cLJ.astype(np.float64)
    else:
        cpar = P.mix_morse(w.enames, P.ELEMENTS_REFERENCE, 1.452, 0.0006808, -1.8).astype(np.float64)
    Qs, QZs = tip_sites
    lvA, lvB = np.array([12.0, 0.0, 0.0]), np.array([0.0, 10.0, 0.0])
    return w, sp, atoms, cpar, np.asarray(Qs, float), np.asarray(QZs, float), (1, 1), lvA, lvB


@pytest.mark.parametrize("model,tip_sites", [(0, ((-0.05, 0, 0, 0), (0, 0, 0, 0))),
                                             (1, ((-0.1, 0.05, 0, 0), (0.0, 0.4, 0, 0)))])
def test_grid_f64_matches_fp64_oracle(model, tip_sites):
    w, sp, atoms, cpar, Qs, QZs, npbc, lvA, lvB = _grid_case(model, tip_sites)
    dev = D.require_gpu()
    G = D.grid_project_f64(torch.as_tensor(atoms, device=dev), torch.as_tensor(cpar, device=dev), model, sp, Qs, QZs,
                           npbc, lvA, lvB, cap=True).cpu().numpy()
    Go = O.grid_project_d(model, atoms, cpar, sp.n, sp.p0, sp.dA, sp.dB, sp.dC, Qs, QZs, npbc, lvA, lvB)
    mag = np.linalg.norm(Go, axis=-1, keepdims=True)
    err = np.abs(G - Go)
    print(f"[f64] model {model}: max |d| {err.max():.3e}, max |d|/|G| {(err / np.maximum(mag, 1e-300)).max():.3e}")
    assert (err <= 1e-11 * mag + 1e-12).all()


def test_relax_f64_matches_fp64_oracle():
    w, sp, atoms, cpar, Qs, QZs, npbc, lvA, lvB = _grid_case(0, ((-0.05, 0, 0, 0), (0, 0, 0, 0)))
    Go = O.grid_project_d(0, atoms, cpar, sp.n, sp.p0, sp.dA, sp.dB, sp.dC, Qs, QZs, npbc, lvA, lvB)
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy).astype(np.float64)
    dev = D.require_gpu()
    for check_before, f2conv in ((True, 1e-24), (False, 1e-8)):
        FEs, PPs = D.relax_tilted_f64(torch.as_tensor(Go, device=dev), sp, torch.as_tensor(pts, device=dev), w.nz, tip,
                                      max_iter=40000, f2conv=f2conv, check_before=check_before)
        FEo, PPo = O.relax_tilted_d(Go, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, tip.tipA, tip.tipB, tip.tipC,
                                    tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, max_iter=40000,
                                    f2conv=np.float32(f2conv), check_before=check_before)
        FEs, PPs = FEs.cpu().numpy(), PPs.cpu().numpy()
        dp = np.abs(PPs[..., :3] - PPo[..., :3]).max()
        df = np.abs(FEs - FEo).max()
        print(f"[f64] relax check_before={check_before}: max |dPP| {dp:.3e}, max |dFE| {df:.3e}, "
              f"mean iters {np.abs(PPo[..., 3]).mean():.1f}")
        assert np.array_equal(PPs[..., 3], PPo[..., 3])
        assert dp <= 1e-9 and df <= 1e-9


def test_relax_f64_f32_anchors_follow_fp32_scan_heights():
    """f32_anchors: the FP64 scan's tip anchors advance in float32 exactly as the FP32 kernel's
    (tip += C*dtip in float32).  On a zero grid the relaxed probe sits at anchor + dpos0, so the
    probe heights expose the anchors: float32-accumulated heights with the flag, p0 + iz*dz in
    FP64 without it (over 40 slices of -0.1 A from 9.5 A they differ by up to ~1e-5 A)."""
    w, sp, *_ = _grid_case(0, ((-0.05, 0, 0, 0), (0, 0, 0, 0)))
    nz = 40
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., -0.1], np.float32))
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, 3, 2).astype(np.float64)
    dev = D.require_gpu()
    G0 = torch.zeros((sp.n[2], sp.n[1], sp.n[0], 4), dtype=torch.float64, device=dev)
    z32 = np.empty(nz)
    t = np.float32(pts[0, 2])
    for k in range(nz):
        z32[k] = float(t)
        t = np.float32(t + np.float32(np.float32(1.0) * np.float32(-0.1)))
    z64 = pts[0, 2] + np.arange(nz) * float(np.float32(-0.1))
    d0z = float(np.float32(w.dpos0[2]))
    for f32, ref in ((True, z32), (False, z64)):
        _, PPs = D.relax_tilted_f64(G0, sp, torch.as_tensor(pts, device=dev), nz, tip, max_iter=40000, f2conv=1e-24,
                                    check_before=True, f32_anchors=f32)
        PPz = PPs[..., 2].cpu().numpy()
        err = np.abs(PPz - (ref[None, :] + d0z)).max()
        print(f"[f64] f32_anchors={f32}: max |z - (anchor + d0)| {err:.2e}")
        assert err <= 1e-9
    assert np.abs(z32 - z64).max() > 1e-6


def test_fp32_scan_with_polish_matches_fp64_scan_at_fp32_anchors():
    """The FP32 scan, its states polished by FP64 Newton steps on the trilinear model, against
    the FP64 scan of the same FP32 grid (widened) with f32_anchors: the same heights, so the
    polished probes and their Fz agree to FP32 resolution; with FP64 anchors (exact heights, up
    to ~1e-5 A away over 40 slices) Fz differs by the anchor drift times the force gradient."""
    w = W.small()
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    dev = D.require_gpu()
    G = D.grid_project(torch.as_tensor(w.atoms, device=dev), torch.as_tensor(w.cLJ, device=dev), 0, sp, w.Qs,
                       w.QZs, w.npbc, w.lvA, w.lvB, cap=True)
    nz = 40
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., -0.1], np.float32))
    pts = torch.as_tensor(D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy), device=dev)
    FEs, PPs = D.relax_tilted(G, sp, pts, nz, tip, 1000, 1e-12, True)
    D.relax_polish(G, sp, pts, nz, tip, PPs, FEs, 3)
    G64, pts64 = G.double().contiguous(), pts.double().contiguous()
    out = {}
    for f32 in (True, False):
        F64, P64 = D.relax_tilted_f64(G64, sp, pts64, nz, tip, max_iter=40000, f2conv=1e-24, check_before=True,
                                      f32_anchors=f32)
        dz = (FEs[..., 2].double() - F64[..., 2]).abs().cpu().numpy()
        dp = (PPs[..., :3].double() - P64[..., :3]).abs().amax(-1).cpu().numpy()
        out[f32] = (float(np.median(dz)), float(np.median(dp)))
        print(f"[f64] f32_anchors={f32}: median |dFz| {out[f32][0]:.2e} eV/A, median |dPP| {out[f32][1]:.2e} A")
    assert (PPs[..., 3] > 0).float().mean().item() > 0.99
    assert out[True][0] <= 1e-8 and out[True][1] <= 1e-6
    assert out[False][0] >= 3 * out[True][0]


def test_f64_entry_points_reject_fp32_tensors():
    w, sp, atoms, cpar, Qs, QZs, npbc, lvA, lvB = _grid_case(0, ((-0.05, 0, 0, 0), (0, 0, 0, 0)))
    dev = D.require_gpu()
    with pytest.raises(ValueError):
        D.grid_project_f64(torch.as_tensor(atoms, dtype=torch.float32, device=dev),
                           torch.as_tensor(cpar, device=dev), 0, sp, Qs, QZs)
