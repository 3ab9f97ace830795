"""The north-star accuracy bound at full size: the naphthalene (200^3 grid, 100x100x40 scan) and
flake (300x300x240 grid, 256x256x60 scan) workloads, a 256-pixel subsample relaxed to |F| < 1e-6
eV/A (check before the update, up to 4000 iterations).

The FP64 oracle advances its tip anchors in float32 as the FP32 kernel does (kernels/AFM.cl:556), so
both relax at the same heights (their drift from p0 + iz*dz, ~1e-5 A over 60 slices, moves a soft
probe by up to k/lambda_min times as much).  A FIRE stop at |F| < 1e-6 leaves a probe |F|/lambda_min
from its equilibrium, lambda_min the
smallest eigenvalue of the total Hessian; on soft states (lambda_min well below k_lat = 0.0156 eV/A^2)
that exceeds 1e-4 A, so the position bound is asserted on the forward with the FP64 Newton polish
(ppm_relax_polish, 3 steps on the trilinear model), the accuracy form of the library, and the FIRE
result is reported beside it with the lambda_min of its worst states.

Bars:
  * the FP32 relaxation + polish on the device vs the FP64 oracle relaxation (to |F| < 1e-10 eV/A)
    on the same FP32 grid: probe positions within 1e-4 A and Fz within 1e-4 eV/A at every state both
    converge (the bound of tests/test_gpu_parity.py::test_relax_vs_fp64_oracle_tight_convergence,
    at the workloads' own sizes); FIRE alone: Fz within 1e-4 eV/A;
  * the whole FP32 path (FP32 grid + FP32 relaxation + polish) vs the whole FP64 path (the FP64 grid
    of ppm_grid_project_f64 -- equal to the FP64 oracle grid to 1e-13, tests/test_gpu_f64.py --
    relaxed by the FP64 oracle): the same bound.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O            # noqa: E402
from spammm_amd import device as D        # noqa: E402
from spammm_amd import workloads as W     # noqa: E402
from tests.test_gpu_parity import dev, grid_gpu, tip_of   # noqa: E402


@pytest.mark.parametrize("name", ["naphthalene", "flake"])
def test_fp32_path_within_north_star_bound_at_full_size(name):
    w = getattr(W, name)()
    G, sp = grid_gpu(w)
    d = dev()
    pts_all = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    sel = np.sort(np.random.default_rng(5).choice(w.n_scan, 256, replace=False))
    pts = pts_all[sel]
    tip = tip_of(w)
    pts_d = torch.as_tensor(pts, device=d)
    FEf, PPf = D.relax_tilted(G, sp, pts_d, w.nz, tip, 4000, 1e-12, True)
    diag = D.relax_state_diag(G, sp, pts_d, w.nz, tip, PPf).cpu().numpy()
    FEs, PPs = FEf.clone(), PPf.clone()
    D.relax_polish(G, sp, pts_d, w.nz, tip, PPs, FEs, 3)
    FEf, PPf, FEs, PPs = FEf.cpu().numpy(), PPf.cpu().numpy(), FEs.cpu().numpy(), PPs.cpu().numpy()
    args = (sp.n, sp.dinvA, sp.dinvB, sp.dinvC)
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), *args, pts, w.nz, tip.tipA, tip.tipB, tip.tipC, tip.stiffness,
                              tip.dpos0, tip.relax, tip.surfFF, 20000, 1e-20, True, "f64", f32_anchors=True)
    conv = (PPs[..., 3] > 0) & (PPo[..., 3] > 0)
    dfire = np.abs(PPf[..., :3] - PPo[..., :3]).max(-1)
    worst = np.argsort(np.where(conv, dfire, 0).ravel())[::-1][:3]
    print(f"[{name}] FIRE alone vs FP64 oracle on the FP32 grid: PP max {dfire[conv].max():.2e} median "
          f"{np.median(dfire[conv]):.2e} A, Fz max {np.abs(FEf[..., 2] - FEo[..., 2])[conv].max():.2e} eV/A; "
          f"worst states |dPP| {dfire.ravel()[worst]} at lambda_min {diag[..., 1].ravel()[worst]} eV/A^2 "
          f"(k_lat {abs(float(tip.stiffness[0])):.4f})", flush=True)
    assert np.abs(FEf[..., 2] - FEo[..., 2])[conv].max() <= 1e-4
    dpos = np.abs(PPs[..., :3] - PPo[..., :3]).max(-1)[conv]
    dFz = np.abs(FEs[..., 2] - FEo[..., 2])[conv]
    print(f"[{name}] FP32 relaxation + polish vs FP64 oracle on the FP32 grid: converged {conv.mean():.4f}; "
          f"PP max {dpos.max():.2e} median {np.median(dpos):.2e} A; Fz max {dFz.max():.2e} eV/A", flush=True)
    assert conv.mean() > 0.98
    assert dpos.max() <= 1e-4 and dFz.max() <= 1e-4
    # the whole FP64 path: FP64 grid (device kernel) + FP64 oracle relaxation
    atoms64 = torch.as_tensor(w.atoms.astype(np.float64), device=d)
    cp64 = torch.as_tensor(w.cLJ.astype(np.float64), device=d)
    G64 = D.grid_project_f64(atoms64, cp64, 0, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB).cpu().numpy()
    FEd, PPd = O.relax_tilted_d(G64, *args, pts.astype(np.float64), w.nz, tip.tipA, tip.tipB, tip.tipC,
                                tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, max_iter=20000, f2conv=1e-20,
                                check_before=True, f32_anchors=True)
    del G64
    conv2 = (PPs[..., 3] > 0) & (PPd[..., 3] > 0)
    dpos2 = np.abs(PPs[..., :3] - PPd[..., :3]).max(-1)[conv2]
    dFz2 = np.abs(FEs[..., 2] - FEd[..., 2])[conv2]
    worst2 = np.argsort(np.where(conv2, np.abs(PPs[..., :3] - PPd[..., :3]).max(-1), 0).ravel())[::-1][:3]
    print(f"[{name}] FP32 path + polish vs FP64 path (FP64 grid + FP64 relaxation): converged {conv2.mean():.4f}; "
          f"PP max {dpos2.max():.2e} median {np.median(dpos2):.2e} A; Fz max {dFz2.max():.2e} "
          f"median {np.median(dFz2):.2e} eV/A; worst states at lambda_min {diag[..., 1].ravel()[worst2]}", flush=True)
    assert conv2.mean() > 0.98
    assert dpos2.max() <= 1e-4 and dFz2.max() <= 1e-4
