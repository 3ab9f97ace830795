"""Substrate placements at their full size: a 60-atom molecule 3 A above a 4-layer rock-salt slab
(8x8 cells, a = 5.64 A, 1024 ions, +-0.8 e, periodic x,y), 902x902x300 grid at 0.05 A, 192x192
pixels x 60 z-steps; placements run on the scan window of the grid.

Bars:
  * the grid of one placement over the slab base (on the window) equals the FP64 oracle grid of
    the combined atom list at 2000 sampled voxels within |d| <= 1e-5 ||(F,E)|| + 4e-6 sum|terms|
    + 1e-6 (the grid bar of tests/test_gpu_poses.py);
  * one placement on the window gives the df of the same placement on the full grid (the window's
    voxels are the full grid's; only the FP32 rounding of its origin differs), and the gradients of a
    loss over the columns stable on both grids within 1e-3 of their norm (near a fold, lambda ~
    1/lambda_min amplifies that rounding);
  * the pose gradient of an L2 loss to the df image of a displaced placement, taken over the pixel
    columns stable in the forward and in the target (diff.column_gate), equals central differences
    of the FP32 forward along -grad within 1e-3 relative at the best-agreeing step of
    eps = 1e-3, 1e-4 A (the reference's plateau rule, tests/SPM/test_afm_morse.py:1670-1716), on
    two placements.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O            # noqa: E402
from spammm_amd import poses as PS        # noqa: E402
from spammm_amd import workloads as W     # noqa: E402
from tests.test_gpu_poses import combined, make_scene   # noqa: E402


@pytest.fixture(scope="module")
def c4():
    sc = W.substrate_scene()
    scene, _ = make_scene(sc, max_iter=1000, gate_eig_frac=0.5, gate_max_jump=0.3, strict=False)
    print(f"[c4] grid {sc['n']}, window {tuple(scene.spec_run.n)}", flush=True)
    return sc, scene


def test_c4_grid_over_base_vs_fp64_oracle(c4):
    sc, scene = c4
    pose = sc["poses"][5]
    q = sc["mol_atoms"][:, 3].astype(np.float64)
    x = PS.place(scene.mol_xyz0, scene.center, pose)
    atoms = torch.as_tensor(np.concatenate([x, q[:, None]], 1).astype(np.float32), device=scene.pts.device)
    from spammm_amd import device as D
    sp = scene.spec_run
    G = D.grid_project(atoms, scene.mol_cpar, 0, sp, sc["Qs"], sc["QZs"], sc["npbc"], sc["lvA"], sc["lvB"], cap=True,
                       base=scene.base)
    nx, ny, nz = sp.n
    rng = np.random.default_rng(11)
    idx = rng.choice(nx * ny * nz, 2000, replace=False)
    ic, ib, ia = np.unravel_index(idx, (nz, ny, nx))
    got = G.reshape(-1, 4)[torch.as_tensor(idx, device=G.device)].cpu().numpy().astype(np.float64)
    ca, cp = combined(sc, pose, q)
    ca32, cp32 = ca.astype(np.float32), cp.astype(np.float32)
    flat = (ia + nx * (ib + ny * ic)).astype(np.int64)
    want = O.grid_voxels(0, ca32, cp32, sp.n, sp.p0, sp.dA, sp.dB, sp.dC, sc["Qs"], sc["QZs"], flat, sc["npbc"],
                         sc["lvA"], sc["lvB"])
    P = (np.asarray(sp.p0, np.float64)[None, :] + np.outer(ia, sp.dA) + np.outer(ib, sp.dB) + np.outer(ic, sp.dC))
    S = O.abs_term_sums(ca32.astype(np.float64), cp32.astype(np.float64), sc["Qs"][0], sc["npbc"], sc["lvA"], sc["lvB"],
                        P)
    scale = np.minimum(1.0, 100.0 / np.maximum(np.abs(want[:, 3]), 1e-300))   # capped voxels scale like the grid
    bound = 1e-5 * np.linalg.norm(want, axis=1) + 4e-6 * S * scale + 1e-6
    err = np.abs(got - want).max(axis=1)
    print(f"[c4] grid over base: max |d| {err.max():.3e}, worst |d|/bound {(err / bound).max():.3f}, "
          f"z range of the sample {P[:, 2].min():.2f}..{P[:, 2].max():.2f} A", flush=True)
    assert (err <= bound).all()


def _loss_grad(scene, sc, k, dpose, columns_of_target=True):
    dev = scene.pts.device
    q0 = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev)
    pose_k = torch.tensor(sc["poses"][k], dtype=torch.float64, device=dev)
    with torch.no_grad():
        target = PS.pose_df(pose_k + dpose, q0, scene).double()
    tcols = scene.last_columns.clone()
    pose = pose_k.clone().requires_grad_(True)
    dfi = PS.pose_df(pose, q0, scene)
    cols = scene.last_columns & tcols if columns_of_target else scene.last_columns.clone()
    loss = 0.5 * (((dfi.double() - target) * cols[..., None]) ** 2).sum()
    loss.backward()
    return pose_k, q0, target, cols, pose.grad.detach(), dfi.detach()


def test_c4_window_equals_full_grid(c4):
    sc, scene = c4
    full, _ = make_scene(sc, max_iter=1000, window_margin=None, strict=False)
    dev = scene.pts.device
    v = torch.as_tensor(np.random.default_rng(8).standard_normal((*sc["nxy"], sc["nz"])), device=dev)
    full.gate_eig_frac = 0.5
    dfs, cols = [], []
    for s in (scene, full):
        with torch.no_grad():
            dfs.append(PS.pose_df(torch.tensor(sc["poses"][3], dtype=torch.float64, device=dev),
                                  torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev), s))
        cols.append(s.last_columns.clone())
    both = (cols[0] & cols[1])[..., None].double()
    out = []
    for s, df0 in zip((scene, full), dfs):
        pose = torch.tensor(sc["poses"][3], dtype=torch.float64, device=dev, requires_grad=True)
        q = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev, requires_grad=True)
        df = PS.pose_df(pose, q, s)
        (df.double() * v * both).sum().backward()
        out.append((df0.cpu().numpy(), pose.grad.cpu().numpy(), q.grad.cpu().numpy()))
    print(f"[c4] columns stable on both grids {int(both.sum())} of {both.numel()}", flush=True)
    del full
    torch.cuda.empty_cache()
    a, b = out
    print(f"[c4] window vs full grid: df {np.abs(a[0] - b[0]).max():.3e} (max {np.abs(b[0]).max():.3e}); "
          f"pose grad {a[1]} vs {b[1]}", flush=True)
    assert np.allclose(a[0], b[0], rtol=1e-4, atol=1e-5 * np.abs(b[0]).max())
    for x, y in zip(a[1:], b[1:]):
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        print(f"[c4] gradient vector, window vs full: |d|/|g| {rel:.2e}", flush=True)
        assert rel <= 1e-3
    assert scene.window_fallbacks == 0


@pytest.mark.parametrize("k", [0, 1])
def test_c4_pose_gradient_directional_fd(c4, k):
    sc, scene = c4
    dpose = torch.tensor([0.1, -0.1, 0.0, 0.02], dtype=torch.float64, device=scene.pts.device)
    pose_k, q0, target, cols, gp, _ = _loss_grad(scene, sc, k, dpose)
    d = -gp / gp.norm()
    ad = -float(gp.norm())
    fds = []
    for eps in (1e-3, 1e-4):
        with torch.no_grad():
            lp = float(0.5 * (((PS.pose_df(pose_k + eps * d, q0, scene).double() - target) * cols[..., None]) ** 2).sum())
            lm = float(0.5 * (((PS.pose_df(pose_k - eps * d, q0, scene).double() - target) * cols[..., None]) ** 2).sum())
        fds.append((lp - lm) / (2 * eps))
    rel = [abs(ad - f) / abs(f) for f in fds]
    print(f"[c4] placement {k}: columns kept {int(cols.sum())} of {cols.numel()}; adjoint {ad:.6e}, "
          f"FD {fds}, rel {rel}", flush=True)
    assert cols.sum() >= 0.8 * cols.numel()
    assert min(rel) <= 1e-3


def test_c4_batch_equals_one_at_a_time(c4):
    """Four placements through pose_df_batch (two in flight, grids kept) at full size: each df
    image equals pose_df's bit for bit and the pose gradients agree to the rounding of the relax
    adjoint's atomic scatter (rtol 1e-5)."""
    sc, scene = c4
    dev = scene.pts.device
    P = torch.tensor(np.asarray(sc["poses"][4:8]), dtype=torch.float64, device=dev)
    q0 = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev)
    v = torch.as_tensor(np.random.default_rng(10).standard_normal((4, *sc["nxy"], sc["nz"])), device=dev)
    ref = []
    for k in range(4):
        pose = P[k].clone().requires_grad_(True)
        df = PS.pose_df(pose, q0, scene)
        (df.double() * v[k]).sum().backward()
        ref.append((df.detach(), pose.grad))
    poses = P.clone().requires_grad_(True)
    df = PS.pose_df_batch(poses, q0, scene)
    (df.double() * v).sum().backward()
    for k, (d, g) in enumerate(ref):
        rel = float((poses.grad[k] - g).norm() / g.norm())
        print(f"[c4] batch placement {k}: df equal {bool(torch.equal(df[k].detach(), d))}, pose grad rel {rel:.2e}",
              flush=True)
        assert torch.equal(df[k].detach(), d)
        assert torch.allclose(poses.grad[k], g, rtol=1e-5, atol=1e-6 * float(g.abs().max()))
