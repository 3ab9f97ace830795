"""Batched placements (poses.pose_df_batch): several placements in flight on their own streams give,
placement by placement, the df image of pose_df bit for bit and its gradients w.r.t. the
placement and the (shared) charges to the rounding of the relax adjoint's atomic scatter (rtol
1e-5, the bar of tests/test_gpu_poses.py::test_kept_states_equal_recomputed_states), with the
grids kept for the backward or recomputed there, with the column gate, and with placements that
fall back to the full grid.  Reduced substrate scene (45x45x52 grid, 10x10 pixels x 8 z-steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spammm_amd import poses as PS        # noqa: E402
from spammm_amd import workloads as W     # noqa: E402
from tests.test_gpu_poses import make_scene   # noqa: E402


def _poses(sc):
    extra = sc["poses"][:2] + np.array([0.05, -0.05, 0.0, 0.1])
    return np.concatenate([sc["poses"], extra])


def _per_placement(scene, sc, P, v):
    dev = scene.pts.device
    out = []
    for k in range(P.shape[0]):
        pose = torch.tensor(P[k], dtype=torch.float64, device=dev, requires_grad=True)
        q = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev, requires_grad=True)
        df = PS.pose_df(pose, q, scene)
        cols = scene.last_columns.clone() if scene.gate_eig_frac > 0 else None
        (df.double() * v[k]).sum().backward()
        out.append((df.detach(), pose.grad, q.grad, cols))
    return out


def _batched(scene, sc, P, v):
    dev = scene.pts.device
    poses = torch.tensor(P, dtype=torch.float64, device=dev, requires_grad=True)
    q = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev, requires_grad=True)
    df = PS.pose_df_batch(poses, q, scene)
    cols = scene.last_columns.clone() if scene.gate_eig_frac > 0 else None
    (df.double() * v).sum().backward()
    return df.detach(), poses.grad, q.grad, cols


def _check(ref, got, tag):
    df, gp, gq, cols = got
    for k, (d, p, q, c) in enumerate(ref):
        assert torch.equal(df[k], d), (tag, k)
        if c is not None:
            assert torch.equal(cols[k], c), (tag, k)
        assert torch.allclose(gp[k], p, rtol=1e-5, atol=1e-6 * float(p.abs().max())), (tag, k, gp[k], p)
    gq_ref = sum(r[2] for r in ref)
    dp = max(float((gp[k] - r[1]).abs().max()) for k, r in enumerate(ref))
    print(f"[poses batch] {tag}: pose grads max |d| {dp:.3e}, charge grad max |d| "
          f"{float((gq - gq_ref).abs().max()):.3e} (max {float(gq_ref.abs().max()):.3e})", flush=True)
    assert torch.allclose(gq, gq_ref, rtol=1e-5, atol=1e-6 * float(gq_ref.abs().max()))


@pytest.mark.parametrize("depth,budget", [(2, None), (3, 0.0)])
def test_batch_equals_per_placement(depth, budget):
    sc = W.small_substrate_scene()
    scene, _ = make_scene(sc, gate_eig_frac=0.5)
    scene.batch_depth = depth
    if budget is not None:
        scene.batch_grid_budget = budget          # no grid kept: the backward recomputes every grid
    P = _poses(sc)
    v = torch.as_tensor(np.random.default_rng(12).standard_normal((P.shape[0], *sc["nxy"], sc["nz"])),
                        device=scene.pts.device)
    ref = _per_placement(scene, sc, P, v)
    got = _batched(scene, sc, P, v)
    _check(ref, got, f"depth {depth}, grid budget {budget}")
    assert scene.window_fallbacks == 0
    with torch.no_grad():                          # no backward: no grid kept, same images
        df = PS.pose_df_batch(torch.tensor(P, dtype=torch.float64, device=scene.pts.device),
                              torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=scene.pts.device),
                              scene)
    assert torch.equal(df, got[0])


def test_batch_with_full_grid_fallback():
    """Zero window margins: every placement's probes come within the guard of a cut face and the
    placement is recomputed on the full grid, in the batch as in pose_df."""
    sc = W.small_substrate_scene()
    tight, _ = make_scene(sc, window_margin=(0.0, 0.0))
    assert tight.windowed
    P = _poses(sc)[:3]
    v = torch.as_tensor(np.random.default_rng(13).standard_normal((P.shape[0], *sc["nxy"], sc["nz"])),
                        device=tight.pts.device)
    ref = _per_placement(tight, sc, P, v)
    n_ref = tight.window_fallbacks
    got = _batched(tight, sc, P, v)
    _check(ref, got, "full-grid fallback")
    assert n_ref >= 1 and tight.window_fallbacks == 2 * n_ref


def test_batch_with_per_placement_charges():
    """Charges (B, na), one row per placement: each row's gradient is that placement's q.grad."""
    sc = W.small_substrate_scene()
    scene, _ = make_scene(sc)
    dev = scene.pts.device
    P = _poses(sc)[:4]
    q0 = sc["mol_atoms"][:, 3]
    Q = np.stack([q0 * (1.0 + 0.1 * k) for k in range(P.shape[0])])
    v = torch.as_tensor(np.random.default_rng(14).standard_normal((P.shape[0], *sc["nxy"], sc["nz"])), device=dev)
    ref = []
    for k in range(P.shape[0]):
        pose = torch.tensor(P[k], dtype=torch.float64, device=dev, requires_grad=True)
        q = torch.tensor(Q[k], dtype=torch.float64, device=dev, requires_grad=True)
        df = PS.pose_df(pose, q, scene)
        (df.double() * v[k]).sum().backward()
        ref.append((df.detach(), pose.grad, q.grad))
    poses = torch.tensor(P, dtype=torch.float64, device=dev, requires_grad=True)
    qb = torch.tensor(Q, dtype=torch.float64, device=dev, requires_grad=True)
    df = PS.pose_df_batch(poses, qb, scene)
    (df.double() * v).sum().backward()
    assert qb.grad.shape == qb.shape
    for k, (d, gp, gq) in enumerate(ref):
        assert torch.equal(df[k].detach(), d), k
        assert torch.allclose(poses.grad[k], gp, rtol=1e-5, atol=1e-6 * float(gp.abs().max())), k
        assert torch.allclose(qb.grad[k], gq, rtol=1e-5, atol=1e-6 * float(gq.abs().max())), k
