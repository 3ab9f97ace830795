"""The differentiable df image at its full size: the 200-atom flake on its 300x300x240 grid,
128x128 pixels x 40 z-steps from tip height 10 A (probe 6 -> 2.1 A), L2 loss to the df image of
the same flake with positions displaced by N(0, 0.1 A).

At these heights tens of thousands of probe states sit near folds of their equilibrium branch:
a parameter change of 1e-3 A moves them to another branch and their df column jumps, so the loss
is only piecewise smooth there.  The loss is therefore taken over the pixel columns that are
stable (diff.stable_columns: smallest Hessian eigenvalue >= 0.5 k_lat at every slice, no probe
jump >= 0.3 A between slices, no state the adjoint gates) in the forward at the evaluated
parameters and in the target's forward; the mask is held fixed for the finite differences.

The forward runs on the scan window: the z planes of the probe heights +- 1 A (59 of 240; the
scan covers the whole periodic x, y cell), the grid's own voxels.

Bars:
  * the window gives the whole grid's df (rtol 1e-4) and the same gradient (1e-3 of its norm) on
    the columns stable on both (near a fold the FP32 rounding of the window's origin may put a probe
    on the other branch);
  * adjoint vs central differences of the FP64 device forward (gradcheck.Fp64Forward) on 32
    random parameters (seed 3; positions and charges): within 1e-3 relative for at least 31 of 32
    by the reference's plateau rule (best-agreeing step of h = 1e-3, 1e-4, 1e-5 A / 1e-4, 1e-5,
    1e-6 e, tests/SPM/test_afm_morse.py:1670-1716), and at h = 1e-4 A / 1e-5 e alone;
  * 50 Adam steps (lr 1e-2) on the positions lower the loss at least 20x and bring both the
    lateral and the vertical RMSD to the displaced positions down.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spammm_amd import device as D        # noqa: E402
from spammm_amd import diff               # noqa: E402
from spammm_amd import gradcheck as GC    # noqa: E402
from spammm_amd import params as P        # noqa: E402
from spammm_amd import workloads as W     # noqa: E402


def c3_case(tip_z0=10.0, window=(2.0, 1.0)):
    dev = D.require_gpu()
    w = W.flake_differentiable(tip_z0)
    prob = diff.Problem.from_workload(w, dev, max_iter=1000, f2conv=1e-12, check_before=True, polish=3, strict=False,
                                      gate_eig_frac=0.5, gate_max_jump=0.3, window_margin=window)
    R0, E0 = P.mix_R0E0(w.enames, P.ELEMENTS_BENCH, 1.66, 0.0091)
    R0, E0 = torch.tensor(R0, device=dev), torch.tensor(E0, device=dev)
    xyz0 = torch.tensor(w.atoms[:, :3], dtype=torch.float64, device=dev)
    q0 = torch.tensor(w.atoms[:, 3], dtype=torch.float64, device=dev)
    xyz_t = torch.tensor(W.perturbed_positions(w), device=dev)
    with torch.no_grad():
        target = diff.df_image(xyz_t, q0, R0, E0, prob).double()
    tcols = prob.last_columns.clone()
    return w, prob, R0, E0, xyz0, q0, xyz_t, target, tcols


def test_c3_window_equals_full_grid():
    """The forward on the scan window (z planes of the probe heights +- 1 A) gives the df of the
    whole grid (the window's voxels are the whole grid's; only the FP32 rounding of its origin
    differs), and the gradient of a loss over the columns stable on both within 1e-3 of its norm."""
    out = []
    for window in ((2.0, 1.0), None):
        w, prob, R0, E0, xyz0, q0, xyz_t, target, tcols = c3_case(window=window)
        x = xyz0.clone().requires_grad_(True)
        q = q0.clone().requires_grad_(True)
        dfi = diff.df_image(x, q, R0, E0, prob)
        out.append((prob, dfi.detach(), prob.last_columns & tcols, x, q, target))
    (pw, dw, cw, _, _, tw), (pf, df_, cf, _, _, tf) = out
    both = cw & cf
    m = both[..., None].expand_as(dw)
    print(f"[c3 window] grid {pw.last_spec.n} of {pf.last_spec.n}, fallbacks {pw.window_fallbacks}; df max |d| "
          f"{float((dw - df_)[m].abs().max()):.3e} on the stable columns, {float((dw - df_).abs().max()):.3e} on all "
          f"(max |df| {float(df_.abs().max()):.3e})", flush=True)
    assert pw.window_fallbacks == 0 and tuple(pw.last_spec.n) != tuple(pf.last_spec.n)
    assert torch.allclose(dw[m], df_[m], rtol=1e-4, atol=1e-5 * float(df_.abs().max()))
    grads = []
    for prob, dfi0, _, x, q, target in out:
        dfi = diff.df_image(x, q, R0, E0, prob)
        diff.masked_l2(dfi, tf, both).backward()
        grads.append(torch.cat([x.grad.flatten(), q.grad]))
    rel = float((grads[0] - grads[1]).norm() / grads[1].norm())
    print(f"[c3 window] columns stable on both {int(both.sum())}; gradient |d|/|g| {rel:.2e}", flush=True)
    assert rel <= 1e-3


def test_c3_adjoint_matches_fp64_fd_at_configured_heights():
    w, prob, R0, E0, xyz0, q0, xyz_t, target, tcols = c3_case()
    x = xyz0.clone().requires_grad_(True)
    q = q0.clone().requires_grad_(True)
    dfi = diff.df_image(x, q, R0, E0, prob)
    cols = prob.last_columns & tcols
    diff.masked_l2(dfi, target, cols).backward()
    kept = int(cols.sum())
    print(f"[c3] columns kept {kept} of {cols.numel()} (forward gate {prob.last_gate}); "
          f"adjoint status {prob.last_status}", flush=True)
    assert kept >= 0.6 * cols.numel()
    assert prob.last_status["nonfinite"] == 0
    na = xyz0.shape[0]
    sel = np.random.default_rng(3).choice(4 * na, 32, replace=False)
    res = GC.fd_check(GC.Fp64Forward(prob, R0, E0), target.reshape(prob.n_scan, prob.nz), xyz0, q0, x.grad, q.grad,
                      sel, columns=cols)
    for s in res["per_step"]:
        print(f"[c3] FD h = {s['h']}: {s['within_tol']}/32 within 1e-3, median {s['median_rel_err']:.2e}, "
              f"max {s['max_rel_err']:.2e}", flush=True)
    print(f"[c3] plateau: {res['plateau_within_tol']}/32, max {res['plateau_max_rel_err']:.2e}", flush=True)
    assert res["plateau_within_tol"] >= 31
    assert res["per_step"][1]["within_tol"] >= 31


def test_c3_adam_recovers_positions():
    w, prob, R0, E0, xyz0, q0, xyz_t, target, tcols = c3_case()
    x = xyz0.clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=1e-2)
    losses = []
    for it in range(50):
        opt.zero_grad()
        dfi = diff.df_image(x, q0, R0, E0, prob)
        L = diff.masked_l2(dfi, target, prob.last_columns & tcols)
        L.backward()
        opt.step()
        losses.append(float(L.detach()))
    d0, d1 = xyz0 - xyz_t, x.detach() - xyz_t
    rxy = [float(d[:, :2].pow(2).sum(1).mean().sqrt()) for d in (d0, d1)]
    rz = [float(d[:, 2].pow(2).mean().sqrt()) for d in (d0, d1)]
    print(f"[c3 adam] loss {losses[0]:.4g} -> {losses[-1]:.4g}; RMSD xy {rxy[0]:.4f} -> {rxy[1]:.4f} A, "
          f"z {rz[0]:.4f} -> {rz[1]:.4f} A", flush=True)
    assert losses[-1] < losses[0] / 20
    assert rxy[1] < rxy[0] and rz[1] < 0.7 * rz[0]
