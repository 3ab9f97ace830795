"""GPU adjoint (implicit-function relax adjoint + grid->atom adjoint, torch.autograd.Function)
against central finite differences of the FP64 oracle forward.

Forward for the finite differences: FP64 grid from FP64 atoms/parameters, FP64 relaxation
on that grid converged to |F| < 1e-12 eV/A (check before the update), df by the Giessibl
operator (FP64 matrix, pinned against quadrature in tests/test_oracle_df.py).
Bar: every checked gradient component within 1e-3 relative of its FD value (denominator
floored at 1e-2 of the largest |FD| component so that near-zero components are judged on
the gradient's scale).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O            # noqa: E402
from spammm_amd import device as D        # noqa: E402
from spammm_amd import df as DF           # noqa: E402
from spammm_amd import diff               # noqa: E402
from spammm_amd import params as P        # noqa: E402
from spammm_amd import workloads as W     # noqa: E402


def setup_case(seed=7):
    w = W.small(seed)
    dev = D.require_gpu()
    prob = diff.Problem.from_workload(w, dev, max_iter=3000, f2conv=1e-12, check_before=True)
    R0, E0 = P.mix_R0E0(w.enames, P.ELEMENTS_BENCH, 1.66, 0.0091)
    theta = {
        "xyz": w.atoms[:, :3].astype(np.float64), "q": w.atoms[:, 3].astype(np.float64),
        "R0": R0, "E0": E0,
    }
    return w, prob, theta


def gpu_df(prob, th, requires_grad=False):
    dev = prob.pts.device
    t = {k: torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=requires_grad) for k, v in th.items()}
    df = diff.df_image(t["xyz"], t["q"], t["R0"], t["E0"], prob)
    return df, t


def fp64_df(w, prob, th):
    atoms = np.concatenate([th["xyz"], th["q"][:, None]], 1)
    cpar = np.stack([2 * th["E0"] * th["R0"] ** 6, th["E0"] * th["R0"] ** 12], 1)
    sp = prob.spec
    G = O.grid_project_d(0, atoms, cpar, sp.n, sp.p0, sp.dA, sp.dB, sp.dC, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB)
    tip = prob.tip
    pts = prob.pts.cpu().numpy()
    FEs, PPs = O.relax_tilted_d(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, prob.nz, tip.tipA, tip.tipB, tip.tipC,
                                tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, max_iter=40000, f2conv=1e-24,
                                check_before=True)
    assert (PPs[..., 3] > 0).all(), "FP64 reference relaxation did not converge"
    M = DF.giessibl_matrix(prob.nz, prob.dz, prob.amp, prob.f0, prob.k_cant)
    return (FEs[..., 2] @ M.T).reshape(prob.nxy[0], prob.nxy[1], prob.nz)


def test_adjoint_matches_fp64_finite_differences():
    w, prob, th = setup_case()
    rng = np.random.default_rng(3)
    # target: df of the same molecule with positions perturbed by N(0, 0.1 A)
    th_t = {k: v.copy() for k, v in th.items()}
    th_t["xyz"] = th["xyz"] + rng.normal(0, 0.1, th["xyz"].shape)
    with torch.no_grad():
        target = gpu_df(prob, th_t)[0].double()
    df, t = gpu_df(prob, th, requires_grad=True)
    loss = 0.5 * ((df.double() - target) ** 2).sum()
    loss.backward()
    print(f"[adjoint] loss {float(loss.detach()):.6e} status {prob.last_status}")
    assert prob.last_status["gated"] == 0
    grads = {k: t[k].grad.cpu().numpy() for k in t}
    tgt = target.cpu().numpy()
    d64 = fp64_df(w, prob, th)
    dgpu = df.detach().cpu().numpy()
    print(f"[adjoint] forward df GPU vs FP64: max {np.abs(dgpu - d64).max():.3e} "
          f"(scale {np.abs(d64).max():.3e}), residual scale {np.abs(d64 - tgt).max():.3e}")

    # The adjoint is the vector-Jacobian product v^T d(df)/d(theta) with v = dL/ddf of the GPU forward.
    # Finite differences of phi(theta) = v . df_FP64(theta) with the same fixed v check that product
    # directly; the FP32 forward's own df error (~1e-5 of df) does not enter.
    v = dgpu - tgt

    def L64(thx):
        return float((v * fp64_df(w, prob, thx)).sum())

    na = th["q"].size
    picks = []
    for k in ("xyz", "q", "R0", "E0"):
        for _ in range({"xyz": 14, "q": 6, "R0": 6, "E0": 6}[k]):
            i = int(rng.integers(na))
            picks.append((k, (i, int(rng.integers(3))) if k == "xyz" else (i,)))
    fd, an = [], []
    for k, idx in picks:
        h = {"xyz": 1e-3, "q": 1e-4, "R0": 1e-3, "E0": 1e-3 * th["E0"][idx[0]]}[k]
        tp = {kk: v.copy() for kk, v in th.items()}
        tm = {kk: v.copy() for kk, v in th.items()}
        tp[k][idx] += h
        tm[k][idx] -= h
        fd.append((L64(tp) - L64(tm)) / (2 * h))
        an.append(grads[k][idx])
    fd, an = np.array(fd), np.array(an)
    floor = 1e-2 * np.abs(fd).max()
    rel = np.abs(an - fd) / np.maximum(np.abs(fd), floor)
    for (k, idx), a, f, r in zip(picks, an, fd, rel):
        print(f"  {k}{idx}: adjoint {a:+.6e}  fd {f:+.6e}  rel {r:.2e}")
    assert rel.max() < 1e-3, f"worst relative error {rel.max():.3e}"


def test_adjoint_descent_direction_on_loss():
    """Along -grad, the loss of the GPU forward decreases at the rate |grad| (directional FD, eps=1e-3 A)."""
    w, prob, th = setup_case(11)
    rng = np.random.default_rng(2)
    th_t = {k: v.copy() for k, v in th.items()}
    th_t["xyz"] = th["xyz"] + rng.normal(0, 0.1, th["xyz"].shape)
    with torch.no_grad():
        target = gpu_df(prob, th_t)[0].double()
    df, t = gpu_df(prob, th, requires_grad=True)
    (0.5 * ((df.double() - target) ** 2).sum()).backward()
    g = t["xyz"].grad.cpu().numpy()
    d = -g / np.linalg.norm(g)
    eps = 1e-3

    def loss_at(x):
        th2 = dict(th)
        th2["xyz"] = x
        with torch.no_grad():
            return float(0.5 * ((gpu_df(prob, th2)[0].double() - target) ** 2).sum())
    slope = (loss_at(th["xyz"] + eps * d) - loss_at(th["xyz"] - eps * d)) / (2 * eps)
    print(f"[descent] directional slope {slope:.6e} vs -|g| {-np.linalg.norm(g):.6e}")
    assert slope < 0
    assert abs(slope + np.linalg.norm(g)) <= 2e-2 * np.linalg.norm(g)


def test_adjoint_recompute_equals_stored():
    """Recomputing the probe positions in the backward gives the same gradient as keeping them."""
    w, prob, th = setup_case(5)
    out = []
    for rec in (False, True):
        prob.recompute = rec
        df, t = gpu_df(prob, th, requires_grad=True)
        (0.5 * (df.double() ** 2).sum()).backward()
        out.append(np.concatenate([t[k].grad.cpu().numpy().ravel() for k in ("xyz", "q", "R0", "E0")]))
    scale = np.abs(out[0]).max()
    assert np.abs(out[0] - out[1]).max() <= 1e-4 * scale


def test_grid_adjoint_without_parameter_columns_equals_full():
    """ppm_grid_adjoint with params = 0 gives the position and charge columns of the params = 1
    call (the same formulas; the compiler may fuse multiply-adds differently in the two forms, so
    to 1e-5 of each column's largest entry) and zero parameter columns; the autograd Function takes
    that form when neither R0 nor E0 wants a gradient, with the same xyz and q gradients."""
    w, prob, th = setup_case()
    dev = prob.pts.device
    rng = np.random.default_rng(4)
    atoms = torch.tensor(np.concatenate([th["xyz"], th["q"][:, None]], 1), dtype=torch.float32, device=dev)
    cpar = diff.lj_cpar(torch.tensor(th["R0"], device=dev), torch.tensor(th["E0"], device=dev))
    sp = prob.spec
    grid = D.grid_project(atoms, cpar, 0, sp, prob.Qs, prob.QZs, prob.npbc, prob.lvA, prob.lvB, cap=True)
    cot = torch.as_tensor(rng.standard_normal(tuple(grid.shape)).astype(np.float32), device=dev)
    cot[: sp.n[2] // 3] = 0.0                       # some octets without a cotangent
    full = D.grid_adjoint(atoms, cpar, 0, sp, prob.Qs, prob.QZs, grid, cot, prob.npbc, prob.lvA, prob.lvB, cap=True)
    part = D.grid_adjoint(atoms, cpar, 0, sp, prob.Qs, prob.QZs, grid, cot, prob.npbc, prob.lvA, prob.lvB, cap=True,
                          params=False)
    scale = full[:, :4].abs().amax(dim=0)
    rel = float(((full[:, :4] - part[:, :4]).abs() / scale).max())
    print(f"[grid adjoint params=0] max |d| / column scale {rel:.2e}", flush=True)
    assert rel <= 1e-5
    assert (part[:, 4:] == 0).all() and (full[:, 4:] != 0).any()
    grads = []
    for want in (True, False):
        t = {k: torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=(want or k in ("xyz", "q")))
             for k, v in th.items()}
        df = diff.df_image(t["xyz"], t["q"], t["R0"], t["E0"], prob)
        (df.double() ** 2).sum().backward()
        grads.append((t["xyz"].grad, t["q"].grad))
    for a, b in zip(*grads):
        assert float((a - b).abs().max() / b.abs().max()) <= 1e-5
