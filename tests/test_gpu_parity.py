"""GPU parity of the HIP path (libppm_hip.so through its C ABI) against the CPU oracle.

Bars:
  * grid projection (FP32 on the GPU vs FP64 oracle): per voxel
    |d(Fx,Fy,Fz,E)| <= 1e-5 * |(Fx,Fy,Fz,E)| + 1e-6 (terms of opposite sign cancel,
    so the bound is relative to the voxel's vector magnitude);
  * relaxation / sampling (FP32 on the GPU vs the FP32 op-order-faithful oracle on the
    same grid): bit-exact (probe positions, iteration counts, FEs);
  * relaxation vs the FP64 oracle at tightened convergence (|F| < 1e-6 eV/A):
    probe positions <= 1e-4 A and Fz <= 1e-4 eV/A;
  * df operators (FP32 GPU vs FP64 NumPy/SciPy oracle): <= 1e-5 relative to max|df|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O          # noqa: E402
from oracle import giessibl_ref as GQ   # noqa: E402
from tests.golden_vectors import df_vectors   # noqa: E402
from spammm_amd import device as D      # noqa: E402
from spammm_amd import workloads as W   # noqa: E402


def dev():
    return D.require_gpu()


def spec_of(w):
    return D.axis_aligned_spec(w.p0, w.L, w.n)


def grid_gpu(w, model=0, cpar=None, Qs=None, QZs=None, npbc=None, cap=True, acc64=False, sp=None):
    d = dev()
    atoms = torch.as_tensor(w.atoms, device=d)
    cp = torch.as_tensor(w.cLJ if cpar is None else cpar, device=d)
    sp = spec_of(w) if sp is None else sp
    G = D.grid_project(atoms, cp, model, sp, w.Qs if Qs is None else Qs, w.QZs if QZs is None else QZs,
                       w.npbc if npbc is None else npbc, w.lvA, w.lvB, cap=cap, acc64=acc64)
    torch.cuda.synchronize()
    return G, sp


def check_grid(Gg, Go):
    Go = Go.astype(np.float64)
    mag = np.linalg.norm(Go, axis=-1, keepdims=True)
    err = np.abs(Gg.astype(np.float64) - Go)
    worst = float((err / (mag + 1e-30)).max())
    assert (err <= 1e-5 * mag + 1e-6).all(), f"grid mismatch, worst relative {worst:.3e}"
    return worst


def morse_params(w):
    c = np.zeros((w.atoms.shape[0], 4), np.float32)
    R0 = np.where(np.array(w.enames) == "H", 1.487, 1.908) + 1.66
    c[:, 0] = R0
    c[:, 1] = np.sqrt(0.0091 * np.where(np.array(w.enames) == "H", 0.00068, 0.0037))
    c[:, 2] = -1.8
    return c


@pytest.mark.parametrize("acc64", [False, True])
@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("case", ["pbc_q0", "nopbc_offsets", "pbc_offsets_nocap", "tilted_c"])
def test_grid_project_small_vs_oracle(model, case, acc64):
    """Both voxel-sum forms (FP32, and FP32 terms summed in FP64) against the FP64 oracle."""
    w = W.small()
    cpar = w.cLJ if model == 0 else morse_params(w)
    Qs, QZs, npbc, cap = w.Qs, w.QZs, w.npbc, True
    sp = spec_of(w)
    if case == "nopbc_offsets":
        Qs, QZs, npbc = np.array([0., -0.1, 0.1, 0.], np.float32), np.array([0., 1.8, 3.6, 0.], np.float32), (0, 0)
    elif case == "pbc_offsets_nocap":
        Qs, QZs, cap = np.array([-0.05, 0.2, 0., 0.], np.float32), np.array([0., 1.2, 0., 0.], np.float32), False
    elif case == "tilted_c":     # dC with an x,y component: the kernels' per-voxel x,y form
        n = tuple(int(v) for v in w.n)
        dC = np.array([0.02, -0.015, float(w.L[2]) / n[2]], np.float32)
        L = np.array([sp.dA * n[0], sp.dB * n[1], dC * n[2]], np.float64)
        sp = D.GridSpec(n, sp.p0, sp.dA, sp.dB, dC, *D.dinv_lvec(np.linalg.inv(L).T, n))
    print(f"[grid] model={model} case={case} acc64={acc64} n={w.n}", flush=True)
    G, sp = grid_gpu(w, model, cpar, Qs, QZs, npbc, cap, acc64=acc64, sp=sp)
    Go = O.grid_project(model, w.atoms, cpar, w.n, sp.p0, sp.dA, sp.dB, sp.dC, Qs, QZs, npbc, w.lvA, w.lvB,
                        cap=cap)
    worst = check_grid(G.cpu().numpy(), Go)
    print(f"[grid] worst relative error {worst:.3e}")


def test_grid_project_full_naphthalene_sampled():
    w = W.naphthalene()
    print(f"[grid] naphthalene full grid {w.n}", flush=True)
    G, sp = grid_gpu(w)
    Gh = G.cpu().numpy().reshape(-1, 4)
    assert np.isfinite(Gh).all()
    idx = np.random.default_rng(5).choice(Gh.shape[0], 3000, replace=False)
    Go = O.grid_voxels(0, w.atoms, w.cLJ, w.n, sp.p0, sp.dA, sp.dB, sp.dC, w.Qs, w.QZs, idx, w.npbc, w.lvA, w.lvB)
    worst = check_grid(Gh[idx], Go)
    print(f"[grid] worst relative error {worst:.3e}")


def tip_of(w, dtip=None):
    return D.Tip(stiffness=w.stiffness, dpos0=w.dpos0,
                 tipC=np.array([0., 0., 1., w.dtip if dtip is None else dtip], np.float32))


def relax_both(w, G, sp, precision="f32", max_iter=None, f2conv=None, check_before=False, scan_ny=0, pts=None,
               oracle_max_iter=None, oracle_f2conv=None):
    d = dev()
    if pts is None:
        pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    tip = tip_of(w)
    mi = w.max_iter if max_iter is None else max_iter
    f2 = w.f2conv if f2conv is None else f2conv
    FEs, PPs = D.relax_tilted(G, sp, torch.as_tensor(pts, device=d), w.nz, tip, mi, f2, check_before, scan_ny)
    torch.cuda.synchronize()
    Gh = G.cpu().numpy()
    FEo, PPo = O.relax_tilted(Gh, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, tip.tipA, tip.tipB, tip.tipC,
                              tip.stiffness, tip.dpos0, tip.relax, tip.surfFF,
                              mi if oracle_max_iter is None else oracle_max_iter,
                              f2 if oracle_f2conv is None else oracle_f2conv, check_before, precision)
    return FEs.cpu().numpy(), PPs.cpu().numpy(), FEo, PPo


def tip_positions_f32(pts, nz, dtip):
    """Tip anchor per (point, z) accumulated in float32 as the kernels do (tipPos += dTip)."""
    tip = np.zeros((pts.shape[0], nz, 3), np.float32)
    cur = pts[:, :3].astype(np.float32).copy()
    step = np.array([0, 0, dtip], np.float32)
    for iz in range(nz):
        tip[:, iz] = cur
        cur = cur + step
    return tip


@pytest.mark.parametrize("check_before", [False, True])
@pytest.mark.parametrize("tiled", [False, True])
def test_relax_tilted_bitexact_vs_fp32_oracle(check_before, tiled):
    w = W.small()
    G, sp = grid_gpu(w)
    print(f"[relax] small check_before={check_before} tiled={tiled}", flush=True)
    FEs, PPs, FEo, PPo = relax_both(w, G, sp, "f32", check_before=check_before, scan_ny=w.nxy[1] if tiled else 0)
    np.testing.assert_array_equal(PPs, PPo)
    np.testing.assert_array_equal(FEs, FEo)
    assert (PPs[..., 3] != 0).all()


def test_relax_tilted_tilted_frame_bitexact():
    """A rotated tip frame and a surface force exercise rotMat/rotMatT and surfFF."""
    w = W.small()
    G, sp = grid_gpu(w)
    d = dev()
    th = 0.2
    A = np.array([np.cos(th), 0., -np.sin(th), 0.], np.float32)
    B = np.array([0., 1., 0., 0.], np.float32)
    C = np.array([np.sin(th), 0., np.cos(th), -0.1], np.float32)
    tip = D.Tip(A, B, C, w.stiffness, w.dpos0, np.array([0.5, 0.1, 0, 0], np.float32),
                np.array([0.01, 0, 0, 0], np.float32))
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    FEs, PPs = D.relax_tilted(G, sp, torch.as_tensor(pts, device=d), w.nz, tip, 100, 1e-8)
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, A, B, C,
                              tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, 100, 1e-8, False, "f32")
    np.testing.assert_array_equal(PPs.cpu().numpy(), PPo)
    np.testing.assert_array_equal(FEs.cpu().numpy(), FEo)


def test_relax_vs_fp64_oracle_tight_convergence():
    """FP32 GPU at |F| < 1e-6 eV/A vs the FP64 oracle converged to |F| < 1e-10 eV/A (so that the bound
    measures the FP32 path): probe positions within 1e-4 A, Fz within 1e-4 eV/A."""
    w = W.small()
    G, sp = grid_gpu(w)
    FEs, PPs, FEo, PPo = relax_both(w, G, sp, "f64", max_iter=4000, f2conv=1e-12, check_before=True,
                                    oracle_max_iter=20000, oracle_f2conv=1e-20)
    conv = (PPs[..., 3] > 0) & (PPo[..., 3] > 0)
    assert conv.mean() > 0.98, f"only {conv.mean():.3f} of states converged at |F|<1e-6"
    dpos = np.abs(PPs[..., :3] - PPo[..., :3]).max(-1)[conv]
    dFz = np.abs(FEs[..., 2] - FEo[..., 2])[conv]
    print(f"[relax f64] PP max {dpos.max():.2e} median {np.median(dpos):.2e}; Fz max {dFz.max():.2e}")
    assert dpos.max() <= 1e-4
    assert dFz.max() <= 1e-4


def test_relax_reference_settings_error_vs_fp64_reported():
    """Reference settings (|F|<1e-4, check after update): FP32 GPU vs FP64 oracle error is reported;
    two correct implementations may stop up to |F|/k apart, so only the median is bounded."""
    w = W.small()
    G, sp = grid_gpu(w)
    FEs, PPs, FEo, PPo = relax_both(w, G, sp, "f64")
    dpos = np.abs(PPs[..., :3] - PPo[..., :3]).max(-1)
    dFz = np.abs(FEs[..., 2] - FEo[..., 2])
    print(f"[relax ref] PP max {dpos.max():.2e} median {np.median(dpos):.2e}; "
          f"Fz max {dFz.max():.2e} median {np.median(dFz):.2e}")
    assert np.median(dpos) < 1e-2


def test_relax_residual_at_converged_states():
    """check_before mode: every converged PP satisfies |F_total| < 1e-4 eV/A (tests/SPM/test_afm_morse.py:907-941)."""
    w = W.small()
    G, sp = grid_gpu(w)
    FEs, PPs, _, _ = relax_both(w, G, sp, "f32", check_before=True)
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    Gh = G.cpu().numpy()
    conv = PPs[..., 3] > 0
    pp = PPs[..., :3].reshape(-1, 3)
    fe = O.interp(Gh, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pp)
    tip = tip_positions_f32(pts, w.nz, w.dtip).astype(np.float64)
    d = pp - tip.reshape(-1, 3)
    r = np.linalg.norm(d, axis=1)
    k = w.stiffness.astype(np.float64)
    d0 = w.dpos0.astype(np.float64)
    ftip = (d - d0[:3]) * k[:3] + d * (k[3] * (r - d0[3]) / r)[:, None]
    res = np.linalg.norm(fe[:, :3] + ftip, axis=1).reshape(conv.shape)
    assert conv.mean() > 0.95
    # FP32 evaluation of |F|^2 < 1e-8 in the kernel vs FP64 re-evaluation: allow 1e-6 of rounding
    assert res[conv].max() < 1e-4 + 1e-6, f"residual {res[conv].max():.3e}"


def test_newton_polish_reaches_fp64_equilibrium():
    """After FIRE (|F|<1e-6) the FP64 Newton polish puts probe positions on the FP64 equilibrium of the
    same trilinear model to well below the FIRE tolerance (<= 1e-5 A)."""
    w = W.small()
    G, sp = grid_gpu(w)
    d = dev()
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    tip = tip_of(w)
    FEs, PPs = D.relax_tilted(G, sp, torch.as_tensor(pts, device=d), w.nz, tip, 4000, 1e-12, True)
    pp_fire = PPs.cpu().numpy().copy()
    D.relax_polish(G, sp, torch.as_tensor(pts, device=d), w.nz, tip, PPs, FEs, 3)
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, tip.tipA, tip.tipB,
                              tip.tipC, tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, 20000, 1e-24, True, "f64")
    P = PPs.cpu().numpy()
    conv = (P[..., 3] > 0) & (PPo[..., 3] > 0)
    e_fire = np.abs(pp_fire[..., :3] - PPo[..., :3]).max(-1)[conv]
    e_pol = np.abs(P[..., :3] - PPo[..., :3]).max(-1)[conv]
    dfz = np.abs(FEs.cpu().numpy()[..., 2] - FEo[..., 2])[conv]
    print(f"[polish] PP err FIRE max {e_fire.max():.2e} -> polished max {e_pol.max():.2e} "
          f"median {np.median(e_pol):.2e}; Fz err max {dfz.max():.2e}")
    assert e_pol.max() <= 1e-5
    assert dfz.max() <= 1e-5


def test_relax_strokes_and_raw_fe_bitexact():
    w = W.small()
    G, sp = grid_gpu(w)
    d = dev()
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    dTip = np.array([0., 0., -0.1, 0.], np.float32)
    st, d0, rl = w.stiffness, np.array([0., 0., -3.0, 3.0], np.float32), np.array([0.5, 0.1, 0, 0], np.float32)
    FEs, disps, PPs = D.relax_strokes(G, sp, torch.as_tensor(pts, device=d), w.nz, dTip, st, d0, rl, 100, 1e-8)
    FEo, dio, PPo = O.relax_strokes(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, dTip, st, d0,
                                    rl, 100, 1e-8, "f32")
    np.testing.assert_array_equal(FEs.cpu().numpy(), FEo)
    np.testing.assert_array_equal(disps.cpu().numpy(), dio)
    np.testing.assert_array_equal(PPs.cpu().numpy(), PPo)
    tip = tip_of(w)
    raw = D.fe_in_strokes_tilted(G, sp, torch.as_tensor(pts, device=d), w.nz, tip, dTip, w.dpos0)
    rawo = O.fe_in_strokes_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, tip.tipA,
                                  tip.tipB, tip.tipC, dTip, w.dpos0, "f32")
    np.testing.assert_array_equal(raw.cpu().numpy(), rawo)


def test_full_naphthalene_scan_subsample_bitexact():
    """Full-size 200^3 grid and 100x100x40 scan: all states finite; 256 pixels bit-exact vs the FP32 oracle."""
    w = W.naphthalene()
    G, sp = grid_gpu(w)
    d = dev()
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    FEs, PPs = D.relax_tilted(G, sp, torch.as_tensor(pts, device=d), w.nz, tip_of(w), w.max_iter, w.f2conv,
                              scan_ny=w.nxy[1])
    FEs, PPs = FEs.cpu().numpy(), PPs.cpu().numpy()
    assert np.isfinite(FEs).all() and np.isfinite(PPs).all()
    conv = (PPs[..., 3] > 0).mean()
    print(f"[full] converged fraction {conv:.4f}, mean iters {np.abs(PPs[..., 3]).mean():.2f}")
    sel = np.random.default_rng(3).choice(w.n_scan, 256, replace=False)
    tip = tip_of(w)
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts[sel], w.nz, tip.tipA,
                              tip.tipB, tip.tipC, tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, w.max_iter,
                              w.f2conv, False, "f32")
    np.testing.assert_array_equal(PPs[sel], PPo)
    np.testing.assert_array_equal(FEs[sel], FEo)


@pytest.mark.parametrize("nxy", [(12, 10), (200, 200), (230, 230), (300, 290), (480, 480)])
def test_relax_with_fused_df_equals_separate_operator(nxy):
    """ppm_relax_strokes_tilted_df: FEs and PPs equal ppm_relax_strokes_tilted's and df equals the
    separate operator launch on its FEs, bit for bit, for each lane-group width (these scans run 4,
    4, 2, 1 and 1 lanes per point: up to 2.5 waves per SIMD); FEs also bit-exact against the FP32
    oracle on a subsample.  One lane per point stages its output runs in LDS while the launch's
    waves stay resident with it: both runs of 300x290 points (1406 waves, 6 per compute unit); of
    480x480 (3600 waves, 15 per unit) the one without df, while the df run (12 Fz floats per lane
    in LDS as well: 14 waves per unit) stages in registers."""
    from spammm_amd import df as DF
    w = W.small()
    G, sp = grid_gpu(w)
    d = dev()
    da = np.array([(w.nxy[0] - 1) * w.scan_da[0] / max(nxy[0] - 1, 1), 0, 0], np.float32)
    db = np.array([0, (w.nxy[1] - 1) * w.scan_db[1] / max(nxy[1] - 1, 1), 0], np.float32)
    pts_h = D.scan_points(w.scan_p0, da, db, *nxy)
    pts = torch.as_tensor(pts_h, device=d)
    tip = tip_of(w)
    ns = nxy[0] * nxy[1]
    op = DF.z_operator("giessibl", w.nz, w.dtip, d, w.amp, w.f0, w.k_cant)
    FE1, PP1 = D.relax_tilted(G, sp, pts, w.nz, tip, 100, 1e-8, scan_ny=nxy[1])
    df1 = op.apply(FE1, (w.nz * 4, 4, 0), ns, 1, offset=2)
    df2 = torch.empty((ns, op.n_out), dtype=torch.float32, device=d)
    FE2, PP2 = D.relax_tilted(G, sp, pts, w.nz, tip, 100, 1e-8, scan_ny=nxy[1], df_op=op, df=df2)
    torch.cuda.synchronize()
    assert torch.equal(FE1, FE2) and torch.equal(PP1, PP2)
    assert torch.equal(df1.reshape(ns, op.n_out), df2)
    sel = np.random.default_rng(1).choice(ns, min(ns, 64), replace=False)
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts_h[sel], w.nz, tip.tipA,
                              tip.tipB, tip.tipC, tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, 100, 1e-8, False,
                              "f32")
    np.testing.assert_array_equal(FE2.cpu().numpy()[sel], FEo)
    np.testing.assert_array_equal(PP2.cpu().numpy()[sel], PPo)


# ── df operators ─────────────────────────────────────────────────────────────

def _cmp(a, b, rel=1e-5):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    assert np.abs(a - b).max() <= rel * scale, f"max err {np.abs(a - b).max():.3e} vs scale {scale:.3e}"


def test_df_gradient_and_amp9_vs_oracle():
    from spammm_amd import afm
    V = df_vectors()
    Fz = V["grad__Fz"]
    _cmp(afm.compute_df(Fz, 0.1), V["grad__df"])
    _cmp(afm.compute_df_amp(Fz, 0.1, 1.0), V["amp1__df"])
    _cmp(afm.compute_df_amp(Fz, -0.1, 0.37), V["amp2__df"])


def test_df_giessibl_vs_quadrature():
    from spammm_amd import afm
    rng = np.random.default_rng(1)
    z = 6.0 - 0.1 * np.arange(40)
    Fz = (0.3 * np.exp(-(z[None, None, :] - 3.0)) + 0.05 * rng.normal(size=(3, 2, 40))).astype(np.float32)
    df = afm.df_giessibl(Fz, -0.1, 1.0, 30000.0, 1800.0)
    ref = GQ.giessibl_df_piecewise(Fz.astype(np.float64), z, 1.0, 30000.0, 1800.0)
    _cmp(df, ref, 2e-5)


def test_df_loss_seed_vs_oracle():
    from spammm_amd import afm
    V = df_vectors()
    L, df, dL = afm.df_loss_seed(V["loss__FEs"], V["loss__target"], -0.1, V["loss__w"])
    assert abs(L - float(V["loss__L"])) <= 1e-5 * abs(float(V["loss__L"]))
    _cmp(df, V["loss__df"])
    _cmp(dL, V["loss__dL"])
    L, df, dL = afm.df_loss_seed(V["loss__FEs"], V["loss__target"], -0.1)
    assert abs(L - float(V["lossd__L"])) <= 1e-5 * abs(float(V["lossd__L"]))
    _cmp(df, V["lossd__df"])
    _cmp(dL, V["lossd__dL"])


def test_df_directional_and_dissipation_vs_oracle():
    from spammm_amd import afm
    V = df_vectors()
    FEs, n, sp = V["dir__FEs"], V["dir__n"], tuple(V["dir__spacing"])
    _cmp(afm.compute_df_dir(FEs, sp, n), V["dir__df_dir"])
    _cmp(afm.compute_df_amp_dir(FEs, sp, n, 0.5), V["dir__df_amp_dir"])
    _cmp(afm.compute_df_amp_dir(FEs, sp, (0., 0., 1.), 0.5), V["dir__df_amp_vertical"])
    _cmp(afm.compute_dissipation(FEs, V["dir__bwd"], V["dir__h_scan"], V["dir__h_df"], 0.4), V["dir__E_diss"])


def test_errors_are_raised_not_swallowed():
    d = dev()
    w = W.small()
    G, sp = grid_gpu(w)
    pts = torch.zeros((4, 4), device=d)
    with pytest.raises(ValueError):
        D.relax_tilted(G, sp, pts, 3, tip_of(w), max_iter=0)
    with pytest.raises(ValueError):
        D.relax_tilted(G, sp, pts.double(), 3, tip_of(w))
    bad = D.axis_aligned_spec(w.p0, w.L, (4, 4, 4))
    bad.n = (0, 4, 4)
    with pytest.raises(ValueError):
        D.grid_project(torch.as_tensor(w.atoms, device=d), torch.as_tensor(w.cLJ, device=d), 0, bad, w.Qs, w.QZs)


def _abs_term_sums(w, P):
    """Per point, the sum over atoms and images of |F| and |E| of the individual LJ and Coulomb
    terms (FP64) — the scale of an FP32 sum's rounding where terms cancel."""
    return O.abs_term_sums(w.atoms, w.cLJ, w.Qs[0], w.npbc, w.lvA, w.lvB, P)


def test_full_flake_grid_sampled_and_scan_subsample_bitexact():
    """The 200-atom flake at its full size (300x300x240 grid, 256x256x60 scan): 2000 sampled voxels
    against the FP64 oracle within 1e-5 |(F,E)| + 4e-6 sum|terms| + 1e-6 (the flake's ~1800 Coulomb
    terms per voxel cancel to ~1e-2 of their magnitudes above the plane, where an FP32 sum's
    rounding follows the terms, not the net value); 128 pixels of the relaxed scan bit-exact
    against the FP32 oracle on the same grid."""
    w = W.flake()
    G, sp = grid_gpu(w)
    Gh = G.cpu().numpy().reshape(-1, 4)
    assert np.isfinite(Gh).all()
    idx = np.random.default_rng(11).choice(Gh.shape[0], 2000, replace=False)
    Go = O.grid_voxels(0, w.atoms, w.cLJ, w.n, sp.p0, sp.dA, sp.dB, sp.dC, w.Qs, w.QZs, idx, w.npbc, w.lvA,
                       w.lvB).astype(np.float64)
    nx, ny, nz = w.n
    ic, ib, ia = np.unravel_index(idx, (nz, ny, nx))
    P = (np.asarray(sp.p0, np.float64)[None, :] + np.outer(ia, sp.dA) + np.outer(ib, sp.dB) + np.outer(ic, sp.dC))
    S = _abs_term_sums(w, P)
    scale = np.minimum(1.0, 100.0 / np.maximum(np.abs(Go[:, 3]), 1e-300))     # capped voxels
    bound = 1e-5 * np.linalg.norm(Go, axis=1) + 4e-6 * S * scale + 1e-6
    err = np.abs(Gh[idx].astype(np.float64) - Go).max(1)
    print(f"[full] flake grid: worst |d|/bound {(err / bound).max():.3e}", flush=True)
    assert (err <= bound).all()
    d = dev()
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    FEs, PPs = D.relax_tilted(G, sp, torch.as_tensor(pts, device=d), w.nz, tip_of(w), w.max_iter, w.f2conv,
                              scan_ny=w.nxy[1])
    FEs, PPs = FEs.cpu().numpy(), PPs.cpu().numpy()
    assert np.isfinite(FEs).all() and np.isfinite(PPs).all()
    sel = np.random.default_rng(4).choice(w.n_scan, 128, replace=False)
    tip = tip_of(w)
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts[sel], w.nz, tip.tipA,
                              tip.tipB, tip.tipC, tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, w.max_iter,
                              w.f2conv, False, "f32")
    np.testing.assert_array_equal(PPs[sel], PPo)
    np.testing.assert_array_equal(FEs[sel], FEo)


def test_full_flake_grid_acc64_vs_fp64():
    """ppm_grid_project_acc64 on the full flake grid: 2000 sampled voxels against the FP64 oracle
    within a quarter of the FP32 sum's bound (1e-5 |(F,E)| + 1e-6 sum|terms| + 1e-6), and, on the
    voxels whose net field cancels below 3 % of sum|terms| (above the molecule), a median error at
    most half the FP32 sum's: what FP64 sums remove is the FP32 running sum's rounding."""
    w = W.flake()
    G32, sp = grid_gpu(w)
    G64, _ = grid_gpu(w, acc64=True)
    G32 = G32.cpu().numpy().reshape(-1, 4)
    G64 = G64.cpu().numpy().reshape(-1, 4)
    assert np.isfinite(G64).all()
    idx = np.random.default_rng(12).choice(G64.shape[0], 2000, replace=False)
    Go = O.grid_voxels(0, w.atoms, w.cLJ, w.n, sp.p0, sp.dA, sp.dB, sp.dC, w.Qs, w.QZs, idx, w.npbc, w.lvA,
                       w.lvB).astype(np.float64)
    nx, ny, nz = w.n
    ic, ib, ia = np.unravel_index(idx, (nz, ny, nx))
    P = (np.asarray(sp.p0, np.float64)[None, :] + np.outer(ia, sp.dA) + np.outer(ib, sp.dB) + np.outer(ic, sp.dC))
    S = _abs_term_sums(w, P)
    scale = np.minimum(1.0, 100.0 / np.maximum(np.abs(Go[:, 3]), 1e-300))
    bound = 1e-5 * np.linalg.norm(Go, axis=1) + 1e-6 * S * scale + 1e-6
    e64 = np.abs(G64[idx].astype(np.float64) - Go).max(1)
    e32 = np.abs(G32[idx].astype(np.float64) - Go).max(1)
    canc = np.linalg.norm(Go, axis=1) < 0.03 * S * scale
    m64, m32 = float(np.median(e64[canc])), float(np.median(e32[canc]))
    print(f"[acc64] worst |d|/bound {(e64 / bound).max():.3e}; cancelling voxels {int(canc.sum())}: "
          f"median |d| {m64:.3e} (FP64 sums) vs {m32:.3e} (FP32 sums)", flush=True)
    assert (e64 <= bound).all()
    assert canc.sum() >= 200 and m64 <= 0.5 * m32
