"""Gridless direct Morse path on the GPU (run_scan_morse_direct / vjp_scan_morse_direct).

Mirrors the reference's tests of the same interface (tests/SPM/test_afm_morse.py:981-1881):
input contract, shapes, padded scans, equilibrium residual at every returned PP, results
independent of the workgroup size, the VJP against the CPU64 oracle channel by channel,
re-relaxed finite differences, end-to-end df-loss finite differences per channel, and the
fail-loud negative cases.  Bars: residual < 1e-4 eV/A (the reference's); GPU forward vs the
FP64 oracle at tight convergence PP <= 1e-4 A and Fz <= 1e-4 eV/A; GPU VJP vs the FP64
analytic oracle at the same probe positions <= 1e-6 relative per channel (both evaluate the
adjoint in FP64) and vs the finite-difference construction <= 2e-3 (the reference's bar);
finite differences of the loss <= 5e-2 (the reference's bar; the measured error is printed).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O                               # noqa: E402
from spammm_amd import afm as afm_mod                        # noqa: E402
from spammm_amd import device as D                           # noqa: E402
from spammm_amd.afm import AFMulator                         # noqa: E402
from tests.direct_cases import (fd_component_vjp, total_force, toy_system)   # noqa: E402


def toy_afm(n_atoms=2, seed=7):
    atoms, cMs, Qs, QZs = toy_system(n_atoms, seed)
    afm = AFMulator(use_morse=True)
    afm.atoms_arr = atoms.astype(np.float32)
    afm.cLJs_arr = cMs.astype(np.float32)
    afm.tipQs = Qs.astype(np.float32)
    afm.tipQZs = QZs.astype(np.float32)
    return afm, atoms, cMs


def scan_geom(atoms, nxy, clearance=5.0, dpos0_z=4.0):
    mn, mx = atoms[:, :3].min(0), atoms[:, :3].max(0)
    p0 = np.array([mn[0] + 0.05 * (mx[0] - mn[0]), mn[1] + 0.05 * (mx[1] - mn[1]), mx[2] + clearance + dpos0_z],
                  np.float32)
    da = np.array([(mx[0] - mn[0]) * 0.9 / max(nxy[0] - 1, 1), 0, 0], np.float32)
    db = np.array([0, (mx[1] - mn[1]) * 0.9 / max(nxy[1] - 1, 1), 0], np.float32)
    return p0, da, db


def run(afm, atoms, nxy=(3, 3), nz=5, dtip=-0.2, **kw):
    p0, da, db = scan_geom(atoms, nxy)
    return afm.run_scan_morse_direct(nxy=nxy, nz=nz, dtip=dtip, scan_p0=p0, scan_da=da, scan_db=db, **kw)


def test_atom_cap_129_raises():
    afm, atoms, _ = toy_afm(129, 129)
    with pytest.raises(ValueError, match="out of range"):
        run(afm, atoms, (2, 2), 3)


def test_explicit_scan_raster_required():
    afm, atoms, _ = toy_afm()
    with pytest.raises(ValueError):
        afm.run_scan_morse_direct(nxy=(2, 2), nz=3, dtip=-0.2, scan_p0=None, scan_da=None, scan_db=None)
    with pytest.raises(ValueError, match="dtip"):
        p0, da, db = scan_geom(atoms, (2, 2))
        afm.run_scan_morse_direct(nxy=(2, 2), nz=3, dtip=0.1, scan_p0=p0, scan_da=da, scan_db=db)


def test_shapes_dtypes():
    afm, atoms, _ = toy_afm()
    FEs, pts, PPs = run(afm, atoms, (4, 5), 6)
    assert FEs.shape == (4, 5, 6, 4) and FEs.dtype == np.float32
    assert PPs.shape == (4, 5, 6, 4) and PPs.dtype == np.float32
    assert pts.shape == (4, 5, 3)
    assert np.isfinite(FEs).all() and np.isfinite(PPs).all()


def test_padded_nscan_finite_and_converged():
    afm, atoms, _ = toy_afm(3, 19)
    FEs, pts, PPs = run(afm, atoms, (5, 7), 4)
    assert np.isfinite(FEs).all() and np.isfinite(PPs).all()
    assert (PPs[..., 3] > 0).all()


def test_residual_at_every_returned_pp():
    afm, atoms, cMs = toy_afm()
    nz, dtip = 5, -0.2
    FEs, pts, PPs = run(afm, atoms, (3, 3), nz, dtip)
    worst = 0.0
    for ix in range(3):
        for iy in range(3):
            tip = pts[ix, iy].astype(np.float64)
            for iz in range(nz):
                G = total_force(atoms, cMs, afm.tipQs, afm.tipQZs, PPs[ix, iy, iz, :3].astype(np.float64), tip,
                                np.array([0., 0., 1., dtip]))
                worst = max(worst, float(np.linalg.norm(G)))
                tip = tip + np.array([0., 0., dtip])
    print(f"[direct] residual worst {worst:.3e}")
    assert worst < 1e-4


def test_results_do_not_depend_on_workgroup_size():
    afm, atoms, _ = toy_afm(3, 23)
    F32, p32, P32 = run(afm, atoms, (4, 4), 6, workgroup_size=32)
    F64, p64, P64 = run(afm, atoms, (4, 4), 6, workgroup_size=64)
    np.testing.assert_array_equal(p32, p64)
    np.testing.assert_array_equal(F32, F64)
    np.testing.assert_array_equal(P32, P64)


def molecule_case(na=128, seed=5, nxy=(16, 12), nz=12):
    """A flat sp2 flake of `na` atoms (Morse parameters by the reference's mixing rules)."""
    from spammm_amd import params as P, workloads as W
    rng = np.random.default_rng(seed)
    en, pos = W.flake_geometry(na, rng)
    atoms = np.zeros((na, 4), np.float32)
    atoms[:, :3] = pos
    atoms[:, 3] = rng.uniform(-0.3, 0.3, na)
    atoms[:, 3] -= atoms[:, 3].mean()
    cMs = P.mix_morse(en, P.ELEMENTS_REFERENCE, 1.452, 0.0006808, -1.8).astype(np.float32)
    span = np.ptp(pos[:, :2], axis=0)
    p0 = np.array([pos[:, 0].min() + 0.1 * span[0], pos[:, 1].min() + 0.1 * span[1], 11.0], np.float32)
    da = np.array([0.8 * span[0] / nxy[0], 0, 0], np.float32)
    db = np.array([0, 0.8 * span[1] / nxy[1], 0], np.float32)
    return atoms, cMs, p0, da, db


def test_large_molecule_vs_fp64_oracle_at_tight_convergence():
    """128 atoms (the limit), 16x12x12 scan: GPU FP32 (|F| < 5e-7) vs the FP64 oracle (|F| < 1e-9)."""
    atoms, cMs, p0, da, db = molecule_case()
    nxy, nz, dtip = (16, 12), 12, -0.1
    pts = D.scan_points(p0, da, db, *nxy)
    tip = D.Tip(stiffness=np.array([-0.015604, -0.015604, -0.015604, -1.2483], np.float32),
                tipC=np.array([0., 0., 1., dtip], np.float32))
    Qs, QZs = np.array([-0.05, 0, 0, 0], np.float32), np.zeros(4, np.float32)
    dev = torch.device("cuda:0")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)  # noqa: E731
    FEs, PPs = D.relax_direct_morse(t(atoms), t(cMs), t(pts), nz, tip, Qs, QZs, max_iter=8000, f2conv=2.5e-13,
                                    scan_ny=nxy[1])
    torch.cuda.synchronize()
    F, Pp = FEs.cpu().numpy(), PPs.cpu().numpy()
    Fo, Po = O.relax_direct(atoms, cMs, Qs, QZs, pts, nz, tip.tipA, tip.tipB, tip.tipC, tip.stiffness, tip.dpos0,
                            tip.relax, tip.surfFF, 20000, 1e-18, "f64")
    conv = (Pp[..., 3] > 0) & (Po[..., 3] > 0)
    print(f"[direct] converged gpu {(Pp[..., 3] > 0).mean():.4f} oracle {(Po[..., 3] > 0).mean():.4f}")
    assert conv.mean() > 0.99
    dpp = np.abs(Pp[..., :3] - Po[..., :3]).max(-1)[conv]
    dfz = np.abs(F[..., 2] - Fo[..., 2])[conv]
    print(f"[direct] PP max {dpp.max():.3e} median {np.median(dpp):.3e}; Fz max {dfz.max():.3e}")
    assert dpp.max() < 1e-4 and dfz.max() < 1e-4
    # the FE output is the pair sum at the returned PP (FP64 re-evaluation, FP32 rounding)
    fe = O.pair_fe_d(1, atoms, cMs, Pp[..., :3].reshape(-1, 3).astype(np.float64), Qs, QZs).reshape(F.shape)
    for c in range(4):
        assert np.abs(F[..., c] - fe[..., c]).max() < 1e-5 * np.abs(fe[..., c]).max() + 1e-6, c


def test_vjp_vs_cpu64_oracle_at_gpu_pps_per_channel():
    afm, atoms, cMs = toy_afm(2, 31)
    nz, dtip = 5, -0.2
    FEs, pts, PPs = run(afm, atoms, (3, 3), nz, dtip)
    target = np.zeros((3, 3, nz))
    _, _, dL = afm_mod.df_loss_seed(FEs, target, dtip)
    g, diag = afm.vjp_scan_morse_direct(PPs, pts, dtip, dL.astype(np.float32))
    assert (diag["status_codes"] == 0).all()
    pts4 = np.zeros((9, 4))
    pts4[:, :3] = pts.reshape(-1, 3)
    tip = dict(tipA=afm.tipA, tipB=afm.tipB, tipC=np.array([0., 0., 1., dtip], np.float32), stiffness=afm.stiffness,
               dpos0=afm.dpos0, surfFF=afm.surfFF)
    atoms, cMs = afm.atoms_arr.astype(np.float64), afm.cLJs_arr.astype(np.float64)   # what the GPU received
    go, _, dgo = O.direct_vjp(atoms, cMs, afm.tipQs, afm.tipQZs, pts4, nz, tip["tipA"], tip["tipB"], tip["tipC"],
                              tip["stiffness"], tip["dpos0"], tip["surfFF"], PPs.reshape(-1, nz, 4),
                              dL.astype(np.float32).reshape(-1, nz, 4))
    assert (dgo[:, 3] == 0).all()
    gf = fd_component_vjp(atoms, cMs, afm.tipQs, afm.tipQZs, pts4, nz, dtip, PPs.reshape(-1, nz, 4).astype(np.float64),
                          dL.astype(np.float32).reshape(-1, nz, 4).astype(np.float64))
    for p, lab in enumerate(("dx", "dy", "dz", "dR0", "dE0", "dQ")):
        scale = np.abs(go[:, p]).max() + 1e-30
        rel = np.abs(g[:, p] - go[:, p]).max() / scale
        relf = np.abs(g[:, p] - gf[:, p]).max() / (np.abs(gf[:, p]).max() + 1e-30)
        print(f"[direct vjp] {lab}: vs analytic oracle {rel:.2e}, vs FD construction {relf:.2e}")
        assert rel < 1e-6
        assert relf < 2e-3


def loss_at(afm, atoms, cMs, v, h, target, nxy, nz, dtip):
    ap, cp = atoms.astype(np.float64).copy(), cMs.astype(np.float64).copy()
    ap[:, :3] += h * v[:, :3]
    cp[:, 0] += h * v[:, 3]
    cp[:, 1] += h * v[:, 4]
    ap[:, 3] += h * v[:, 5]
    afm.atoms_arr, afm.cLJs_arr = ap.astype(np.float32), cp.astype(np.float32)
    FE, _, _ = run(afm, atoms, nxy, nz, dtip)
    return afm_mod.df_loss_seed(FE, target, dtip)[0]


def test_vjp_re_relax_directional_fd():
    afm, atoms, cMs = toy_afm(2, 31)
    afm.f2conv = 1e-12                     # tight forward so that small h resolves
    afm.max_iter = 4000
    nxy, nz, dtip = (3, 3), 5, -0.2
    FEs, pts, PPs = run(afm, atoms, nxy, nz, dtip)
    target = np.zeros((*nxy, nz))
    _, _, dL = afm_mod.df_loss_seed(FEs, target, dtip)
    g, _ = afm.vjp_scan_morse_direct(PPs, pts, dtip, dL.astype(np.float32))
    v = np.random.default_rng(55).standard_normal((2, 6)) * 0.01
    v[:, 5] = 0
    lin = float((g * v).sum())
    rels = []
    for h in (1e-1, 3e-2, 1e-2):
        fd = (loss_at(afm, atoms, cMs, v, h, target, nxy, nz, dtip) -
              loss_at(afm, atoms, cMs, v, -h, target, nxy, nz, dtip)) / (2 * h)
        rels.append(abs(fd - lin) / (abs(fd) + 1e-30))
        print(f"[direct re-relax] h={h:.0e} fd={fd:.6e} vjp={lin:.6e} rel={rels[-1]:.2e}")
    assert min(rels) < 5e-2


def test_df_loss_end_to_end_per_channel():
    afm, atoms, cMs = toy_afm(2, 31)
    afm.f2conv = 1e-12
    afm.max_iter = 4000
    nxy, nz, dtip = (3, 3), 5, -0.2
    FEs, pts, PPs = run(afm, atoms, nxy, nz, dtip)
    Fz = FEs[..., 2].astype(np.float64)
    target = np.zeros_like(Fz)
    target[:, :, 1:-1] = -(Fz[:, :, 2:] - Fz[:, :, :-2]) / (2 * abs(dtip))
    target += np.random.default_rng(33).standard_normal(target.shape) * 0.01
    _, _, dL = afm_mod.df_loss_seed(FEs, target, dtip)
    g, _ = afm.vjp_scan_morse_direct(PPs, pts, dtip, dL.astype(np.float32))
    h = {0: 1e-2, 1: 1e-2, 2: 1e-2, 3: 1e-2, 4: 1e-4, 5: 1e-3}
    worst = 0.0
    for i in range(2):
        for p in range(6):
            v = np.zeros((2, 6))
            v[i, p] = 1.0
            fd = (loss_at(afm, atoms, cMs, v, h[p], target, nxy, nz, dtip) -
                  loss_at(afm, atoms, cMs, v, -h[p], target, nxy, nz, dtip)) / (2 * h[p])
            rel = abs(fd - g[i, p]) / (abs(fd) + 1e-30)
            print(f"[direct e2e] atom {i} ch {p}: vjp {g[i, p]:.5e} fd {fd:.5e} rel {rel:.2e}")
            if abs(fd) > 1e-6:
                worst = max(worst, rel)
    assert worst < 5e-2


def test_vjp_nonconverged_and_shape_mismatch_raise():
    afm, atoms, _ = toy_afm()
    FEs, pts, PPs = run(afm, atoms, (2, 2), 3)
    bad = PPs.copy()
    bad[0, 0, 0, 3] = -5.0
    with pytest.raises(ValueError):
        afm.vjp_scan_morse_direct(bad, pts, -0.2, np.zeros_like(FEs))
    with pytest.raises(ValueError):
        afm.vjp_scan_morse_direct(PPs, pts, -0.2, np.zeros((3, 3, 3, 4), np.float32))


def test_vjp_stale_forward_state_raises():
    afm, atoms, _ = toy_afm()
    FEs, pts, PPs = run(afm, atoms, (2, 2), 3)
    afm.atoms_arr = afm.atoms_arr.copy()
    afm.atoms_arr[0, 0] += 0.01
    with pytest.raises(RuntimeError, match="stale"):
        afm.vjp_scan_morse_direct(PPs, pts, -0.2, np.zeros_like(FEs))


def test_direct_autograd_function_matches_facade_vjp():
    """diff.direct_df_image backward equals vjp_scan_morse_direct fed the transposed df operator."""
    from spammm_amd import df as DF, diff
    afm, atoms, cMs = toy_afm(3, 19)
    afm.f2conv, afm.max_iter = 1e-12, 4000
    nxy, nz, dtip = (4, 3), 8, -0.1
    p0, da, db = scan_geom(atoms, nxy)
    FEs, pts, PPs = afm.run_scan_morse_direct(nxy=nxy, nz=nz, dtip=dtip, scan_p0=p0, scan_da=da, scan_db=db)
    dev = torch.device("cuda:0")
    prob = diff.DirectProblem(pts=torch.as_tensor(D.scan_points(p0, da, db, *nxy), device=dev), nxy=nxy, nz=nz,
                              tip=afm._tip(dtip), Qs=afm.tipQs, QZs=afm.tipQZs, max_iter=4000, f2conv=1e-12)
    a32 = afm.atoms_arr.astype(np.float64)
    xyz = torch.tensor(a32[:, :3], device=dev, requires_grad=True)
    q = torch.tensor(a32[:, 3], device=dev, requires_grad=True)
    R0 = torch.tensor(afm.cLJs_arr[:, 0].astype(np.float64), device=dev, requires_grad=True)
    E0 = torch.tensor(afm.cLJs_arr[:, 1].astype(np.float64), device=dev, requires_grad=True)
    dfi = diff.direct_df_image(xyz, q, R0, E0, prob)
    v = torch.randn(dfi.shape, device=dev, generator=torch.Generator(device=dev).manual_seed(3))
    (dfi * v).sum().backward()
    assert prob.last_status["gated"] == 0
    op = DF.z_operator("giessibl", nz, dtip, dev, 1.0, 30000.0, 1800.0, transpose=True)
    U = np.zeros((*nxy, nz, 4), np.float32)
    U[..., 2] = op.apply(v.float().reshape(-1, nz).contiguous(), (nz, 1, 0), nxy[0] * nxy[1], 1)[..., 0] \
        .cpu().numpy().reshape(*nxy, nz)
    g, _ = afm.vjp_scan_morse_direct(PPs, pts, dtip, U)
    np.testing.assert_allclose(xyz.grad.cpu().numpy(), g[:, :3], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(R0.grad.cpu().numpy(), g[:, 3], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(E0.grad.cpu().numpy(), g[:, 4], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(q.grad.cpu().numpy(), g[:, 5], rtol=1e-6, atol=1e-12)
