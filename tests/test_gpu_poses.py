"""Molecule placements over a substrate: grid over a base and the pose adjoint.

Bars: the grid projected over a raw substrate base equals the FP64 oracle grid of the
combined atom list per voxel within |d| <= 1e-5 ||(F,E)|| + 4e-6 sum|terms| + 1e-6 (the FP32 grid bar of
tests/test_gpu_parity.py); the pose/charge adjoint equals central finite differences of the
FP64 oracle forward (phi = v . df with fixed v) within 1e-3 relative, denominator floored at
1e-2 of the largest |FD| (the bar of tests/test_gpu_adjoint.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O            # noqa: E402
from spammm_amd import device as D        # noqa: E402
from spammm_amd import df as DF           # noqa: E402
from spammm_amd import poses as PS        # noqa: E402
from spammm_amd import workloads as W     # noqa: E402


def make_scene(sc, **kw):
    dev = D.require_gpu()
    spec = D.axis_aligned_spec(sc["p0"], sc["L"], sc["n"])
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)  # noqa: E731
    kl = sc["k_lat_Nm"] / 16.02176634
    kr = sc["k_rad_Nm"] / 16.02176634
    tip = D.Tip(stiffness=np.array([-kl, -kl, -kl, -kr], np.float32),
                dpos0=np.array([0, 0, -sc["bond"], sc["bond"]], np.float32),
                tipC=np.array([0., 0., 1., sc["dtip"]], np.float32))
    pts = D.scan_points(sc["scan_p0"], sc["scan_da"], sc["scan_db"], *sc["nxy"])
    args = dict(max_iter=3000, f2conv=1e-12)
    args.update(kw)
    scene = PS.Scene(spec=spec, sub_atoms=t(sc["sub_atoms"]), sub_cpar=t(sc["sub_cLJ"]),
                     mol_xyz0=sc["mol_atoms"][:, :3].astype(np.float64), mol_cpar=t(sc["mol_cLJ"]), pts=t(pts),
                     nxy=sc["nxy"], nz=sc["nz"], tip=tip, Qs=sc["Qs"], QZs=sc["QZs"], npbc=sc["npbc"],
                     lvA=sc["lvA"], lvB=sc["lvB"], **args)
    return scene.prepare(), pts


def combined(sc, pose, q):
    x = PS.place(sc["mol_atoms"][:, :3].astype(np.float64), sc["mol_atoms"][:, :3].astype(np.float64).mean(0),
                 np.asarray(pose, np.float64))
    mol = np.concatenate([x, np.asarray(q, np.float64)[:, None]], 1)
    atoms = np.concatenate([sc["sub_atoms"].astype(np.float64), mol])
    cpar = np.concatenate([sc["sub_cLJ"], sc["mol_cLJ"]]).astype(np.float64)
    return atoms, cpar


def test_grid_over_base_equals_combined_grid():
    sc = W.small_substrate_scene()
    scene, _ = make_scene(sc)
    pose = sc["poses"][1]
    q = sc["mol_atoms"][:, 3].astype(np.float64)
    x = PS.place(scene.mol_xyz0, scene.center, pose)
    atoms = torch.as_tensor(np.concatenate([x, q[:, None]], 1).astype(np.float32), device=scene.pts.device)
    assert scene.windowed
    G = D.grid_project(atoms, scene.mol_cpar, 0, scene.spec_run, sc["Qs"], sc["QZs"], sc["npbc"], sc["lvA"],
                       sc["lvB"], cap=True, base=scene.base).cpu().numpy().astype(np.float64)
    ca, cp = combined(sc, pose, q)
    sp = scene.spec_run
    Go = O.grid_project_d(0, ca.astype(np.float32).astype(np.float64), cp.astype(np.float32).astype(np.float64),
                          sp.n, sp.p0, sp.dA, sp.dB, sp.dC, sc["Qs"], sc["QZs"], sc["npbc"], sc["lvA"], sc["lvB"])
    # The ionic slab's terms cancel to a small net field above it, so the FP32 sum's rounding is
    # bounded by the sum of the terms' magnitudes, not by the net value: |d| <= 1e-5 |net| +
    # 4e-6 sum|terms| + 1e-6 (a missing or doubled atom term, ~1/300 of the sum, is far above it).
    S = abs_term_sums(ca.astype(np.float32).astype(np.float64), cp.astype(np.float32).astype(np.float64), sp, sc)
    mag = np.linalg.norm(Go, axis=-1, keepdims=True)
    err = np.abs(G - Go)
    bound = 1e-5 * mag + 4e-6 * S + 1e-6
    print(f"[poses] grid over base: max |d| {err.max():.3e}, worst |d|/bound {(err / bound).max():.3e}")
    assert (err <= bound).all()


def abs_term_sums(atoms, cpar, sp, sc, cap=True):
    """Per voxel, the sum over atoms and images of |F_term| and |E_term| (FP64, NumPy), capped
    voxels scaled like the grid (100/|E|, if cap) so that the bound follows the stored values."""
    nx, ny, nz = sp.n
    ia, ib, ic = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    P = (np.asarray(sp.p0, np.float64)[None, :] + np.outer(ia.ravel(), sp.dA) + np.outer(ib.ravel(), sp.dB)
         + np.outer(ic.ravel(), sp.dC))
    SF = np.zeros(len(P))
    SE = np.zeros(len(P))
    net = np.zeros(len(P))
    qs = float(sc["Qs"][0]) * 14.3996448915
    for kx in range(-sc["npbc"][0], sc["npbc"][0] + 1):
        for ky in range(-sc["npbc"][1], sc["npbc"][1] + 1):
            sh = kx * np.asarray(sc["lvA"], np.float64) + ky * np.asarray(sc["lvB"], np.float64)
            for (x, y, z, q), (c6, c12) in zip(atoms, cpar):
                d = P - (np.array([x, y, z]) + sh)
                r2 = (d * d).sum(1) + 1e-4
                ir2 = 1.0 / r2
                ir6 = ir2 ** 3
                e_lj = (c12 * ir6 - c6) * ir6
                f_lj = np.abs((12 * c12 * ir6 - 6 * c6) * ir6 * ir2) * np.sqrt(r2)
                e_c = q * qs * np.sqrt(ir2)
                SF += f_lj + np.abs(e_c) * ir2 * np.sqrt(r2)
                SE += np.abs(e_lj) + np.abs(e_c)
                net += e_lj + e_c
    scale = np.minimum(1.0, 100.0 / np.maximum(np.abs(net), 1e-300)) if cap else 1.0
    S = (SF + SE) * scale
    return S.reshape(nx, ny, nz).transpose(2, 1, 0)[..., None]


def fp64_df(sc, scene, pose, q):
    ca, cp = combined(sc, pose, q)
    sp = scene.spec
    G = O.grid_project_d(0, ca, cp, sp.n, sp.p0, sp.dA, sp.dB, sp.dC, sc["Qs"], sc["QZs"], sc["npbc"], sc["lvA"],
                         sc["lvB"])
    tip = scene.tip
    pts = scene.pts.cpu().numpy()
    FEs, PPs = O.relax_tilted_d(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, scene.nz, tip.tipA, tip.tipB, tip.tipC,
                                tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, max_iter=40000, f2conv=1e-24,
                                check_before=True)
    assert (PPs[..., 3] > 0).all(), "FP64 reference relaxation did not converge"
    M = DF.giessibl_matrix(scene.nz, scene.dz, scene.amp, scene.f0, scene.k_cant)
    return (FEs[..., 2] @ M.T).reshape(scene.nxy[0], scene.nxy[1], scene.nz)


def test_pose_adjoint_matches_fp64_finite_differences():
    sc = W.small_substrate_scene()
    scene, _ = make_scene(sc)
    dev = scene.pts.device
    pose0 = sc["poses"][2].astype(np.float64)
    q0 = sc["mol_atoms"][:, 3].astype(np.float64)
    pose = torch.tensor(pose0, device=dev, requires_grad=True)
    q = torch.tensor(q0, device=dev, requires_grad=True)
    df = PS.pose_df(pose, q, scene)
    v = np.random.default_rng(4).standard_normal(df.shape)
    (df.double() * torch.as_tensor(v, device=dev)).sum().backward()
    assert scene.last_status["gated"] == 0, scene.last_status
    gp, gq = pose.grad.cpu().numpy(), q.grad.cpu().numpy()

    def phi(p, qq):
        return float((v * fp64_df(sc, scene, p, qq)).sum())

    fd, an, names = [], [], []
    for k, h in ((0, 1e-3), (1, 1e-3), (2, 1e-3), (3, 1e-4)):
        pp, pm = pose0.copy(), pose0.copy()
        pp[k] += h
        pm[k] -= h
        fd.append((phi(pp, q0) - phi(pm, q0)) / (2 * h))
        an.append(gp[k])
        names.append(("sx", "sy", "sz", "theta")[k])
    for i in range(3):
        qp, qm = q0.copy(), q0.copy()
        qp[i] += 1e-4
        qm[i] -= 1e-4
        fd.append((phi(pose0, qp) - phi(pose0, qm)) / 2e-4)
        an.append(gq[i])
        names.append(f"q{i}")
    fd, an = np.array(fd), np.array(an)
    floor = 1e-2 * np.abs(fd).max()
    rel = np.abs(an - fd) / np.maximum(np.abs(fd), floor)
    for n, a, f, r in zip(names, an, fd, rel):
        print(f"[poses] {n}: adjoint {a:+.6e} fd {f:+.6e} rel {r:.2e}")
    assert rel.max() < 1e-3


def test_backward_after_another_placement_recomputes_the_grid():
    """The shared grid buffer is overwritten by a later placement's forward; the backward of the
    earlier placement must then rebuild its own grid and give the same gradients (to the
    rounding of the relax adjoint's atomic scatter; the other placement's grid would be far off)."""
    sc = W.small_substrate_scene()
    scene, _ = make_scene(sc)
    dev = scene.pts.device
    q0 = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev)
    v = torch.as_tensor(np.random.default_rng(6).standard_normal((*sc["nxy"], sc["nz"])), device=dev)

    def grads(interleave):
        pose = torch.tensor(sc["poses"][1], dtype=torch.float64, device=dev, requires_grad=True)
        q = q0.clone().requires_grad_(True)
        df = PS.pose_df(pose, q, scene)
        if interleave:
            with torch.no_grad():
                PS.pose_df(torch.tensor(sc["poses"][3], dtype=torch.float64, device=dev), q0, scene)
        (df.double() * v).sum().backward()
        return pose.grad.cpu().numpy(), q.grad.cpu().numpy()

    gp_a, gq_a = grads(False)
    gp_b, gq_b = grads(True)
    print(f"[poses] pose grad {gp_a} (interleaved {gp_b})")
    for a, b in ((gp_a, gp_b), (gq_a, gq_b)):
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(a).max())


def test_ewald_substrate_base_equals_lj_plus_ewald_oracle():
    """sub_coulomb = "ewald": the raw base is the slab's LJ image sum (no tip charges) plus the
    2D Ewald electrostatics of the tip charges; checked on 3000 voxels against the FP64 LJ oracle
    plus the FP64 Ewald oracle (oracle/ewald_ref.py) within the grid bar of this module plus 2e-5
    of the largest Ewald channel (the Ewald grid's FP32 bar, tests/test_gpu_ewald.py)."""
    from oracle import ewald_ref as EW
    sc = W.small_substrate_scene()
    scene, _ = make_scene(sc, sub_coulomb="ewald", ewald_n_harm=8)
    base = scene.base.cpu().numpy().astype(np.float64)
    sp = scene.spec_run
    sub = sc["sub_atoms"].astype(np.float64)
    cp = sc["sub_cLJ"].astype(np.float64)
    lj = O.grid_project_d(0, sub, cp, sp.n, sp.p0, sp.dA, sp.dB, sp.dC, np.zeros(4), np.zeros(4), sc["npbc"],
                          sc["lvA"], sc["lvB"], cap=False)
    slab = EW.Slab(sub, sc["lvA"][:2], sc["lvB"][:2], 8)
    nx, ny, nz = sp.n
    rng = np.random.default_rng(5)
    idx = rng.choice(nx * ny * nz, min(3000, nx * ny * nz), replace=False)
    ic, ib, ia = np.unravel_index(idx, (nz, ny, nx))
    P = (np.asarray(sp.p0, np.float64)[None, :] + np.outer(ia, sp.dA) + np.outer(ib, sp.dB) + np.outer(ic, sp.dC))
    ew = np.zeros((len(P), 4))
    for qq, dz in zip(sc["Qs"], sc["QZs"]):
        if qq != 0:
            ew[:, 3] += qq * slab.phi_full(P + [0, 0, dz])
            ew[:, :3] += qq * slab.field_full(P + [0, 0, dz])
    want = lj[ic, ib, ia] + ew
    got = base[ic, ib, ia]
    S = abs_term_sums(sub, cp, sp, dict(sc, Qs=np.zeros(4), QZs=np.zeros(4)), cap=False)[ic, ib, ia]
    bound = 1e-5 * np.linalg.norm(lj[ic, ib, ia], axis=-1, keepdims=True) + 4e-6 * S + 1e-6 + \
        2e-5 * np.abs(ew).max(0)
    err = np.abs(got - want)
    print(f"[poses] ewald base: max |d| {err.max():.3e}, worst |d|/bound {(err / bound).max():.3e}")
    assert (err <= bound).all()


def _df_and_grads(scene, sc, k, v):
    dev = scene.pts.device
    pose = torch.tensor(sc["poses"][k], dtype=torch.float64, device=dev, requires_grad=True)
    q = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev, requires_grad=True)
    df = PS.pose_df(pose, q, scene)
    (df.double() * v).sum().backward()
    return df.detach().cpu().numpy(), pose.grad.cpu().numpy(), q.grad.cpu().numpy()


def test_scan_window_equals_full_grid_and_falls_back():
    """Placements on the scan window give the df and gradients of the full grid (the window's
    voxels are the full grid's; only the FP32 rounding of the window origin differs); with zero
    margins the probes come within the guard of the window's faces and the placement is
    recomputed on the full grid, with the full grid's results."""
    sc = W.small_substrate_scene()
    win, _ = make_scene(sc)
    full, _ = make_scene(sc, window_margin=None)
    tight, _ = make_scene(sc, window_margin=(0.0, 0.0))
    assert win.windowed and not full.windowed and tight.windowed
    assert win.base.numel() < full.base.numel() // 3
    v = torch.as_tensor(np.random.default_rng(8).standard_normal((*sc["nxy"], sc["nz"])), device=win.pts.device)
    for k in (0, 2):
        a = _df_and_grads(win, sc, k, v)
        b = _df_and_grads(full, sc, k, v)
        c = _df_and_grads(tight, sc, k, v)
        print(f"[poses] window vs full, placement {k}: df {np.abs(a[0] - b[0]).max():.3e} "
              f"(max {np.abs(b[0]).max():.3e}), pose grad {a[1]} vs {b[1]}")
        for x, y in zip(a, b):
            assert np.allclose(x, y, rtol=1e-4, atol=1e-5 * np.abs(y).max())
        assert np.array_equal(c[0], b[0])
        for x, y in zip(c[1:], b[1:]):
            assert np.allclose(x, y, rtol=1e-5, atol=1e-6 * np.abs(y).max())
    assert win.window_fallbacks == 0 and tight.window_fallbacks == 2


def test_kept_states_equal_recomputed_states():
    """The backward on the forward's kept probe positions gives the gradients of the backward
    that recomputes the relaxation (same positions bit for bit; the relax adjoint's atomic
    scatter order is the only difference), also after another placement overwrote the grid."""
    sc = W.small_substrate_scene()
    kept, _ = make_scene(sc)
    rec, _ = make_scene(sc, keep_states=False)
    v = torch.as_tensor(np.random.default_rng(9).standard_normal((*sc["nxy"], sc["nz"])), device=kept.pts.device)
    for k in (1, 3):
        a = _df_and_grads(kept, sc, k, v)
        b = _df_and_grads(rec, sc, k, v)
        np.testing.assert_array_equal(a[0], b[0])
        for x, y in zip(a[1:], b[1:]):
            assert np.allclose(x, y, rtol=1e-5, atol=1e-6 * np.abs(y).max())
    dev = kept.pts.device
    pose = torch.tensor(sc["poses"][0], dtype=torch.float64, device=dev, requires_grad=True)
    q = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev, requires_grad=True)
    df = PS.pose_df(pose, q, kept)
    with torch.no_grad():
        PS.pose_df(torch.tensor(sc["poses"][2], dtype=torch.float64, device=dev), q.detach(), kept)
    (df.double() * v).sum().backward()
    ref = _df_and_grads(rec, sc, 0, v)
    assert np.allclose(pose.grad.cpu().numpy(), ref[1], rtol=1e-5, atol=1e-6 * np.abs(ref[1]).max())
