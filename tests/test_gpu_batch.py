"""Batched persistent forward vs the plain per-image path and the oracle.

Bars: every molecule's FEs, probe positions and df of the batched launch equal the plain
per-image launches bit for bit (same arithmetic, different work decomposition), for every
launch shape and schedule, and for every replay of the captured hipGraph; a molecule whose
probes leave the projected z window is detected and recomputed; one molecule is also checked
bit-exact against the FP32 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O                 # noqa: E402
from spammm_amd import batch as B              # noqa: E402
from spammm_amd import device as D             # noqa: E402
from spammm_amd import workloads as W          # noqa: E402


def make_batch(n=7, **kw):
    mols = [W.small_batch_molecule(i) for i in range(n)]
    return B.BatchForward(mols, **kw)


def plain_reference(bf):
    FEs, df = bf.plain_loop()
    torch.cuda.synchronize()
    return FEs.cpu().numpy(), df.cpu().numpy()


@pytest.mark.parametrize("cfg", [
    B.LaunchConfig(8, 8, 3, "fused"), B.LaunchConfig(4, 16, 2, "fused"), B.LaunchConfig(8, 4, 1, "two_phase"),
    B.LaunchConfig(4, 8, 4, "two_phase"), B.LaunchConfig(4, 8, 2, "fused", lag=1),
    B.LaunchConfig(8, 4, 3, "fused", lag=3), B.LaunchConfig(4, 8, 3, "two_phase", lag=2)])
def test_batch_equals_plain_loop(cfg):
    bf = make_batch()
    print(f"[batch] {cfg.key()} windows {bf.windows}", flush=True)
    df = bf.forward(cfg)
    torch.cuda.synchronize()
    assert bf.last_fallback.size == 0
    FEp, dfp = plain_reference(bf)
    np.testing.assert_array_equal(bf.FEs.cpu().numpy(), FEp)
    np.testing.assert_array_equal(df.cpu().numpy(), dfp)
    assert (bf.PPs.cpu().numpy()[..., 3] > 0).mean() > 0.95


def test_batch_member_bitexact_vs_fp32_oracle():
    bf = make_batch(3)
    bf.forward(B.LaunchConfig())
    i = 2
    w, sp, d = bf.molecules[i], bf.specs[i], bf.descs_host[i]
    G = D.grid_project(torch.as_tensor(w.atoms, device=bf.device), torch.as_tensor(w.cLJ, device=bf.device), 0, sp,
                       w.Qs, w.QZs, w.npbc, w.lvA, w.lvB)
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    t = bf.tip
    FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, bf.nz, t.tipA, t.tipB, t.tipC,
                              t.stiffness, t.dpos0, t.relax, t.surfFF, bf.max_iter, bf.f2conv, False, "f32")
    p0, p1 = d.pix_off, d.pix_off + d.scan_nx * d.scan_ny
    np.testing.assert_array_equal(bf.FEs[p0:p1].cpu().numpy(), FEo)
    np.testing.assert_array_equal(bf.PPs[p0:p1].cpu().numpy(), PPo)


def test_window_violation_is_detected_and_recomputed():
    bf = make_batch(4, z_margin=-0.5)     # windows thinner than the probe range: probes must leave them
    df = bf.forward(B.LaunchConfig())
    torch.cuda.synchronize()
    assert bf.last_fallback.size > 0, "no molecule was flagged"
    FEp, dfp = plain_reference(bf)
    np.testing.assert_array_equal(bf.FEs.cpu().numpy(), FEp)
    np.testing.assert_array_equal(df.cpu().numpy(), dfp)


def test_graph_replays_equal_eager():
    bf = make_batch(5)
    cfg = B.LaunchConfig(8, 8, 2, "fused")
    bf.launch(cfg)
    torch.cuda.synchronize()
    ref_df = bf.df.clone()
    ref_fe = bf.FEs.clone()
    bf.capture(cfg)
    for rep in range(3):
        bf.df.zero_()
        bf.FEs.zero_()
        junk = torch.randn(1 << 20, device=bf.device).sum()      # unrelated eager work between replays
        bf.graph.replay()
        torch.cuda.synchronize()
        err, flags = bf.status()
        assert err == 0 and not flags.any()
        assert torch.equal(bf.df, ref_df), f"replay {rep} differs"
        assert torch.equal(bf.FEs, ref_fe), f"replay {rep} differs"
        del junk


def test_autotune_picks_a_config_with_identical_results(tmp_path, monkeypatch):
    from spammm_amd import autotune
    monkeypatch.setenv("SPAMMM_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    bf = make_batch(4)
    cand = {"vz": (4, 8), "tile_w": (8, 16), "workers_per_cu": (2,), "schedule": ("fused", "two_phase")}
    best, res = autotune.tune(bf, cand, reps=1)
    assert len(res) == 8 and best.key() in res
    best2, _ = autotune.tune(bf, cand, reps=1)      # served from the cache
    assert best2.key() == best.key()
    ref = bf.forward(best).clone()
    best3, whole = autotune.refine(bf, best, lags=(2, 3), reps=1)   # every variant checked bit for bit inside
    assert best3.key() in whole and len(whole) >= 3
    assert torch.equal(bf.forward(best3), ref)


def test_grid_and_relax_windows_equal_the_whole_grid():
    """ppm_grid_project_window stores the same planes as ppm_grid_project bit for bit; the windowed
    relaxation equals the whole-grid one where every probe stays inside, and raises its flag when
    the window is too thin."""
    w = W.small()
    d = D.require_gpu()
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    atoms, cpar = torch.as_tensor(w.atoms, device=d), torch.as_tensor(w.cLJ, device=d)
    Gf = D.grid_project(atoms, cpar, 0, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB)
    pts = torch.as_tensor(D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy), device=d)
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
    FEf, PPf = D.relax_tilted(Gf, sp, pts, w.nz, tip, 100, 1e-8, scan_ny=w.nxy[1])
    pz = PPf[..., 2].cpu().numpy()
    step = float(sp.dC[2])
    for klo, nzw in ((3, 37), (int((pz.min() - sp.p0[2]) / step) - 3, int((pz.max() - pz.min()) / step) + 8)):
        Gw = D.grid_project(atoms, cpar, 0, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB, window=(klo, nzw))
        assert torch.equal(Gw, Gf[klo:klo + nzw]), (klo, nzw)
        flag = torch.zeros(1, dtype=torch.int32, device=d)
        FEw, PPw = D.relax_tilted(Gw, sp, pts, w.nz, tip, 100, 1e-8, scan_ny=w.nxy[1], window=(klo, nzw), flag=flag)
        print(f"[window] planes [{klo}, {klo + nzw}) of {sp.n[2]}: flag {int(flag.item())}", flush=True)
        assert int(flag.item()) == 0
        assert torch.equal(FEw, FEf) and torch.equal(PPw, PPf)
    klo = int((pz.mean() - sp.p0[2]) / step)
    Gw = D.grid_project(atoms, cpar, 0, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB, window=(klo, 3))
    flag = torch.zeros(1, dtype=torch.int32, device=d)
    D.relax_tilted(Gw, sp, pts, w.nz, tip, 100, 1e-8, scan_ny=w.nxy[1], window=(klo, 3), flag=flag)
    assert int(flag.item()) != 0


def test_flake_grid_windows_of_any_depth_equal_the_whole_grid():
    """The LJ grid takes 16 or 8 voxels per thread by launch shape (on 300^2 columns: 8 for windows of
    16, 40 and 62 planes, 16 for 96, 149 and the whole 240); every form stores the planes of the whole
    grid bit for bit."""
    w = W.flake()
    d = D.require_gpu()
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    atoms, cpar = torch.as_tensor(w.atoms, device=d), torch.as_tensor(w.cLJ, device=d)
    Gf = D.grid_project(atoms, cpar, 0, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB)
    for klo, nzw in ((0, 16), (57, 40), (100, 62), (5, 96), (91, 149)):
        Gw = D.grid_project(atoms, cpar, 0, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB, window=(klo, nzw))
        print(f"[window] flake planes [{klo}, {klo + nzw}): max |d| {(Gw - Gf[klo:klo + nzw]).abs().max().item():.3e}",
              flush=True)
        assert torch.equal(Gw, Gf[klo:klo + nzw]), (klo, nzw)


def test_windowed_plain_loop_equals_batch_and_whole_grids():
    bf = make_batch(6)
    df = bf.forward(B.LaunchConfig(8, 8, 2, "fused"))
    FEw, dfw = bf.plain_loop(windowed=True)
    torch.cuda.synchronize()
    assert bf.last_plain_fallback.size == 0
    assert torch.equal(FEw, bf.FEs) and torch.equal(dfw, df)
    for windowed in (False, True):         # a molecule's relaxation beside the next one's projection
        FE2, df2 = bf.plain_loop(windowed=windowed, streams=2)
        torch.cuda.synchronize()
        assert torch.equal(FE2, bf.FEs) and torch.equal(df2, df), windowed
    bf2 = make_batch(4, z_margin=-0.5)     # thin windows: the windowed plain loop must fall back as well
    FEw2, dfw2 = bf2.plain_loop(windowed=True)
    torch.cuda.synchronize()
    assert bf2.last_plain_fallback.size > 0
    FEp, dfp = plain_reference(bf2)
    np.testing.assert_array_equal(FEw2.cpu().numpy(), FEp)
    np.testing.assert_array_equal(dfw2.cpu().numpy(), dfp)


def test_batched_forward_at_full_size_32_members():
    """32 members of the 512-molecule batch (U{10..150} atoms, cells U(16,32) A x 20 A on 0.1 A grids,
    128x128 pixels x 40 z-steps), autotuned launch shape, captured graph: bit-exact against the
    plain per-image loop over whole grids, no window fallback, no wait timeout; a 64-pixel
    subsample of two members bit-exact against the FP32 oracle on the member's whole grid."""
    from spammm_amd import autotune
    mols = W.batch(32)
    bf = B.BatchForward(mols, want_pps=True)
    best, _ = autotune.tune(bf, use_cache=False, reps=1)
    bf.capture(best)
    bf.graph.replay()
    torch.cuda.synchronize()
    err, flags = bf.status()
    print(f"[c2] launch {best.key()}, err {err}, window flags {int(np.count_nonzero(flags))}", flush=True)
    assert err == 0 and not flags.any()
    FEp, dfp = bf.plain_loop()
    torch.cuda.synchronize()
    assert torch.equal(bf.FEs, FEp) and torch.equal(bf.df, dfp)
    PPs = bf.PPs.cpu().numpy()
    assert np.isfinite(PPs).all() and (PPs[..., 3] > 0).mean() > 0.99
    t = bf.tip
    for i in (0, 17):
        w, sp, d = bf.molecules[i], bf.specs[i], bf.descs_host[i]
        G = D.grid_project(torch.as_tensor(w.atoms, device=bf.device), torch.as_tensor(w.cLJ, device=bf.device), 0,
                           sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB)
        pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
        sel = np.random.default_rng(i).choice(pts.shape[0], 64, replace=False)
        FEo, PPo = O.relax_tilted(G.cpu().numpy(), sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts[sel], bf.nz, t.tipA,
                                  t.tipB, t.tipC, t.stiffness, t.dpos0, t.relax, t.surfFF, bf.max_iter, bf.f2conv,
                                  False, "f32")
        p0 = d.pix_off
        np.testing.assert_array_equal(bf.FEs[p0:p0 + pts.shape[0]].cpu().numpy()[sel], FEo)
        np.testing.assert_array_equal(PPs[p0:p0 + pts.shape[0]][sel], PPo)
        print(f"[c2] member {i}: {w.atoms.shape[0]} atoms, grid {tuple(sp.n)}, 64 pixels bit-exact", flush=True)
