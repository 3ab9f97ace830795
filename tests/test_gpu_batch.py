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
    B.LaunchConfig(4, 8, 4, "two_phase")])
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
