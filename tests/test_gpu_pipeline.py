"""Consecutive plain scans on two or three streams (spammm_amd.pipeline) give, image by image, the
bits of the same scan run alone on one stream: the naphthalene workload (200^3 grid, 100x100 pixels
x 40 z-steps) with five different atom sets (positions displaced by N(0, 0.1 A), charges kept), the
Giessibl df fused into the relaxation, and the outputs of each image read between submissions while
later images are in flight."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from spammm_amd import _lib, df as DF, device as D, workloads as W   # noqa: E402
from spammm_amd.pipeline import ScanPipeline                          # noqa: E402


def _atom_sets(w, n, dev):
    out = []
    for k in range(n):
        a = w.atoms.copy()
        a[:, :3] = W.perturbed_positions(w, 0.1, seed=20 + k).astype(np.float32)
        out.append(torch.as_tensor(a, device=dev))
    return out


def _alone(w, sp, atoms, cpar, pts, tip, dfop):
    G = D.grid_project(atoms, cpar, _lib.PPM_MODEL_LJ, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB, True)
    df = torch.empty((pts.shape[0], dfop.n_out), dtype=torch.float32, device=pts.device)
    FEs, PPs = D.relax_tilted(G, sp, pts, w.nz, tip, w.max_iter, w.f2conv, False, w.nxy[1], df_op=dfop, df=df)
    return FEs, PPs, df


@pytest.mark.parametrize("depth", [2, 3])
def test_pipeline_images_equal_scans_alone(depth):
    dev = D.require_gpu()
    w = W.naphthalene()
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    cpar = torch.as_tensor(w.cLJ, device=dev)
    pts = torch.as_tensor(D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy), device=dev)
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
    dfop = DF.z_operator("giessibl", w.nz, w.dtip, dev, w.amp, w.f0, w.k_cant)
    sets = _atom_sets(w, 5, dev)
    want = [_alone(w, sp, a, cpar, pts, tip, dfop) for a in sets]
    torch.cuda.synchronize()
    pipe = ScanPipeline(sp, pts, w.nz, tip, _lib.PPM_MODEL_LJ, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB, True, w.max_iter,
                        w.f2conv, False, w.nxy[1], df_op=dfop, depth=depth)
    got, pending = [], []
    for k, a in enumerate(sets):
        pending.append(pipe.submit(a, cpar))
        if len(pending) == depth:               # the oldest slot is reused by the next submission
            sl = pending.pop(0)
            sl.wait()
            got.append((sl.FEs.clone(), sl.PPs.clone(), sl.df.clone()))
    for sl in pending:
        sl.wait()
        got.append((sl.FEs.clone(), sl.PPs.clone(), sl.df.clone()))
    torch.cuda.synchronize()
    assert len(got) == len(want)
    for k, (g, r) in enumerate(zip(got, want)):
        for name, x, y in zip(("FEs", "PPs", "df"), g, r):
            print(f"[pipeline depth {depth}] image {k} {name}: max |d| {(x - y).abs().max().item():.3e}", flush=True)
            assert torch.equal(x, y), (k, name)
    # the images differ from each other (a pipeline that mixed up its buffers could not pass)
    assert not torch.equal(want[0][2], want[1][2])
    pipe.close()
    assert torch.equal(pipe.slots[(len(sets) - 1) % depth].df, want[-1][2])    # buffers outlive the streams
    with pytest.raises(RuntimeError):
        pipe.submit(sets[0], cpar)


def test_pipeline_rejects_one_buffer_set():
    dev = D.require_gpu()
    w = W.small()
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    pts = torch.as_tensor(D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy), device=dev)
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
    with pytest.raises(ValueError):
        ScanPipeline(sp, pts, w.nz, tip, depth=1)
