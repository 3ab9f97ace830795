"""The reference's end-to-end callers, as call sequences on the facade.

The reference's orchestration functions stay in the reference; they reach this path only through
the operator API they call.  These tests drive ``spammm_amd.afm`` in the order two of those
callers do and check what each call returns:

* run_morse_pp_afm (spammm/SPM/AFM_utils.py:4058-4108): Morse atoms from atomic numbers,
  setup_grid_world, make_forcefield, setup_fdbm_grid_from_img, forward and reverse scan_fdbm over
  scan axes padded by the lateral oscillation amplitude, the crop, shared_postprocess.  Bars: the
  Morse grid equals the FP64 oracle (the grid bar of tests/test_gpu_parity.py) on the planes the
  probes reach; the forward stroke equals the FP32 oracle's relaxStrokes on the same grid bit for
  bit; the crop of the padded scan equals the unpadded scan bit for bit (pixels are independent,
  so a wrong pad or crop fails).
* run_morse_coulomb_afm (spammm/SPM/AFM_utils.py:3932-3942): load_molecule, assign_params,
  setup_grid, make_forcefield, run_scan, compute_df(Fz, |dtip|).  Bars: Fz equals the FP32
  oracle's relaxStrokesTilted on the same grid bit for bit; df equals numpy.gradient of that Fz
  to 1e-5 of its largest value.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O                 # noqa: E402
from spammm_amd import afm as AFM              # noqa: E402
from spammm_amd import device as D             # noqa: E402
from spammm_amd import params as P             # noqa: E402
from spammm_amd import postprocess as PP       # noqa: E402
from spammm_amd import workloads as W          # noqa: E402

Z = {"H": 1, "C": 6}


def test_morse_pp_call_sequence():
    en, pos = W.naphthalene_geometry()
    types = np.array([Z[e] for e in en])
    origin, step, ngrid = np.array([-8.0, -7.0, -3.0]), 0.2, (80, 70, 60)
    amp, osc = 0.3, np.array([1.0, 0.0, 2.0]) / np.sqrt(5.0)
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.2, 4.0, 0.1, amp=amp, osc_dir=osc)
    xs, ys = np.arange(-2.0, 2.01, 0.2).astype(np.float32), np.arange(-1.0, 1.01, 0.2).astype(np.float32)

    a = AFM.AFMulator(use_morse=True)
    a.atoms_arr, a.cLJs_arr = P.morse_atoms_from_Z(pos, types)
    a.setup_grid_world(origin, np.array(ngrid) * step, ngrid)
    a.make_forcefield()
    a.setup_fdbm_grid_from_img(a.FF, ngrid, origin, step)
    kw = dict(mol_z=float(pos[:, 2].max()), K_LAT=0.0156, K_RAD=1.25, bond_length=4.0, osc_dir=osc)
    pad = int(np.ceil(amp * osc[0] / step))                 # lateral amplitude in grid steps
    xs_pad = (xs[0] + step * (np.arange(len(xs) + 2 * pad) - pad)).astype(np.float32)
    f, d = a.scan_fdbm(xs, ys, h_scan, **kw)
    b, _ = a.scan_fdbm(xs, ys, h_scan, reverse=True, **kw)
    fp, dp = a.scan_fdbm(xs_pad, ys, h_scan, **kw)
    np.testing.assert_array_equal(fp[pad:pad + len(xs)], f)
    np.testing.assert_array_equal(dp["dz"][pad:pad + len(xs)], d["dz"])
    assert not np.array_equal(f, b)                          # the two strokes stop at different states

    # the forward stroke against the FP32 oracle's relaxStrokes on the same grid
    s = a.fdbm_spec
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    hs = np.sort(h_scan)
    dh = float(hs[-1] - hs[0]) / (len(hs) - 1)
    pts = np.zeros((gx.size, 4), np.float32)
    pts[:, 0], pts[:, 1] = gx.ravel(), gy.ravel()
    pts[:, 2] = float(hs[-1]) + kw["mol_z"] + 4.0
    FEo, _, _ = O.relax_strokes(a.img_FF_fdbm.cpu().numpy(), s.n, s.dinvA, s.dinvB, s.dinvC, pts, len(hs),
                                np.array([0, 0, -dh, 0], np.float32),
                                np.array([-0.0156, -0.0156, -0.0156, -1.25], np.float32),
                                np.array([0, 0, -4.0, 4.0], np.float32), a.relax_pars, a.max_iter, a.f2conv, "f32")
    np.testing.assert_array_equal(f, FEo.reshape(len(xs), len(ys), len(hs), 4)[:, :, ::-1])

    r = PP.shared_postprocess(f, PP.ScanSpec(scan_xs=xs, scan_ys=ys, h_df=h_df, h_Fz=h_Fz, h_scan=h_scan,
                                             amplitude=amp, osc_dir=osc, K_LAT=0.0156, K_RAD=1.25,
                                             bond_length=4.0), FEs_bwd=b, tip_disp=d, backend_name="morse")
    assert r.df.shape == (len(xs), len(ys), len(h_df)) and np.isfinite(r.df).all() and np.isfinite(r.E_diss).all()

    # the Morse grid against the FP64 oracle, 1.5 A and more above the molecule (the probes stay above
    # 2.9 A; in the molecule's own plane the capped forces of neighbouring atoms cancel and an FP32 sum,
    # the oracle's FP32 mode included, exceeds this bar)
    sg = a.spec
    k0 = int(np.ceil((1.5 - origin[2]) / step))
    G = a.FF.cpu().numpy().astype(np.float64)[k0:]
    Go = O.grid_project_d(1, a.atoms_arr.astype(np.float64), a.cLJs_arr.astype(np.float64), sg.n, sg.p0, sg.dA,
                          sg.dB, sg.dC, a.tipQs, a.tipQZs, a.npbc, a.lvA, a.lvB)[k0:]
    mag = np.linalg.norm(Go, axis=-1, keepdims=True)
    assert (np.abs(G - Go) <= 1e-5 * mag + 1e-6).all()


def test_morse_coulomb_call_sequence(tmp_path):
    w = W.naphthalene()
    xyz = tmp_path / "naph.xyz"
    with open(xyz, "w") as fh:
        fh.write(f"{len(w.enames)}\nnaphthalene\n")
        for e, (x, y, z, q) in zip(w.enames, w.atoms):
            fh.write(f"{e} {x:.6f} {y:.6f} {z:.6f} {q:.6f}\n")
    a = AFM.AFMulator(use_morse=True)
    a.load_molecule(str(xyz))
    a.assign_params()
    a.setup_grid(n=(64, 64, 48), margin=5.0, z_top=8.0)
    a.make_forcefield()
    FEs, _ = a.run_scan(nxy=(12, 10), nz=10, dtip=-0.1)
    Fz = FEs[:, :, :, 2]
    df = AFM.compute_df(Fz, 0.1)
    s = a.spec
    scan_p0, scan_da, scan_db = a._auto_scan(12, 10)
    pts = D.scan_points(scan_p0, scan_da, scan_db, 12, 10)
    tip = a._tip(-0.1)
    FEo, _ = O.relax_tilted(a.FF.cpu().numpy(), s.n, s.dinvA, s.dinvB, s.dinvC, pts, 10, tip.tipA, tip.tipB,
                            tip.tipC, tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, a.max_iter, a.f2conv, False,
                            "f32")
    np.testing.assert_array_equal(Fz, FEo[..., 2].reshape(12, 10, 10))
    ref = -np.gradient(Fz.astype(np.float64), 0.1, axis=2)
    assert np.abs(df - ref).max() <= 1e-5 * np.abs(ref).max()
