"""The reference's end-to-end callers on the device path (spammm_amd.runners).

* run_morse_pp_afm: the Morse grid equals the FP64 oracle grid (the grid bar of
  tests/test_gpu_parity.py); the cropped forward / reverse strokes of the amplitude-padded scan
  equal scan_fdbm over the unpadded scan bit for bit (pixels are independent, so a wrong crop
  offset or pad width fails); df, Fz and E_diss equal shared_postprocess of those strokes bit for
  bit; the reference's margin check and extras.
* run_morse_coulomb_afm: Fz equals the FP32 oracle relaxation on the same grid bit for bit, df
  equals the oracle compute_df, heights follow the reference's formula, the .npz holds the
  returned arrays.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import df_ref as R                 # noqa: E402
from oracle import oracle as O                 # noqa: E402
from spammm_amd import postprocess as PP       # noqa: E402
from spammm_amd import runners as RN           # noqa: E402
from spammm_amd import workloads as W          # noqa: E402
from spammm_amd.afm import AFMulator           # noqa: E402

Z = {"H": 1, "C": 6}


def _spec(osc=(1.0, 0.0, 2.0), amp=0.3):
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.2, 4.0, 0.1, amp=amp, osc_dir=osc)
    # the scan axes step by the grid step, as the reference's pipeline builds them: the padded axes
    # of run_morse_pp_afm continue scan_xs[0] in grid steps (spammm/SPM/AFM_utils.py:4086-4087)
    return PP.ScanSpec(scan_xs=np.arange(-2.0, 2.01, 0.2), scan_ys=np.arange(-1.0, 1.01, 0.2), h_df=h_df,
                       h_Fz=h_Fz, h_scan=h_scan, amplitude=amp, osc_dir=osc, K_LAT=0.0156, K_RAD=1.25,
                       bond_length=4.0, scan_margin=2.0)


def test_run_morse_pp_afm(tmp_path):
    en, pos = W.naphthalene_geometry()
    types = np.array([Z[e] for e in en])
    origin, step, ngrid = (-8.0, -7.0, -3.0), 0.2, (80, 70, 60)
    sp = _spec()
    r = RN.run_morse_pp_afm("naph", pos, types, origin, step, ngrid, str(tmp_path), scan_spec=sp)
    nx, ny = len(sp.scan_xs), len(sp.scan_ys)
    assert r.backend_name == "morse" and r.path == "Morse" and r.tag == "naph"
    assert r.df.shape == (nx, ny, len(sp.h_df)) and r.FEs.shape == (nx, ny, len(sp.h_scan), 4)
    assert np.isfinite(r.df).all() and np.isfinite(r.E_diss).all()

    # the same grid and unpadded strokes through the facade
    a = AFMulator(use_morse=True)
    a.atoms_arr, a.cLJs_arr = RN.morse_atoms_from_Z(pos, types)
    a.setup_grid_world(origin, np.array(ngrid) * step, ngrid)
    a.make_forcefield()
    a.setup_fdbm_grid_from_img(a.FF, ngrid, origin, step)
    kw = dict(mol_z=float(pos[:, 2].max()), K_LAT=sp.K_LAT, K_RAD=sp.K_RAD, bond_length=sp.bond_length,
              osc_dir=sp.osc_dir)
    f, d = a.scan_fdbm(sp.scan_xs.astype(np.float32), sp.scan_ys.astype(np.float32), sp.h_scan, **kw)
    b, _ = a.scan_fdbm(sp.scan_xs.astype(np.float32), sp.scan_ys.astype(np.float32), sp.h_scan, reverse=True, **kw)
    np.testing.assert_array_equal(r.FEs, f)
    np.testing.assert_array_equal(r.tip_disp["dz"], d["dz"])
    ref = PP.shared_postprocess(f, sp, FEs_bwd=b, tip_disp=d, backend_name="morse")
    for k in ("df", "Fz", "E_diss", "heights", "heights_Fz"):
        np.testing.assert_array_equal(r[k], ref[k])

    # the Morse grid against the FP64 oracle, on the planes 1.5 A and more above the molecule (the
    # probes stay above 2.9 A; in the molecule's own plane the capped forces of neighbouring atoms
    # cancel and an FP32 sum, the oracle's FP32 mode included, exceeds this bar by up to 3.4x)
    s = a.spec
    k0 = int(np.ceil((1.5 - origin[2]) / step))
    G = a.FF.cpu().numpy().astype(np.float64)[k0:]
    Go = O.grid_project_d(1, a.atoms_arr.astype(np.float64), a.cLJs_arr.astype(np.float64), s.n, s.p0, s.dA, s.dB,
                          s.dC, a.tipQs, a.tipQZs, a.npbc, a.lvA, a.lvB)[k0:]
    mag = np.linalg.norm(Go, axis=-1, keepdims=True)
    assert (np.abs(G - Go) <= 1e-5 * mag + 1e-6).all()

    with pytest.raises(ValueError, match="margin"):
        RN.run_morse_pp_afm("naph", pos, types, origin, step, ngrid, str(tmp_path), scan_spec=_spec(amp=5.0),
                            margin=4.0)
    with pytest.raises(ValueError, match="PNG"):
        RN.run_morse_pp_afm("naph", pos, types, origin, step, ngrid, str(tmp_path), scan_spec=sp, plots={"df"})


def test_run_morse_coulomb_afm(tmp_path):
    w = W.naphthalene()
    xyz = tmp_path / "naph.xyz"
    with open(xyz, "w") as fh:
        fh.write(f"{len(w.enames)}\nnaphthalene\n")
        for e, (x, y, z, q) in zip(w.enames, w.atoms):
            fh.write(f"{e} {x:.6f} {y:.6f} {z:.6f} {q:.6f}\n")
    out = RN.run_morse_coulomb_afm(str(xyz), str(tmp_path / "out"), n=(64, 64, 48), nxy=(12, 10), nz_scan=10,
                                   dtip=-0.1)
    a = out["afmulator"]
    s = a.spec
    scan_p0, scan_da, scan_db = a._auto_scan(12, 10)
    from spammm_amd import device as D
    pts = D.scan_points(scan_p0, scan_da, scan_db, 12, 10)
    tip = a._tip(-0.1)
    FEo, _ = O.relax_tilted(a.FF.cpu().numpy(), s.n, s.dinvA, s.dinvB, s.dinvC, pts, 10, tip.tipA, tip.tipB,
                            tip.tipC, tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, a.max_iter, a.f2conv, False,
                            "f32")
    Fz_o = FEo[..., 2].reshape(12, 10, 10)
    np.testing.assert_array_equal(out["Fz"], Fz_o)
    assert np.allclose(out["df"], R.compute_df(Fz_o.astype(np.float64), 0.1), rtol=1e-5,
                       atol=1e-5 * np.abs(out["df"]).max())
    assert np.allclose(out["heights"], 9.0 - 0.1 * np.arange(10))
    z = np.load(out["npz"])
    for k in ("Fz", "df", "heights"):
        np.testing.assert_array_equal(z[k], out[k])
    assert os.path.exists(out["summary"]) and out["pot"] == "Morse" and out["pngs"] == []
