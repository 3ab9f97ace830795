"""shared_postprocess on the device Δf operators (spammm/SPM/AFM_utils.py:3362-3414).

Bars: closed forms — a force linear along the oscillation direction, F.n = a + c (r.n), has
Δf = -c at every height whose oscillation stays inside the volume (the nine-node average of a
constant), and a constant forward-minus-backward force d has E_diss = |d.n| * 2 A_z per height
(the trapezoid rule is exact for a constant); the composition against the oracle's Δf and
dissipation restatements within 1e-5 of the largest value (the Δf bar of tests/test_gpu_parity.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

pytest.importorskip("torch")

from spammm_amd import postprocess as PP       # noqa: E402
from tests.golden_vectors import df_vectors    # noqa: E402


def spec(h_df, h_Fz, h_scan, amp, osc, nx=5, ny=4):
    return PP.ScanSpec(scan_xs=0.2 * np.arange(nx), scan_ys=0.3 * np.arange(ny), h_df=h_df, h_Fz=h_Fz,
                       h_scan=h_scan, amplitude=amp, osc_dir=osc, K_LAT=0.0156, K_RAD=1.25, bond_length=4.0)


def test_linear_force_and_constant_dissipation_closed_forms():
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.0, 5.0, 0.25, amp=0.5)
    nx, ny, nz = 5, 4, h_scan.size
    FEs = np.zeros((nx, ny, nz, 4), np.float32)
    FEs[..., 2] = 0.7 - 0.3 * h_scan[None, None, :]            # Fz = a + c z, c = -0.3
    bwd = FEs.copy()
    bwd[..., 2] -= 0.125                                      # F_fwd - F_bwd = (0, 0, 0.125)
    r = PP.shared_postprocess(FEs, spec(h_df, h_Fz, h_scan, 0.5, (0., 0., 2.)), FEs_bwd=bwd, backend_name="morse")
    assert r.df.shape == (nx, ny, h_df.size) and r.Fz.shape == (nx, ny, h_Fz.size)
    inner = slice(1, -1)                                      # oscillations that stay inside h_scan
    assert np.allclose(r.df[:, :, inner], 0.3, atol=1e-5)
    assert np.array_equal(r.Fz, FEs[..., 2][:, :, 0:h_Fz.size])   # h_Fz = h_df - A = h_scan[0 .. 8]
    assert np.allclose(r.E_diss, 0.125 * 1.0, rtol=1e-6)       # |d.n| * 2 A_z, A_z = 0.5
    assert r.amp_align and r.backend_name == "morse" and r["df"] is r.df
    assert set(r.tip_disp) == {"dx", "dy", "dz"} and r.tip_disp["dz"].shape == (nx, ny, nz)


def test_lateral_oscillation_closed_form():
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.0, 4.0, 0.25, amp=0.4, osc_dir=(1., 0., 0.))
    nx, ny = 12, 3
    xs = 0.2 * np.arange(nx)
    FEs = np.zeros((nx, ny, h_scan.size, 4), np.float32)
    FEs[..., 0] = (1.5 + 2.0 * xs)[:, None, None]             # Fx = b + c x, c = 2
    r = PP.shared_postprocess(FEs, spec(h_df, h_Fz, h_scan, 0.4, (1., 0., 0.), nx, ny))
    assert not r.amp_align and np.array_equal(r.E_diss, np.zeros_like(r.E_diss))
    # samples shifted by up to 0.4 A = 2 x-steps (+1 for the gradient) stay inside for 3 <= ix <= 8
    assert np.allclose(r.df[3:9], -2.0, atol=1e-5)


def test_composition_matches_oracle():
    """shared_postprocess of the fixture volume against the oracle's compute_df_amp_dir and
    compute_dissipation of it (tests/golden/df_vectors.npz, case post)."""
    V = df_vectors()
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.0, 4.0, 0.1, amp=0.3, osc_dir=(1., 2., 3.))
    FEs, bwd = V["post__FEs"], V["post__bwd"]
    nx, ny = FEs.shape[:2]
    sp = spec(h_df, h_Fz, h_scan, 0.3, (1., 2., 3.), nx, ny)
    r = PP.shared_postprocess(FEs, sp, FEs_bwd=bwd)
    idf = [int(np.argmin(np.abs(h_scan - h))) for h in h_df]
    iFz = [int(np.argmin(np.abs(h_scan - h))) for h in h_Fz]
    for a, b in ((r.df, V["post__df_amp_dir"][:, :, idf]), (r.E_diss, V["post__E_diss"])):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()
    assert np.array_equal(r.Fz, FEs[..., 2][:, :, iFz])
