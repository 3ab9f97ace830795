"""The reference's scan-contract and texel-centre tests (tests/SPM/test_scan_contract.py:21-106,
tests/SPM/test_texel_centers.py:17-94) on the facade's host functions: the vectorised scan
point builder against a nested loop, the ScanSpec / ScanResult fields and defaults, and the
half-texel offset of the grid's sampling rows (AFMulator._make_dinv_axis_aligned / _lvec)."""
import numpy as np

from spammm_amd import postprocess as PP
from spammm_amd.afm import AFMulator, build_scan_points_vectorized


def test_scan_points_match_nested_loop_5x5x3():
    xs = np.array([0.0, 0.1, 0.2, 0.3, 0.4], np.float32)
    ys = np.array([1.0, 1.1, 1.2, 1.3, 1.4], np.float32)
    hs = np.array([3.0, 3.5, 4.0], np.float32)
    ref = np.zeros((len(xs) * len(ys) * len(hs), 4), np.float32)
    k = 0
    for x in xs:
        for y in ys:
            for h in hs:
                ref[k, :3] = (x, y, h)
                k += 1
    pts = build_scan_points_vectorized(xs, ys, hs)
    assert pts.shape == ref.shape and pts.dtype == np.float32 and pts.flags["C_CONTIGUOUS"]
    np.testing.assert_array_equal(pts, ref)


def test_scan_points_w_zero_and_single_height():
    pts = build_scan_points_vectorized(np.linspace(0, 1, 4, dtype=np.float32), np.linspace(0, 1, 3, dtype=np.float32),
                                       np.array([2.0, 3.0], np.float32))
    assert not pts[:, 3].any()
    one = build_scan_points_vectorized(np.array([0.0, 0.5, 1.0], np.float32), np.array([0.0, 0.5], np.float32),
                                       np.array([4.0], np.float32))
    assert one.shape == (6, 4) and np.all(one[:, 2] == 4.0)


def test_scanspec_fields_and_scanresult_defaults():
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.7, 4.7, 0.1, amp=1.0, amp_align=True)
    spec = PP.ScanSpec(scan_xs=np.arange(-2.0, 2.0, 0.1, dtype=np.float32),
                       scan_ys=np.arange(-1.0, 1.0, 0.1, dtype=np.float32), h_df=h_df, h_Fz=h_Fz, h_scan=h_scan,
                       amplitude=1.0, osc_dir=(0.0, 0.0, 1.0), K_LAT=0.0312, K_RAD=20.0, bond_length=3.0)
    for name in ("scan_xs", "scan_ys", "h_df", "h_Fz", "h_scan", "amplitude", "osc_dir", "K_LAT", "K_RAD",
                 "bond_length", "scan_margin"):
        assert hasattr(spec, name)
    assert spec.scan_margin == 2.0 and spec.amplitude == 1.0 and spec.osc_dir == (0.0, 0.0, 1.0)
    z = np.zeros((2, 2, 3), np.float32)
    res = PP.ScanResult(df=z, Fz=z, heights=np.arange(3.0), heights_Fz=np.arange(3.0), scan_xs=np.arange(2.0),
                        scan_ys=np.arange(2.0), amp_align=True, FEs=np.zeros((2, 2, 5, 4), np.float32),
                        tip_disp={k: np.zeros((2, 2, 5)) for k in ("dx", "dy", "dz")}, E_diss=z)
    assert res.backend_name == "unknown" and res.fft_path == "none"


def test_dinv_half_texel_offsets_and_diagonal():
    L, n = (10.0, 20.0, 30.0), (50, 100, 150)
    for origin in (np.zeros(3, np.float32), np.array([-5.0, 3.0, 1.0], np.float32)):
        rows = AFMulator._make_dinv_axis_aligned(L, origin, n)
        for a, r in enumerate(rows):
            assert abs(r[3] - (-origin[a] / L[a] + 0.5 / n[a])) < 1e-6
            assert r[3] != 0.0 or origin[a] != 0.0
            assert abs(r[a] - 1.0 / L[a]) < 1e-6
            assert all(r[b] == 0.0 for b in range(3) if b != a)
    rows = AFMulator._make_dinv_lvec(np.eye(3, dtype=np.float32), n)
    for a, r in enumerate(rows):
        assert abs(r[3] - 0.5 / n[a]) < 1e-6


def test_dinv_maps_voxels_to_texel_centres():
    L, n = (10.0, 20.0, 30.0), (50, 100, 150)
    dA, _, _ = AFMulator._make_dinv_axis_aligned(L, np.zeros(3, np.float32), n)
    for i in (0, 1, 25, 49):
        pos = i * L[0] / n[0]
        assert abs(pos * dA[0] + dA[3] - (i + 0.5) / n[0]) < 1e-5
