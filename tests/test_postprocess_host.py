"""Height ladders and the scan-result contract (host side; spammm/SPM/AFM_utils.py:653-681,
:3307-3359).  The ladder values are pinned by closed forms: h_df = h_min + k*h_step, the
vertical amplitude A*|n_z| of a unit oscillation direction lowers h_Fz and widens h_scan."""
import numpy as np
import pytest

from spammm_amd import postprocess as PP


def test_vertical_ladders_are_amp_aligned():
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.0, 5.0, 0.25, amp=0.5)
    assert h_df.dtype == h_Fz.dtype == h_scan.dtype == np.float32
    assert np.array_equal(h_df, 3.0 + 0.25 * np.arange(9, dtype=np.float32))
    assert np.array_equal(h_Fz, h_df - np.float32(0.5))
    assert np.array_equal(h_scan, 2.5 + 0.25 * np.arange(13, dtype=np.float32))


def test_lateral_and_tilted_oscillation():
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(4.0, 6.0, 0.5, amp=1.0, osc_dir=(1.0, 0.0, 0.0))
    assert np.array_equal(h_Fz, h_df) and np.array_equal(h_scan, h_df)    # no vertical amplitude
    # n = (3,0,4)/5: vertical amplitude 0.8 A for amp = 1 (direction need not be unit)
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(4.0, 6.0, 0.2, amp=1.0, osc_dir=(3.0, 0.0, 4.0))
    assert np.allclose(h_Fz, h_df - 0.8, atol=1e-6)
    assert abs(h_scan[0] - 3.2) < 1e-6 and abs(h_scan[-1] - 6.8) < 1e-5
    assert np.allclose(np.diff(h_scan), 0.2, atol=1e-5)
    _, h_Fz2, _ = PP.afm_df_height_stacks(4.0, 6.0, 0.2, amp=1.0, amp_align=False, osc_dir=(3.0, 0.0, 4.0))
    assert np.array_equal(h_Fz2, h_df)


def test_ladder_errors():
    for bad in ((0.0, 0.0, 0.0), (np.nan, 0.0, 1.0), (1.0, 0.0), (0.0, 0.0, np.inf)):
        with pytest.raises(ValueError, match="osc_dir"):
            PP.afm_df_height_stacks(3.0, 5.0, 0.25, osc_dir=bad)
    with pytest.raises(ValueError, match="denser"):
        PP.afm_df_height_stacks(3.0, 3.0, 0.1, amp=0.0)      # a single height: h_scan has 1 entry


def test_scan_result_reads_like_a_dict():
    z = np.zeros((2, 2, 3), np.float32)
    r = PP.ScanResult(df=z, Fz=z, heights=np.zeros(3), heights_Fz=np.zeros(3), scan_xs=np.zeros(2),
                      scan_ys=np.zeros(2), amp_align=True, FEs=np.zeros((2, 2, 3, 4)), tip_disp={}, E_diss=z)
    assert r["df"] is z and "E_diss" in r and "nothing" not in r and r.get("nothing", 7) == 7
    assert r.backend_name == "unknown" and r["fft_path"] == "none"


def test_nearest_picks_the_first_on_a_tie():
    assert PP._nearest(np.array([0.0, 1.0, 2.0]), [0.5, 1.6, -3.0, 9.0]) == [0, 2, 0, 2]


def test_morse_atoms_from_Z():
    """morse_atoms_from_Z equals the element-name mixing rule (zero charges) and rejects an atomic
    number without an element entry."""
    from spammm_amd import params as P
    pos = np.array([[0., 0., 0.], [1.1, 0., 0.], [0., 1.2, 0.3]])
    atoms, cMs = P.morse_atoms_from_Z(pos, [6, 1, 8])
    ref = P.mix_morse(["C", "H", "O"], P.ELEMENTS_REFERENCE, 1.452, 0.0006808, -1.8)
    assert np.array_equal(cMs, ref) and np.array_equal(atoms[:, :3], pos.astype(np.float32))
    assert not atoms[:, 3].any()
    with pytest.raises(KeyError):
        P.morse_atoms_from_Z(pos[:1], [92])
