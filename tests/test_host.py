"""Host-side checks that need no GPU: the C ABI library and the facade's host logic."""
import ctypes
import os
import re

import numpy as np
import pytest

from spammm_amd import _lib, device as D, params as P, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ppm_api.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ppm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libppm_hip.so not built (python -m spammm_amd.build)"
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 8
    for n in names:
        assert hasattr(L, n), f"{n} declared in ppm_api.h but not exported"
    assert set(names) == set(_lib.EXPORTS)


def test_version_and_error_string_without_gpu():
    L = _lib.load()
    assert L.ppm_version() == 2
    assert isinstance(L.ppm_last_error(), bytes)


def test_invalid_arguments_rejected_on_host():
    """Argument validation runs before any device work, so it is testable without a GPU (the device
    addresses below are never dereferenced)."""
    L = _lib.load()
    s = _lib.PpmGridSpec()
    s.nx, s.ny, s.nz = 0, 4, 4
    ctl = D.relax_ctl()
    tip = D.Tip().c()
    A, B = 0x10000, 0x20000          # 16-byte aligned stand-ins for device buffers
    call = lambda pts, fes, pps: L.ppm_relax_strokes_tilted(A, ctypes.byref(s), pts, 4, 3, ctypes.byref(tip),
                                                            ctypes.byref(ctl), fes, pps, None)
    rc = call(A, B, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG
    assert b"grid dims" in L.ppm_last_error()
    s.nx = 4
    ctl.max_iter = 0
    rc = call(A, B, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG and b"max_iter" in L.ppm_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)
    ctl.max_iter = 10
    ctl.lanes = 3
    rc = call(A, B, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG and b"lanes" in L.ppm_last_error()
    ctl.lanes = 0
    # float4 rows must be 16-byte aligned: misaligned points, FEs or PPs are rejected
    for args in ((A + 4, B, None), (A, B + 8, None), (A, B, B + 0x1000 + 12)):
        rc = call(*args)
        assert rc == _lib.PPM_ERR_INVALID_ARG and b"16-byte aligned" in L.ppm_last_error(), args


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_lib.PpmGridSpec) == 16 + 7 * 16
    assert ctypes.sizeof(_lib.PpmTip) == 7 * 16
    assert ctypes.sizeof(_lib.PpmRelaxCtl) == 28


def test_scan_points_match_reference_loop():
    """Vectorized scan starts equal the reference's double loop (spammm/SPM/AFM.py:977-980) bit for bit."""
    p0 = np.array([0.3, -1.7, 9.1], np.float32)
    da = np.array([0.21, 0.0, 0.0], np.float32)
    db = np.array([0.0, 0.19, 0.01], np.float32)
    pts = D.scan_points(p0, da, db, 7, 5)
    k = 0
    for ix in range(7):
        for iy in range(5):
            np.testing.assert_array_equal(pts[k, :3], p0 + da * ix + db * iy)
            k += 1
    assert (pts[:, 3] == 0).all()


def test_mixing_rules():
    c = P.mix_lj(["C", "H"], P.ELEMENTS_BENCH, 1.66, 0.0091)
    R0, E0 = 1.66 + 1.908, np.sqrt(0.0091 * 0.0037)
    np.testing.assert_allclose(c[0], [2 * E0 * R0 ** 6, E0 * R0 ** 12], rtol=1e-6)
    m = P.mix_morse(["H"], P.ELEMENTS_REFERENCE, 1.452, 0.0006808, -1.8)
    np.testing.assert_allclose(m[0], [1.452 + 1.443, np.sqrt(0.0006808 * 0.00190802059), -1.8, 0], rtol=1e-6)
    assert abs(P.stiffness_Nm_to_eVA2(0.25) - 0.25 / 16.02176634) < 1e-15


def test_workloads_match_configuration():
    w = W.naphthalene()
    assert w.enames.count("C") == 10 and w.enames.count("H") == 8
    assert abs(w.atoms[:, 3].sum()) < 1e-6 and np.abs(w.atoms[:, 3]).max() <= 0.3   # U(-0.15,0.15) - mean
    C = w.atoms[:10, :3]
    d = np.linalg.norm(C[:, None] - C[None], axis=-1)
    assert np.isclose(np.sort(d[d > 0])[0], 1.40, atol=1e-6)
    assert (np.sum(np.isclose(d, 1.40, atol=1e-6), axis=1) >= 2).all()
    H = w.atoms[10:, :3]
    dh = np.linalg.norm(H[:, None] - C[None], axis=-1).min(1)
    np.testing.assert_allclose(dh, 1.08, atol=1e-6)
    assert w.n == (200, 200, 200) and w.nxy == (100, 100) and w.nz == 40
    f = W.flake()
    assert f.atoms.shape[0] == 200 and abs(f.atoms[:, 3].sum()) < 1e-5
    heavy = [e for e in f.enames if e != "H"]
    assert abs(sum(e in ("N", "O") for e in heavy) - 0.1 * len(heavy)) <= 1
    assert f.n == (300, 300, 240) and f.nxy == (256, 256) and f.nz == 60


def test_batch_struct_layouts_match_header():
    from spammm_amd import batch as B
    assert ctypes.sizeof(B.PpmMolDesc) == 16 * 4 + 8 + 12 * 16
    assert B.PpmMolDesc.grid_off.offset == 64
    assert ctypes.sizeof(B.PpmBatchParams) == 12 * 4 + 2 * 16 + 7 * 16 + 28


def test_batch_z_window_covers_probe_heights():
    from spammm_amd import batch as B
    w = W.batch_molecule(3)
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
    klo, nzw = B.z_window(w, sp, tip, w.nz, 1.5)
    z = sp.p0[2] + sp.dC[2] * np.array([klo, klo + nzw - 1])
    probe_hi = w.scan_p0[2] - w.bond
    probe_lo = probe_hi + (w.nz - 1) * w.dtip
    assert z[0] <= probe_lo - 1.5 and z[1] >= probe_hi + 1.5 + sp.dC[2]
    assert nzw < w.n[2] // 2
    assert B.z_window(w, sp, tip, w.nz, 100.0) == (0, w.n[2])


def test_batch_workload_matches_configuration():
    ms = W.batch(16)
    na = [m.atoms.shape[0] for m in ms]
    assert min(na) >= 10 and max(na) <= 150
    for m in ms:
        assert 160 <= m.n[0] <= 320 and m.n[0] == m.n[1] and m.n[2] == 200
        assert m.nxy == (128, 128) and m.nz == 40


def test_batch_z_window_asymmetric_margins():
    from spammm_amd import batch as B
    w = W.batch_molecule(5)
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
    klo_s, nzw_s = B.z_window(w, sp, tip, w.nz, 1.0)
    klo_a, nzw_a = B.z_window(w, sp, tip, w.nz, (1.0, 1.0))
    assert (klo_s, nzw_s) == (klo_a, nzw_a)
    klo, nzw = B.z_window(w, sp, tip, w.nz, (0.2, 0.6))
    z = sp.p0[2] + sp.dC[2] * np.array([klo, klo + nzw - 1])
    probe_hi = w.scan_p0[2] - w.bond
    probe_lo = probe_hi + (w.nz - 1) * w.dtip
    assert z[0] <= probe_lo - 0.2 and z[1] >= probe_hi + 0.6 + sp.dC[2]
    assert nzw < nzw_s


def test_rocksalt_slab_and_placements():
    en, sub = W.rocksalt_slab()
    assert sub.shape == (1024, 4) and abs(sub[:, 3].sum()) < 1e-4
    assert set(en) == {"Na", "Cl"} and abs(np.abs(sub[:, 3]) - 0.8).max() < 1e-6
    top = sub[sub[:, 2] == 0]
    d = np.linalg.norm(top[:, None, :2] - top[None, :, :2], axis=-1)
    nn = np.isclose(d, 5.64 / 2, atol=1e-5)
    assert (top[:, 3][:, None] * top[:, 3][None, :])[nn].max() < 0      # nearest neighbours alternate
    sc = W.substrate_scene()
    assert sc["n"] == (902, 902, 300) and sc["mol_atoms"].shape == (60, 4) and sc["poses"].shape == (64, 4)
    assert abs(sc["mol_atoms"][:, 2].mean() - 3.0) < 0.1
    from spammm_amd import poses as PS
    x0 = sc["mol_atoms"][:, :3].astype(np.float64)
    c = x0.mean(0)
    x = PS.place(x0, c, sc["poses"][5])
    d0 = np.linalg.norm(x0[:, None] - x0[None], axis=-1)
    d1 = np.linalg.norm(x[:, None] - x[None], axis=-1)
    np.testing.assert_allclose(d0, d1, atol=1e-12)                    # rigid
    np.testing.assert_allclose(x.mean(0) - c, [*sc["poses"][5][:2], 0.0], atol=1e-12)


def test_ewald_sizes_and_validation_on_host():
    """ppm_ewald_* sizes and argument checks (no device work)."""
    L = _lib.load()
    assert L.ppm_ewald_num_g(3) == 48 and L.ppm_ewald_num_g(32) == 65 * 65 - 1
    assert L.ppm_ewald_system_bytes(4, 3) >= 4 * 48 * 16 + 48 * (4 + 16)
    assert L.ppm_ewald_system_bytes(4, 0) == 0 and b"n_harm" in L.ppm_last_error()
    assert L.ppm_ewald_grid_workspace_size(8, 10) >= 10 * 288 * 16 + 10 * 8
    D2 = ctypes.c_double * 2
    a, b, deg = D2(4.0, 0.0), D2(0.0, 4.0), D2(8.0, 0.0)
    rc = L.ppm_ewald_prepare(1, 1, 2, ctypes.byref(a), ctypes.byref(deg), 3, 0.0, 1, 1 << 20, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG and b"degenerate" in L.ppm_last_error()
    rc = L.ppm_ewald_prepare(1, 1, 2, ctypes.byref(a), ctypes.byref(b), 3, 0.0, 1, 16, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG and b"too small" in L.ppm_last_error()
    rc = L.ppm_ewald_eval(7, 1, 1, 1, 1 << 20, 2, 3, ctypes.byref(a), ctypes.byref(b), 1, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG and b"mode" in L.ppm_last_error()
    rc = L.ppm_ewald_eval(0, 1, 1, 1, 64, 2, 3, ctypes.byref(a), ctypes.byref(b), 1, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG and b"system buffer too small" in L.ppm_last_error()
    s = _lib.PpmGridSpec()
    s.nx, s.ny, s.nz = 4, 4, 4
    s.dA, s.dB, s.dC = _lib.f4((1, 0, 0)), _lib.f4((0, 1, 0)), _lib.f4((0.2, 0, 1))
    rc = L.ppm_ewald_tip_grid(1, 1 << 20, 2, 3, ctypes.byref(a), ctypes.byref(b), ctypes.byref(s), _lib.f4((1,)),
                              _lib.f4((0,)), 0, 1, 1 << 20, 1, None)
    assert rc == _lib.PPM_ERR_INVALID_ARG and b"dC along z" in L.ppm_last_error()
