"""FDBM force-field producer on the GPU against the FP64 NumPy oracle (oracle/fdbm_ref.py, pinned
by closed forms in tests/test_oracle_fdbm.py).

Bars: FFT products and the fields built from them within 2e-5 of the largest |value| (FP32
FFTs of ~1e5-point grids, as the reference's complex64 path); dispersion within 1e-5 relative
per voxel; pad/roll and flip bit-exact; the periodic central-difference grid within 1e-5 of the
largest force; scans over the produced grid bit-identical to scans over the same grid uploaded
from the host.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import fdbm_ref as F            # noqa: E402
from spammm_amd import _lib, afm, fdbm      # noqa: E402
from spammm_amd import device as D          # noqa: E402
from fdbm_cases import case                 # noqa: E402


def close(a, b, rel):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err, scale = np.abs(a - b).max(), np.abs(b).max()
    print(f"[fdbm] max err {err:.3e} of {scale:.3e}")
    assert err <= rel * scale


@pytest.mark.parametrize("shape", [(24, 20, 30), (15, 21, 35)])
def test_fft_modes_match_numpy(shape):
    rng = np.random.default_rng(7)
    a = rng.normal(size=shape).astype(np.float32)
    b = rng.normal(size=shape).astype(np.float32)
    step = 0.25
    dev = D.require_gpu()
    ta, tb = torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev)
    a0 = ta.clone()
    A, B = np.fft.fftn(a.astype(np.float64)), np.fft.fftn(b.astype(np.float64))
    k2 = F._k2(shape, step)
    k2i = np.where(k2 == 0, 0.0, 1.0 / np.where(k2 == 0, 1.0, k2))
    expect = {_lib.PPM_FFT_CORRELATE: A * np.conj(B), _lib.PPM_FFT_CONVOLVE: A * B,
              _lib.PPM_FFT_POISSON: A * k2i, _lib.PPM_FFT_POISSON_CONVOLVE: A * B * k2i}
    for mode, spec in expect.items():
        out = fdbm.fft_combine(mode, ta, tb, step, 1.7).cpu().numpy()
        close(out, 1.7 * np.real(np.fft.ifftn(spec)), 2e-5)
    assert torch.equal(ta, a0), "the forward transform modified its input"


def test_fft_plans_shared_by_threads_on_their_own_streams_and_released():
    """Four host threads, each on its own (default-priority) stream, run FFT products of the same
    shape at once: each result equals the single-threaded one bit for bit (a thread's stream binding
    and transforms run as a unit under the plan lock); after ppm_fdbm_release_plans the next call
    recreates the plans with the same results."""
    import threading
    shape = (24, 20, 30)
    dev = D.require_gpu()
    rng = np.random.default_rng(3)
    ins = [(torch.as_tensor(rng.normal(size=shape).astype(np.float32), device=dev),
            torch.as_tensor(rng.normal(size=shape).astype(np.float32), device=dev)) for _ in range(4)]
    ref = [fdbm.fft_combine(_lib.PPM_FFT_POISSON_CONVOLVE, a, b, 0.25, 1.3) for a, b in ins]
    torch.cuda.synchronize()
    outs = [None] * 4

    def work(i):
        st = torch.cuda.Stream(device=dev)
        with torch.cuda.stream(st):
            for _ in range(5):
                outs[i] = fdbm.fft_combine(_lib.PPM_FFT_POISSON_CONVOLVE, ins[i][0], ins[i][1], 0.25, 1.3)
        st.synchronize()

    th = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for r, o in zip(ref, outs):
        assert torch.equal(r, o)
    fdbm.release_plans()
    again = fdbm.fft_combine(_lib.PPM_FFT_POISSON_CONVOLVE, ins[0][0], ins[0][1], 0.25, 1.3)
    torch.cuda.synchronize()
    assert torch.equal(again, ref[0])


def test_fields_match_oracle():
    c = case()
    st, dV = c["step"], c["step"] ** 3
    target = c["shape"]
    ox, oy, oz, rx, ry, rz = fdbm.pad_roll_shifts(c["tip_tot"], target)
    tip_tot = F.pad_roll(c["tip_tot"].astype(np.float64), target, (ox, oy, oz), (rx, ry, rz))
    tip_del = F.pad_roll(c["tip_del"].astype(np.float64), target, (ox, oy, oz), (rx, ry, rz))
    dev = D.require_gpu()
    g_tt = fdbm.pad_roll(torch.as_tensor(c["tip_tot"], device=dev), target, (ox, oy, oz), (rx, ry, rz))
    assert np.array_equal(g_tt.cpu().numpy(), tip_tot.astype(np.float32))
    assert np.array_equal(fdbm.flip3(g_tt).cpu().numpy(), F.flip3(tip_tot).astype(np.float32))
    # Pauli overlap and its scaling
    ov = fdbm.compute_pauli_overlap(c["rho_scf"], tip_tot.astype(np.float32), st)
    ov_o = F.pauli_overlap(c["rho_scf"], tip_tot.astype(np.float32), st)
    close(ov, ov_o, 2e-5)
    Ep = fdbm.scale_pauli_field(ov_o.astype(np.float32), st, 18.0, 0.9, return_grads=False)
    close(Ep, F.pauli_scale(ov_o.astype(np.float32), 18.0, 0.9), 2e-5)
    # Poisson, electrostatic convolution (rolled and unrolled tip), fused path
    V = fdbm.fft_poisson(c["rho_diff"], st)
    close(V, F.fft_poisson(c["rho_diff"], st), 2e-5)
    close(fdbm.compute_es_conv_field(V, tip_del.astype(np.float32), st, tip_rolled=True, return_grads=False),
          F.es_conv(V, tip_del, st, tip_rolled=True), 2e-5)
    close(fdbm.compute_es_conv_field(V, tip_del.astype(np.float32), st, tip_rolled=False, return_grads=False),
          F.es_conv(V, tip_del, st, tip_rolled=False), 2e-5)
    close(fdbm.compute_es_fused_field(c["rho_diff"], tip_del.astype(np.float32), st),
          F.es_fused(c["rho_diff"], tip_del, st), 2e-5)
    # dispersion (per voxel relative)
    w = fdbm._c6_atoms(c["pos"], c["Z"], None, 30.0, 1.0)
    Ev = fdbm.compute_dispersion_grid(c["pos"], c["Z"], c["origin"], st, target, return_grads=False)
    Evo = F.dispersion(w.astype(np.float64), c["origin"], st, target)
    assert (np.abs(Ev - Evo) <= 1e-5 * np.abs(Evo) + 1e-12).all()


def test_stage3_end_to_end_and_scan():
    c = case()
    st, target = c["step"], c["shape"]
    A, beta = 18.0, 0.9
    a = afm.AFMulator(use_morse=False)
    _, Ep, Ee, Ev, Ft = fdbm.stage3_fdbm_fields_fast(a, c["rho_scf"], c["rho_diff"], c["tip_tot"], c["tip_del"],
                                                     c["origin"], st, target, c["pos"], c["Z"], A, beta)
    def placed(raw):        # each tip is centred and rolled by its own |peak| (the reference's rule)
        ox, oy, oz, rx, ry, rz = fdbm.pad_roll_shifts(raw, target)
        return F.pad_roll(raw.astype(np.float64), target, (ox, oy, oz), (rx, ry, rz))
    tt, td = placed(c["tip_tot"]), placed(c["tip_del"])
    Ep_o = F.pauli_scale(F.pauli_overlap(c["rho_scf"], tt, st), A, beta)
    Ee_o = F.es_fused(c["rho_diff"], td, st)
    Ev_o = F.dispersion(fdbm._c6_atoms(c["pos"], c["Z"], None, 30.0, 1.0).astype(np.float64), c["origin"], st,
                        target)
    close(Ep, Ep_o, 2e-5)
    close(Ee, Ee_o, 2e-5)
    close(Ev, Ev_o, 1e-5)
    G_o = F.gradient_periodic(Ep_o + Ee_o + Ev_o, st)
    for ch in range(4):
        close(Ft[..., ch], G_o[..., ch], 3e-5)
    # the device grid installed by stage 3 relaxes exactly like the same grid uploaded from the host
    xs = np.linspace(-1.0, 1.0, 6).astype(np.float32)
    ys = np.linspace(-0.8, 0.8, 5).astype(np.float32)
    hs = (0.2 * np.arange(8) + 4.0).astype(np.float32)
    kw = dict(mol_z=0.0, K_LAT=0.03, K_RAD=20.0, bond_length=3.0)
    FE1, d1 = a.scan_fdbm(xs, ys, hs, **kw)
    b = afm.AFMulator(use_morse=False)
    b.setup_fdbm_grid(Ft, c["origin"], st)
    FE2, d2 = b.scan_fdbm(xs, ys, hs, **kw)
    assert np.array_equal(FE1, FE2) and np.array_equal(d1["dz"], d2["dz"])
    assert np.isfinite(FE1).all()


def test_gradient_cl_matches_periodic_central_difference():
    rng = np.random.default_rng(2)
    E = rng.normal(size=(10, 12, 14)).astype(np.float32)
    a = afm.AFMulator(use_morse=False)
    G = a.compute_gradient_cl(E, 0.3)
    close(G, F.gradient_periodic(E, 0.3), 1e-6)
    assert G.shape == (10, 12, 14, 4)


def test_errors():
    dev = D.require_gpu()
    x = torch.zeros((4, 4, 4), device=dev)
    with pytest.raises(ValueError):
        fdbm.fft_combine(_lib.PPM_FFT_CORRELATE, x, None, 0.1, 1.0)      # b required
    with pytest.raises(ValueError):
        fdbm.fft_combine(7, x, x, 0.1, 1.0)                              # unknown mode
    with pytest.raises(ValueError):
        fdbm.fft_poisson(np.zeros((4, 4, 4)), (0.1, 0.2, 0.1))            # anisotropic step
