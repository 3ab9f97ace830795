"""The 2D Ewald oracle (oracle/ewald_ref.py) against independent facts.

* Above the slab the vacuum form (coefficients C_G e^{|G| z_i}) equals the full form (|z - z_i|).
* The full form equals the direct Coulomb sum over images inside growing circles (the reference's
  brute force, spammm/surfaces/Ewald2D.py:176-204), above the slab and between its layers, to 1e-9.
* In the vacuum the potential is harmonic (finite-difference Laplacian), periodic along a and b,
  and the analytic field equals -grad phi by central differences.
"""
import numpy as np

from oracle import ewald_ref as E
from spammm_amd import workloads as WL


def slab(n_harm=10):
    en, ions = WL.rocksalt_slab(cells=1, layers=2)
    return E.Slab(ions.astype(np.float64), (5.64, 0.0), (0.0, 5.64), n_harm)


P_ABOVE = np.array([[0.3, 0.7, 2.5], [1.1, 2.0, 3.5], [2.0, 0.5, 5.0]])


def test_vacuum_equals_full_above_the_slab():
    s = slab()
    assert np.allclose(s.phi_vacuum(P_ABOVE), s.phi_full(P_ABOVE), rtol=0, atol=1e-12)


def test_full_form_equals_brute_force_image_sum():
    s = slab(24)     # points between layers need more harmonics (|z - z_i| ~ 1.3 A)
    P = np.vstack([P_ABOVE, [[0.9, 1.7, -1.3], [2.2, 0.4, -1.6]]])     # two points between the layers
    d80 = np.abs(s.phi_brute(P, 80) - s.phi_full(P))
    d40 = np.abs(s.phi_brute(P, 40) - s.phi_full(P))
    assert (d80 < 1e-9).all() and (d80 < d40).all()


def test_harmonic_periodic_and_field():
    s = slab()
    h = 1e-3
    lap = -6 * s.phi_full(P_ABOVE)
    for ax in range(3):
        e = np.zeros(3); e[ax] = h
        lap += s.phi_full(P_ABOVE + e) + s.phi_full(P_ABOVE - e)
    assert np.abs(lap / h ** 2).max() < 1e-4 * np.abs(s.field_full(P_ABOVE)).max()
    assert np.allclose(s.phi_full(P_ABOVE + [5.64, 0, 0]), s.phi_full(P_ABOVE), atol=1e-12)
    assert np.allclose(s.phi_full(P_ABOVE + [0, 5.64, 0]), s.phi_full(P_ABOVE), atol=1e-12)
    fd = np.zeros((len(P_ABOVE), 3))
    for ax in range(3):
        e = np.zeros(3); e[ax] = 1e-5
        fd[:, ax] = -(s.phi_full(P_ABOVE + e) - s.phi_full(P_ABOVE - e)) / 2e-5
    assert np.allclose(s.field_full(P_ABOVE), fd, atol=1e-7 * np.abs(fd).max() + 1e-9)
