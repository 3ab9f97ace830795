"""Pins of the oracle's damped-velocity relaxations (CPU only).

The reference builds its relaxation kernels with OPT_FIRE=0 for AFMulator(use_fire=False)
(spammm/SPM/AFM.py:320-321): ``v *= (1 - damp); v += f dt; pos += v dt`` with constant dt, damp
(kernels/AFM.cl:378-382, :536-540), and runs relaxStrokes2D for scan_fdbm_2d (kernels/AFM.cl:396-450,
spammm/SPM/AFM.py:644-701).  In a force field that is affine in the probe position and a linear
spring, every component of that iteration is the linear recurrence

    (e, v)_{n+1} = M (e, v)_n,   M = [[1 - k dt^2, (1 - d) dt], [-k dt, 1 - d]],

for the distance e to the analytic equilibrium (k = spring stiffness minus the field's gradient).
The tests below take positions, iteration counts (stop at the first force below the threshold,
tested after the update) and outputs from powers of M, not from the oracle's loop.
"""
import numpy as np
import pytest

from oracle import oracle as O
from spammm_amd import device as D   # host-side geometry helpers only (no GPU use)

from test_oracle_relax import affine_grid


def recurrence(e0, k, dt, damp, f2conv, max_iter, f_of_e):
    """Iterates (e, v) by M per component (arrays over components) until |f(e_n)|^2 < f2conv, tested
    after the update as the kernels do; returns (e after the last step, iterations, converged,
    e at the start of the last iteration)."""
    e = np.array(e0, dtype=np.float64)
    v = np.zeros_like(e)
    M = [np.array([[1.0 - kk * dt * dt, (1.0 - damp) * dt], [-kk * dt, 1.0 - damp]]) for kk in k]
    for it in range(max_iter):
        f = f_of_e(e)
        e_start = e.copy()
        ev = np.array([M[c] @ np.array([e[c], v[c]]) for c in range(len(e))])
        e, v = ev[:, 0], ev[:, 1]
        if float(np.sum(f * f)) < f2conv:
            return e, it + 1, True, e_start
    return e, max_iter, False, e_start


def same_count(got, want, precision):
    """Iteration counts: exact in FP64; FP32 sampling rounds the probe coordinate (~5e-8 eV/A of force
    on a 1e-5 residual), which can move the stop by one step."""
    return got == want if precision == "f64" else abs(got - want) <= 1


def _field(n, p0, L, Fx, Fy, Fz, E):
    coef = np.array([Fx, Fy, Fz, E], dtype=np.float64)
    return affine_grid(n, p0, L, coef), coef


@pytest.mark.parametrize("precision,tol", [("f64", 1e-6), ("f32", 2e-4)])
def test_damped_strokes_follow_the_linear_recurrence(precision, tol):
    """relaxStrokes with OPT_FIRE=0 in a uniform field and a linear spring (k_rad = 0, dpos0 = 0):
    per axis f = F0 - k (x - tip), so x* = tip + F0/k and e = x - x* obeys M; the next slice starts
    from the last position moved by dTip with v = 0."""
    n, p0, L = (16, 16, 32), np.array([-4.0, -4.0, -8.0]), np.array([8.0, 8.0, 16.0])
    F0 = np.array([0.012, -0.007, 0.02])
    G, coef = _field(n, p0, L, [0, 0, 0, F0[0]], [0, 0, 0, F0[1]], [0, 0, 0, F0[2]], [0, 0, 0, -0.3])
    a, b, c = D.dinv_axis_aligned(L, p0, n)
    k, dt, damp = 0.5, 0.3, 0.2
    st = np.array([-k, -k, -k, 0.0], np.float32)
    d0 = np.zeros(4, np.float32)
    relax = np.array([dt, damp, 0.03, 0.3], np.float32)
    dTip = np.array([0.0, 0.0, -0.1, 0.0], np.float32)
    pts = np.array([[0.1, -0.2, 4.0, 0.0], [-0.3, 0.25, 3.5, 0.0]], np.float32)
    nz, f2conv, max_iter = 5, 1e-10, 400
    FEs, disps, PPs = O.relax_strokes(G, n, a, b, c, pts, nz, dTip, st, d0, relax, max_iter, f2conv, precision,
                                      damped=True)
    kk = np.full(3, k)
    for ip in range(2):
        tip = pts[ip, :3].astype(np.float64)
        e = tip - (tip + F0 / k)                 # the probe starts at the tip (dpos0 = 0)
        for iz in range(nz):
            e, iters, conv, _ = recurrence(e, kk, dt, damp, f2conv, max_iter, lambda ee: -k * ee)
            assert conv and same_count(PPs[ip, iz, 3], iters, precision)
            pos = tip + F0 / k + e
            np.testing.assert_allclose(PPs[ip, iz, :3], pos, rtol=0, atol=tol)
            np.testing.assert_allclose(disps[ip, iz, :3], pos - tip, rtol=0, atol=tol)
            np.testing.assert_allclose(FEs[ip, iz], [*F0, -0.3], rtol=0, atol=1e-6)
            tip = tip + dTip[:3]
    # the same scan with FIRE takes other iteration counts (the integrator switch is live)
    _, _, PPf = O.relax_strokes(G, n, a, b, c, pts, nz, dTip, st, d0, relax, max_iter, f2conv, precision)
    assert not np.array_equal(PPf[..., 3], PPs[..., 3])


@pytest.mark.parametrize("precision,tol", [("f64", 1e-6), ("f32", 2e-4)])
def test_damped_tilted_follows_the_linear_recurrence(precision, tol):
    """relaxStrokesTilted with OPT_FIRE=0, identity frame, in a field affine in x (Fx = F0x + g x):
    the x component obeys M with k - g, y and z with k; FEs = the sample after the last update."""
    n, p0, L = (16, 16, 32), np.array([-4.0, -4.0, -8.0]), np.array([8.0, 8.0, 16.0])
    F0, g = np.array([0.012, -0.007, 0.02]), 0.1
    G, coef = _field(n, p0, L, [g, 0, 0, F0[0]], [0, 0, 0, F0[1]], [0, 0, 0, F0[2]], [0.1, 0, 0, -0.3])
    a, b, c = D.dinv_axis_aligned(L, p0, n)
    k, dt, damp = 0.5, 0.3, 0.2
    st = np.array([-k, -k, -k, 0.0], np.float32)
    d0 = np.zeros(4, np.float32)
    relax = np.array([dt, damp, 0.0, 0.0], np.float32)
    tipA, tipB = np.array([1, 0, 0, 0], np.float32), np.array([0, 1, 0, 0], np.float32)
    tipC = np.array([0, 0, 1, -0.1], np.float32)
    pts = np.array([[0.4, -0.2, 4.0, 0.0]], np.float32)
    nz, f2conv, max_iter = 4, 1e-10, 400
    FEs, PPs = O.relax_tilted(G, n, a, b, c, pts, nz, tipA, tipB, tipC, st, d0, relax, np.zeros(4), max_iter,
                              f2conv, False, precision, damped=True)
    keff = np.array([k - g, k, k])
    tip = pts[0, :3].astype(np.float64)
    for iz in range(nz):
        xs = np.array([(F0[0] + k * tip[0]) / (k - g), tip[1] + F0[1] / k, tip[2] + F0[2] / k])
        e = (PPs[0, iz - 1, :3] + tipC[:3] * tipC[3] - xs) if iz else tip - xs
        e, iters, conv, _ = recurrence(e, keff, dt, damp, f2conv, max_iter, lambda ee: -keff * ee)
        assert conv and same_count(PPs[0, iz, 3], iters, precision)
        pos = xs + e
        np.testing.assert_allclose(PPs[0, iz, :3], pos, rtol=0, atol=tol)
        np.testing.assert_allclose(FEs[0, iz, :3], [F0[0] + g * pos[0], F0[1], F0[2]], rtol=0, atol=10 * tol)
        np.testing.assert_allclose(FEs[0, iz, 3], -0.3 + 0.1 * pos[0], rtol=0, atol=10 * tol)
        tip = tip + tipC[:3] * tipC[3]


@pytest.mark.parametrize("precision,tol", [("f64", 1e-6), ("f32", 2e-4)])
def test_relax_2d_follows_the_linear_recurrence(precision, tol):
    """relaxStrokes2D in Fx = ax0 + gx x + cz z, Fy = ay0 + gy y: at height pz the lateral equilibrium
    is x* = (ax0 + cz pz + k ax)/(k - gx); each slice restarts from the anchor with v = 0; the
    probe height of slice iz is az - iz*dh; disps = (x - ax, y - ay, 0, 0)."""
    n, p0, L = (20, 20, 24), np.array([-5.0, -5.0, -4.0]), np.array([10.0, 10.0, 12.0])
    ax0, gx, cz, ay0, gy = 0.02, 0.08, -0.04, -0.015, -0.06
    G, coef = _field(n, p0, L, [gx, 0, cz, ax0], [0, gy, 0, ay0], [0, 0, 0.01, 0.05], [0.1, -0.1, 0.2, 0.3])
    a, b, c = D.dinv_axis_aligned(L, p0, n)
    k_lat, dh, dt, damp = 0.4, 0.2, 0.3, 0.2
    pts = np.array([[0.5, -0.4, 4.0, 0.0], [-1.0, 1.2, 3.0, 0.0]], np.float32)
    nz, f2conv, max_iter = 6, 1e-10, 600
    FEs, disps, PPs = O.relax_2d(G, n, a, b, c, pts, nz, k_lat, dh, dt, damp, max_iter, f2conv, precision)
    keff = np.array([k_lat - gx, k_lat - gy])
    for ip in range(2):
        ax, ay, az = pts[ip, :3].astype(np.float64)
        for iz in range(nz):
            pz = np.float32(az) - np.float32(iz) * np.float32(dh) if precision == "f32" else az - iz * dh
            xs = np.array([(ax0 + cz * pz + k_lat * ax) / (k_lat - gx), (ay0 + k_lat * ay) / (k_lat - gy)])
            e, iters, conv, _ = recurrence(np.array([ax, ay]) - xs, keff, dt, damp, f2conv, max_iter,
                                           lambda ee: -keff * ee)
            assert conv and same_count(PPs[ip, iz, 3], iters, precision)
            px, py = xs + e
            np.testing.assert_allclose(PPs[ip, iz, :3], [px, py, pz], rtol=0, atol=tol)
            np.testing.assert_allclose(disps[ip, iz], [px - ax, py - ay, 0.0, 0.0], rtol=0, atol=tol)
            ref = coef @ np.array([px, py, pz, 1.0])
            np.testing.assert_allclose(FEs[ip, iz], ref, rtol=0, atol=10 * tol)


def test_relax_2d_without_early_stop_is_the_fixed_step_loop():
    """With f2conv = 0 the loop never stops early: max_iter steps, PPs.w = -max_iter (not converged),
    and the position after N steps is M^N applied to the anchor's distance to the equilibrium (the
    N_RELAX-step loop of the CPU twin pp_relax_2d, spammm/SPM/AFM.py:3977-4029, with its
    v = 0.8 v + 0.3 F, x += 0.3 v, i.e. dt = 0.3, damp = 0.2)."""
    n, p0, L = (20, 20, 24), np.array([-5.0, -5.0, -4.0]), np.array([10.0, 10.0, 12.0])
    G, coef = _field(n, p0, L, [0.005, 0, 0, 0.02], [0, 0, 0, -0.01], [0, 0, 0, 0.05], [0, 0, 0, 0.3])
    a, b, c = D.dinv_axis_aligned(L, p0, n)
    k_lat, N = 0.03, 50
    pts = np.array([[0.5, -0.4, 4.0, 0.0]], np.float32)
    FEs, disps, PPs = O.relax_2d(G, n, a, b, c, pts, 2, k_lat, 0.1, 0.3, 0.2, N, 0.0, "f64")
    keff = np.array([k_lat - 0.005, k_lat])
    xs = np.array([(0.02 + k_lat * 0.5) / keff[0], (-0.01 + k_lat * -0.4) / keff[1]])
    M = [np.array([[1 - kk * 0.09, 0.8 * 0.3], [-kk * 0.3, 0.8]]) for kk in keff]
    e = [np.linalg.matrix_power(M[i], N) @ np.array([[0.5, -0.4][i] - xs[i], 0.0]) for i in range(2)]
    assert np.all(PPs[0, :, 3] == -N)
    np.testing.assert_allclose(PPs[0, 0, :2], xs + np.array([e[0][0], e[1][0]]), rtol=0, atol=1e-6)
