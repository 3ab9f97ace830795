"""Pins of the oracle's trilinear sampling and FIRE relaxation (CPU only).

Sampling (interpFE + sampler_1, kernels/AFM.cl:49,100-105):
  * texel-centre mapping of the half-texel dinv (tests/SPM/test_texel_centers.py:17-94):
    voxel i at p0 + i*step maps to (i+0.5)/n and samples exactly the voxel value;
  * exactness on affine fields inside the grid; REPEAT wrap on every axis.
Relaxation (relaxStrokesTilted, kernels/AFM.cl:478-562; relaxStrokes :345-394):
  * analytic equilibrium in a uniform force field (harmonic springs => closed form);
  * lateral symmetry: a pixel on a mirror plane of the field has zero lateral displacement;
  * equilibrium residual < 1e-4 eV/A at converged states (tests/SPM/test_afm_morse.py:907-941);
  * step-trace reproducibility (tests/SPM/test_afm_morse.py:944-961);
  * FP32 and FP64 restatements agree at tight convergence.
"""
import numpy as np
import pytest

from oracle import oracle as O
from spammm_amd import device as D   # host-side geometry helpers only (no GPU use)


def affine_grid(n, p0, L, coef):
    """Grid whose 4 channels are affine in position: ch_q = coef[q] . (x,y,z,1)."""
    nx, ny, nz = n
    step = np.asarray(L) / np.asarray(n)
    ia, ib, ic = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = np.stack([p0[0] + ia * step[0], p0[1] + ib * step[1], p0[2] + ic * step[2], np.ones_like(ia, float)], -1)
    G = pos @ np.asarray(coef).T          # (nx,ny,nz,4)
    return np.ascontiguousarray(G.transpose(2, 1, 0, 3).astype(np.float32))


def test_dinv_half_texel_mapping():
    L, origin, n = (10.0, 20.0, 30.0), np.array([-5.0, 3.0, 1.0]), (50, 100, 150)
    a, b, c = D.dinv_axis_aligned(L, origin, n)
    for i in [0, 1, 25, 49]:
        pos = origin[0] + i * L[0] / n[0]
        assert abs(pos * a[0] + a[3] - (i + 0.5) / n[0]) < 1e-5
    assert abs(b[3] - (-origin[1] / L[1] + 0.5 / n[1])) < 1e-6
    assert abs(c[3] - (-origin[2] / L[2] + 0.5 / n[2])) < 1e-6
    assert a[1] == a[2] == 0 and b[0] == b[2] == 0 and c[0] == c[1] == 0
    A2, B2, C2 = D.dinv_lvec(np.eye(3, dtype=np.float32), n)
    assert abs(A2[3] - 0.5 / n[0]) < 1e-7 and abs(C2[3] - 0.5 / n[2]) < 1e-7


def test_interp_reproduces_voxels_and_affine_fields():
    n, p0, L = (8, 7, 9), np.array([-2.0, -1.0, 0.5]), np.array([4.0, 3.5, 4.5])
    coef = np.array([[0.3, -0.2, 0.7, 0.1], [1.0, 0.0, 0.0, -0.5], [0.0, 0.5, -1.0, 2.0], [0.2, 0.2, 0.2, 0.2]])
    G = affine_grid(n, p0, L, coef)
    a, b, c = D.dinv_axis_aligned(L, p0, n)
    step = L / np.array(n)
    # texel centres reproduce the stored values
    idx = np.array([[0, 0, 0], [3, 2, 5], [7, 6, 8]])
    pos = p0 + idx * step
    vals = O.interp(G, n, a, b, c, pos)
    for k, (i, j, m) in enumerate(idx):
        np.testing.assert_allclose(vals[k], G[m, j, i], rtol=1e-5, atol=1e-5)
    # interior points of an affine field are exact up to float32 grid/coordinate rounding
    rng = np.random.default_rng(0)
    pos = p0 + rng.uniform(0.0, 1.0, (200, 3)) * (np.array(n) - 1) * step
    vals = O.interp(G, n, a, b, c, pos)
    ref = np.concatenate([pos, np.ones((200, 1))], 1) @ coef.T
    np.testing.assert_allclose(vals, ref, rtol=0, atol=2e-5)


def test_interp_repeat_wraps_every_axis():
    n, p0, L = (6, 5, 4), np.zeros(3), np.array([3.0, 2.5, 2.0])
    G = np.random.default_rng(1).normal(size=(4, 5, 6, 4)).astype(np.float32)
    a, b, c = D.dinv_axis_aligned(L, p0, n)
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 1, (50, 3)) * L
    base = O.interp(G, n, a, b, c, pos)
    for ax in range(3):
        sh = np.zeros(3)
        sh[ax] = L[ax]
        np.testing.assert_allclose(O.interp(G, n, a, b, c, pos + sh), base, rtol=1e-4, atol=1e-4)
        np.testing.assert_allclose(O.interp(G, n, a, b, c, pos - 2 * sh), base, rtol=1e-4, atol=1e-4)
    # half a step beyond the last voxel = mean of last and first voxel (wrap)
    step = L / np.array(n)
    p = np.array([[(n[0] - 0.5) * step[0], 0.0, 0.0]])
    np.testing.assert_allclose(O.interp(G, n, a, b, c, p)[0], 0.5 * (G[0, 0, -1] + G[0, 0, 0]), rtol=1e-5, atol=1e-6)


def uniform_field_case(F0, k_lat=0.05, k_rad=1.5, L=4.0):
    n, p0, Lc = (16, 16, 32), np.array([-4.0, -4.0, -8.0]), np.array([8.0, 8.0, 16.0])
    coef = np.zeros((4, 4))
    coef[0, 3], coef[1, 3], coef[2, 3] = F0
    G = affine_grid(n, p0, Lc, coef)
    a, b, c = D.dinv_axis_aligned(Lc, p0, n)
    st = np.array([-k_lat, -k_lat, -k_lat, -k_rad], np.float32)
    d0 = np.array([0, 0, -L, L], np.float32)
    return G, n, (a, b, c), st, d0


def analytic_equilibrium(F0, k_lat, k_rad, L):
    """Solve F0 + k*(d - d0) + k_r*d*(|d|-L)/|d| = 0 (stiffness negative) for d by Newton in FP64."""
    from scipy.optimize import fsolve
    k = np.array([-k_lat] * 3)
    d0 = np.array([0, 0, -L])

    def g(d):
        r = np.linalg.norm(d)
        return F0 + (d - d0) * k + d * (-k_rad * (r - L) / r)
    return fsolve(g, d0 + 0.01, xtol=1e-14)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_uniform_field_equilibrium_closed_form(precision):
    F0 = np.array([0.01, -0.004, 0.02])
    G, n, dinv, st, d0 = uniform_field_case(F0)
    pts = np.array([[0.1, -0.2, 4.0, 0], [0.3, 0.25, 3.5, 0]], np.float32)
    tipC = np.array([0, 0, 1, -0.1], np.float32)
    FEs, PPs = O.relax_tilted(G, n, *dinv, pts, 5, [1, 0, 0, 0], [0, 1, 0, 0], tipC, st, d0, [0.5, 0.1, 0, 0],
                              [0, 0, 0, 0], max_iter=20000, f2conv=1e-20 if precision == "f64" else 1e-12,
                              check_before=True, precision=precision)
    dref = analytic_equilibrium(F0, 0.05, 1.5, 4.0)
    assert (PPs[..., 3] > 0).all()
    for i in range(2):
        for iz in range(5):
            tip = pts[i, :3] + np.array([0, 0, -0.1]) * iz
            tol = 1e-6 if precision == "f64" else 3e-5
            np.testing.assert_allclose(PPs[i, iz, :3] - tip, dref, atol=tol)
            np.testing.assert_allclose(FEs[i, iz, :3], F0, atol=1e-6)


def test_f32_anchors_follow_the_float32_tip_accumulation():
    """relax_tilted_d with f32_anchors: in a uniform field the relaxed probe sits at the closed-form
    offset from anchors accumulated in float32 (tip += dTip, kernels/AFM.cl:556); without it, from the
    exact anchors p0 + iz*dtip.  Over 60 slices of -0.1 A the two anchor sequences differ by more than
    the 1e-9 A tolerance, so each mode is told apart."""
    F0 = np.array([0.01, -0.004, 0.02])
    G, n, dinv, st, d0 = uniform_field_case(F0)
    pts = np.array([[0.1, -0.2, 6.3, 0]], np.float32)
    nz = 60
    tipC = np.array([0, 0, 1, -0.1], np.float32)
    dref = analytic_equilibrium(F0.astype(np.float32).astype(np.float64), float(np.float32(0.05)), 1.5, 4.0)  # FP32 inputs
    anch32 = np.zeros((nz, 3), np.float32)
    cur = pts[0, :3].copy()
    for iz in range(nz):
        anch32[iz] = cur
        cur = cur + tipC[:3] * tipC[3]
    anch64 = pts[0, :3].astype(np.float64) + np.outer(np.arange(nz), [0, 0, float(tipC[3])])
    assert np.abs(anch32 - anch64).max() > 1e-7
    print(f"anchor drift of the float32 accumulation: {np.abs(anch32 - anch64).max():.2e} A")
    for f32, anch in ((True, anch32.astype(np.float64)), (False, anch64)):
        FEs, PPs = O.relax_tilted_d(G.astype(np.float64), n, *dinv, pts, nz, [1, 0, 0, 0], [0, 1, 0, 0], tipC, st, d0,
                                    [0.5, 0.1, 0, 0], [0, 0, 0, 0], max_iter=20000, f2conv=1e-26, check_before=True,
                                    f32_anchors=f32)
        assert (PPs[..., 3] > 0).all()
        np.testing.assert_allclose(PPs[0, :, :3] - anch, np.tile(dref, (nz, 1)), rtol=0, atol=1e-9)


def test_mirror_symmetric_pixel_has_no_lateral_displacement():
    """One LJ atom on the z axis, pixel on the axis: x,y of the probe stay exactly on the axis."""
    from oracle.oracle import grid_project
    atoms = np.array([[0.0, 0.0, 0.0, 0.2]], np.float32)
    R0, E0 = 3.5, 0.005
    c = np.array([[2 * E0 * R0 ** 6, E0 * R0 ** 12]], np.float32)
    n, p0, L = (32, 32, 40), np.array([-4.0, -4.0, -2.0]), np.array([8.0, 8.0, 10.0])
    step = L / np.array(n)
    G = grid_project(0, atoms, c, n, p0, [step[0], 0, 0], [0, step[1], 0], [0, 0, step[2]], [-0.05, 0, 0, 0],
                     [0, 0, 0, 0])
    dinv = D.dinv_axis_aligned(L, p0, n)
    pts = np.array([[0.0, 0.0, 8.5, 0]], np.float32)
    st = np.array([-0.03, -0.03, -0.03, -1.0], np.float32)
    d0 = np.array([0, 0, -4, 4], np.float32)
    for prec in ("f64", "f32"):
        FEs, PPs = O.relax_tilted(G, n, *dinv, pts, 20, [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, -0.1], st, d0,
                                  [0.5, 0.1, 0, 0], [0, 0, 0, 0], precision=prec)
        assert np.abs(PPs[0, :, :2]).max() < 1e-6
        assert np.abs(FEs[0, :, :2]).max() < 1e-5


def small_case(seed=7):
    from spammm_amd import workloads as W
    w = W.small(seed)
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    G = O.grid_project(0, w.atoms, w.cLJ, w.n, sp.p0, sp.dA, sp.dB, sp.dC, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB)
    pts = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    return w, sp, G, pts


def residual(w, sp, G, pts, PPs):
    pp = PPs[..., :3].reshape(-1, 3).astype(np.float64)
    fe = O.interp(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pp)
    tip = (pts[:, None, :3] + np.arange(w.nz)[None, :, None] * np.array([0, 0, w.dtip])).reshape(-1, 3)
    d = pp - tip
    r = np.linalg.norm(d, axis=1)
    k = w.stiffness.astype(np.float64)
    d0 = w.dpos0.astype(np.float64)
    f = fe[:, :3] + (d - d0[:3]) * k[:3] + d * (k[3] * (r - d0[3]) / r)[:, None]
    return np.linalg.norm(f, axis=1).reshape(PPs.shape[:2])


def test_relax_residual_below_tolerance():
    w, sp, G, pts = small_case()
    FEs, PPs = O.relax_tilted(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, [1, 0, 0, 0], [0, 1, 0, 0],
                              [0, 0, 1, w.dtip], w.stiffness, w.dpos0, [0.5, 0.1, 0, 0], [0, 0, 0, 0],
                              max_iter=w.max_iter, f2conv=w.f2conv, check_before=True)
    conv = PPs[..., 3] > 0
    assert conv.mean() > 0.95
    res = residual(w, sp, G, pts, PPs)
    assert res[conv].max() < 1e-4 + 1e-6


def test_relax_step_trace_reproducible_and_fp32_fp64_agree():
    w, sp, G, pts = small_case(3)
    args = (G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, w.dtip],
            w.stiffness, w.dpos0, [0.5, 0.1, 0, 0], [0, 0, 0, 0])
    a = O.relax_tilted(*args, max_iter=8, precision="f32")
    b = O.relax_tilted(*args, max_iter=8, precision="f32")
    np.testing.assert_array_equal(a[1], b[1])
    F64, P64 = O.relax_tilted(*args, max_iter=4000, f2conv=1e-14, check_before=True, precision="f64")
    F32, P32 = O.relax_tilted(*args, max_iter=4000, f2conv=1e-12, check_before=True, precision="f32")
    conv = (P64[..., 3] > 0) & (P32[..., 3] > 0)
    assert conv.mean() > 0.98
    assert np.abs(P64[..., :3] - P32[..., :3]).max(-1)[conv].max() < 1e-4
    assert np.abs(F64[..., 2] - F32[..., 2])[conv].max() < 1e-4


def test_relax_strokes_matches_tilted_in_identity_frame():
    """relaxStrokes (check after the update, FE from the last loop sample) and relaxStrokesTilted in the
    identity frame reach the same equilibria; both stop when |F|^2 < f2conv."""
    w, sp, G, pts = small_case(5)
    dTip = np.array([0, 0, w.dtip, 0], np.float32)
    FEs, disps, PPs = O.relax_strokes(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, dTip, w.stiffness, w.dpos0,
                                      [0.5, 0.1, 0, 0], 4000, 1e-14)
    FEt, PPt = O.relax_tilted(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, w.nz, [1, 0, 0, 0], [0, 1, 0, 0],
                              [0, 0, 1, w.dtip], w.stiffness, w.dpos0, [0.5, 0.1, 0, 0], [0, 0, 0, 0], 4000, 1e-14)
    conv = (PPs[..., 3] > 0) & (PPt[..., 3] > 0)
    assert conv.mean() > 0.98
    assert np.abs(PPs[..., :3] - PPt[..., :3]).max(-1)[conv].max() < 1e-5
    tip = pts[:, None, :3] + np.arange(w.nz)[None, :, None] * dTip[:3]
    np.testing.assert_allclose(disps[..., :3], PPs[..., :3] - (tip + w.dpos0[:3]), atol=1e-5)


def test_fe_in_strokes_equals_interp_along_stroke():
    w, sp, G, pts = small_case()
    A, B = np.array([1, 0, 0, 0.]), np.array([0, 1, 0, 0.])
    C = np.array([0, 0, 1, 0.])
    dTip = np.array([0, 0, -0.2, 0], np.float32)
    FEs = O.fe_in_strokes_tilted(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts, 6, A, B, C, dTip, w.dpos0)
    pos = pts[:, None, :3] + w.dpos0[:3] + np.arange(6)[None, :, None] * dTip[:3]
    ref = O.interp(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pos.reshape(-1, 3)).reshape(FEs.shape)
    np.testing.assert_allclose(FEs, ref, rtol=1e-6, atol=1e-6)
