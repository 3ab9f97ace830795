"""CPU oracle for the probe-particle AFM path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this module, and only as the checker (or, for ``cpu_baseline``, as
the timed CPU restatement).  The product package ``spammm_amd`` never imports
it; the product path fails loudly when its HIP library is missing.

ctypes wrapper over ``oracle/_build/libppm_oracle.so`` (built from
``oracle/ppm_oracle.cpp`` by ``oracle/Makefile``).  Every function restates the
reference (file:line cited in ``ppm_oracle.cpp`` and below).

Parity pins (what checks each oracle function against something other than its
own formula; the tests live in ``tests/test_oracle_*.py``):

* ``pair_fe`` (getLJ / getMorse / getCoulombAFM, kernels/AFM.cl:730-757):
  F = -grad E by central finite difference; single-atom closed forms (LJ and
  Morse minimum at r=R0 with E=-E0, Coulomb q/r); translation, permutation and
  charge-linearity invariants (tests/SPM/test_afm_morse.py:694-857);
  periodic images checked against explicitly replicated atom lists.
* ``grid_project`` (evalLJC_QZs_toImg / evalMorseC_QZs_toImg,
  kernels/AFM.cl:916-1056): equals ``pair_fe`` at the voxel positions
  (p0 + dA*ia + dB*ib + dC*ic) followed by the |E|>100 rescale; voxel layout
  x-fastest; FP32 vs FP64 agreement.
* ``interp`` (sampler_1 trilinear REPEAT, kernels/AFM.cl:49,100-105): exact
  on affine fields, reproduces voxel values at texel centres
  (tests/SPM/test_texel_centers.py:17-94), REPEAT wrap on every axis.  The
  reference's hardware filter weight precision is not pinned by any reference
  test: parity unpinned for that property only.
* ``relax_tilted`` / ``relax_strokes`` (kernels/AFM.cl:478-562, :345-394):
  equilibrium residual |F_total(PP)| < 1e-4 eV/A at every converged state
  (tests/SPM/test_afm_morse.py:907-941); harmonic-field closed form (linear
  force field => analytic equilibrium); on-axis symmetric pixel => zero
  lateral displacement; step-trace reproducibility; FP32 and FP64 modes agree
  at tight convergence.
* ``fe_in_strokes_tilted`` (kernels/AFM.cl:210-233): equals ``interp`` at the
  stroke positions, rotated into the tip frame.
* ``relax_direct`` (relaxStrokesTiltedMorseDirect, kernels/AFM.cl:1419-1511):
  residual |G(PP)| < 1e-4 at every converged state, recomputed independently
  (tests/SPM/test_afm_morse.py:1048-1089); equals ``relax_tilted``'s loop with the
  direct pair sum as sample (same code path); FP32 and FP64 modes agree at tight
  convergence.
* ``direct_vjp`` (kernels/AFM.cl:1711-2029): channel by channel against the
  finite-difference component construction of the reference's CPU64 VJP oracle
  (tests/SPM/test_afm_morse.py:1240-1309: J, dO/dq, dG/dtheta, dO/dtheta by central
  differences of ``pair_fe_d``), and against re-relaxed finite differences of
  ``relax_direct_d`` with a plateau over three step sizes
  (tests/SPM/test_afm_morse.py:1523-1582); gates checked on constructed states.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libppm_oracle.so")
_lib = None

_f = ctypes.POINTER(ctypes.c_float)
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_l = ctypes.POINTER(ctypes.c_long)


def build(force: bool = False) -> str:
    """Compile the oracle library with its Makefile (g++, OpenMP)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_float = ctypes.c_int, ctypes.c_float
        L.oracle_pair_fe_f64.argtypes = [c_int, _f, _f, c_int, c_int, _f, _f, c_int, c_int, _f, _f, _d, c_int, _d]
        L.oracle_pair_fe_f64d.argtypes = [c_int, _d, _d, c_int, c_int, _d, _d, c_int, c_int, _d, _d, _d, c_int, _d]
        for nm in ("oracle_grid_project_f64", "oracle_grid_project_f32"):
            getattr(L, nm).argtypes = [c_int, _f, _f, c_int, c_int, _f, _f, c_int, c_int, _f, _f, _i,
                                       _f, _f, _f, _f, c_int, _f]
        L.oracle_grid_voxels_f64.argtypes = [c_int, _f, _f, c_int, c_int, _f, _f, c_int, c_int, _f, _f, _i,
                                             _f, _f, _f, _f, c_int, _l, c_int, _d]
        L.oracle_interp_f64.argtypes = [_f, _i, _f, _f, _f, _d, c_int, _d]
        for nm in ("oracle_relax_tilted_f64", "oracle_relax_tilted_f32"):
            getattr(L, nm).argtypes = [_f, _i, _f, _f, _f, _f, c_int, c_int, _f, _f, _f, _f, _f, _f, _f,
                                       c_int, c_float, c_int, _f, _f, c_int, c_int]
        for nm in ("oracle_relax_strokes_f64", "oracle_relax_strokes_f32"):
            getattr(L, nm).argtypes = [_f, _i, _f, _f, _f, _f, c_int, c_int, _f, _f, _f, _f, c_int, c_float,
                                       _f, _f, _f, c_int]
        for nm in ("oracle_relax_2d_f64", "oracle_relax_2d_f32"):
            getattr(L, nm).argtypes = [_f, _i, _f, _f, _f, _f, c_int, c_int, c_float, c_float, c_float, c_float,
                                       c_int, c_float, _f, _f, _f]
        for nm in ("oracle_fe_in_strokes_tilted_f64", "oracle_fe_in_strokes_tilted_f32"):
            getattr(L, nm).argtypes = [_f, _i, _f, _f, _f, _f, c_int, c_int, _f, _f, _f, _f, _f, _f]
        L.oracle_grid_project_f64g.argtypes = [c_int, _d, _d, c_int, c_int, _d, _d, c_int, c_int, _d, _d, _i,
                                               _d, _d, _d, _d, c_int, _d]
        L.oracle_relax_tilted_f64g.argtypes = [_d, _i, _d, _d, _d, _d, c_int, c_int, _d, _d, _d, _d, _d, _d, _d,
                                               c_int, ctypes.c_double, c_int, _d, _d, c_int]
        for nm in ("oracle_relax_direct_f64", "oracle_relax_direct_f32"):
            getattr(L, nm).argtypes = [_f, _f, c_int, _f, _f, _f, c_int, c_int, _f, _f, _f, _f, _f, _f, _f,
                                       c_int, c_float, _f, _f]
        L.oracle_relax_direct_f64d.argtypes = [_d, _d, c_int, _d, _d, _d, c_int, c_int, _d, _d, _d, _d, _d, _d, _d,
                                               c_int, ctypes.c_double, _d, _d]
        L.oracle_direct_vjp_f64.argtypes = [_d, _d, c_int, _d, _d, _d, c_int, c_int, _d, _d, _d, _d, _d, _d,
                                            _d, _d, _d, _d, _d]
        _lib = L
    return _lib


def _cf(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if n is not None:
        a = np.ascontiguousarray(np.resize(a.ravel(), n).astype(np.float32)) if a.size != n else a.ravel()
    return a


def _cd(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a, t):
    return a.ctypes.data_as(t)


def _vec(a, n):
    out = np.zeros(n, dtype=np.float32)
    a = np.asarray(a, dtype=np.float32).ravel()
    out[: min(n, a.size)] = a[: min(n, a.size)]
    return out


def pair_fe(model, atoms, cpar, queries, Qs, QZs, npbc=(0, 0), lvA=(0, 0, 0), lvB=(0, 0, 0)):
    """Direct sample FE (Fx,Fy,Fz,E) at points, FP64, uncapped.  model 0=LJ, 1=Morse."""
    atoms = _cf(atoms)
    cpar = _cf(cpar)
    na = atoms.shape[0]
    cs = cpar.shape[1] if cpar.ndim == 2 else 0
    q = _cd(queries).reshape(-1, 3)
    out = np.zeros((q.shape[0], 4), dtype=np.float64)
    Qs, QZs, lvA, lvB = _vec(Qs, 4), _vec(QZs, 4), _vec(lvA, 3), _vec(lvB, 3)
    lib().oracle_pair_fe_f64(int(model), _p(atoms, _f), _p(cpar, _f), cs, na, _p(Qs, _f), _p(QZs, _f),
                             int(npbc[0]), int(npbc[1]), _p(lvA, _f), _p(lvB, _f), _p(q, _d), q.shape[0],
                             _p(out, _d))
    return out


def pair_fe_d(model, atoms, cpar, queries, Qs, QZs, npbc=(0, 0), lvA=(0, 0, 0), lvB=(0, 0, 0)):
    """As ``pair_fe`` with double-precision atoms/params (for finite differences)."""
    atoms = _cd(atoms)
    cpar = _cd(cpar)
    na = atoms.shape[0]
    cs = cpar.shape[1]
    q = _cd(queries).reshape(-1, 3)
    out = np.zeros((q.shape[0], 4), dtype=np.float64)
    Qs = _cd(np.resize(np.asarray(Qs, float), 4))
    QZs = _cd(np.resize(np.asarray(QZs, float), 4))
    lvA = _cd(np.asarray(lvA, float).reshape(3))
    lvB = _cd(np.asarray(lvB, float).reshape(3))
    lib().oracle_pair_fe_f64d(int(model), _p(atoms, _d), _p(cpar, _d), cs, na, _p(Qs, _d), _p(QZs, _d),
                              int(npbc[0]), int(npbc[1]), _p(lvA, _d), _p(lvB, _d), _p(q, _d), q.shape[0],
                              _p(out, _d))
    return out


def grid_project(model, atoms, cpar, n, p0, dA, dB, dC, Qs, QZs, npbc=(0, 0), lvA=(0, 0, 0), lvB=(0, 0, 0),
                 cap=True, precision="f64"):
    """Full grid (nz,ny,nx,4) float32 output, computed in ``precision``."""
    atoms = _cf(atoms)
    cpar = _cf(cpar)
    cs = cpar.shape[1]
    nn = np.asarray(n, dtype=np.int32).ravel()[:3].copy()
    out = np.zeros((int(nn[2]), int(nn[1]), int(nn[0]), 4), dtype=np.float32)
    args = [_vec(x, 3) for x in (p0, dA, dB, dC)]
    Qs, QZs, lvA, lvB = _vec(Qs, 4), _vec(QZs, 4), _vec(lvA, 3), _vec(lvB, 3)
    fn = lib().oracle_grid_project_f64 if precision == "f64" else lib().oracle_grid_project_f32
    fn(int(model), _p(atoms, _f), _p(cpar, _f), cs, atoms.shape[0], _p(Qs, _f), _p(QZs, _f), int(npbc[0]),
       int(npbc[1]), _p(lvA, _f), _p(lvB, _f), _p(nn, _i), *[_p(a, _f) for a in args], int(bool(cap)),
       _p(out, _f))
    return out


def grid_voxels(model, atoms, cpar, n, p0, dA, dB, dC, Qs, QZs, idx, npbc=(0, 0), lvA=(0, 0, 0),
                lvB=(0, 0, 0), cap=True):
    """FP64 values (k,4) of the grid at flat voxel indices ``idx``."""
    atoms = _cf(atoms)
    cpar = _cf(cpar)
    cs = cpar.shape[1]
    nn = np.asarray(n, dtype=np.int32).ravel()[:3].copy()
    idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
    out = np.zeros((idx.size, 4), dtype=np.float64)
    args = [_vec(x, 3) for x in (p0, dA, dB, dC)]
    Qs, QZs, lvA, lvB = _vec(Qs, 4), _vec(QZs, 4), _vec(lvA, 3), _vec(lvB, 3)
    lib().oracle_grid_voxels_f64(int(model), _p(atoms, _f), _p(cpar, _f), cs, atoms.shape[0], _p(Qs, _f),
                                 _p(QZs, _f), int(npbc[0]), int(npbc[1]), _p(lvA, _f), _p(lvB, _f), _p(nn, _i),
                                 *[_p(a, _f) for a in args], int(bool(cap)), _p(idx, _l), idx.size, _p(out, _d))
    return out


def _grid_args(grid, n, dinvA, dinvB, dinvC):
    grid = _cf(grid)
    nn = np.asarray(n, dtype=np.int32).ravel()[:3].copy()
    assert grid.size == 4 * int(np.prod(nn.astype(np.int64))), "grid size does not match n"
    return grid, nn, _vec(dinvA, 4), _vec(dinvB, 4), _vec(dinvC, 4)


def interp(grid, n, dinvA, dinvB, dinvC, pos):
    g, nn, a, b, c = _grid_args(grid, n, dinvA, dinvB, dinvC)
    pos = _cd(pos).reshape(-1, 3)
    out = np.zeros((pos.shape[0], 4), dtype=np.float64)
    lib().oracle_interp_f64(_p(g, _f), _p(nn, _i), _p(a, _f), _p(b, _f), _p(c, _f), _p(pos, _d), pos.shape[0],
                            _p(out, _d))
    return out


def relax_tilted(grid, n, dinvA, dinvB, dinvC, pts, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF,
                 max_iter=128, f2conv=1e-8, check_before=False, precision="f64", damped=False, f32_anchors=False):
    """Returns (FEs (n_scan,nz,4), PPs (n_scan,nz,4)) float32.  ``damped``: the damped-velocity
    integrator of the OPT_FIRE=0 build (kernels/AFM.cl:536-540) instead of update_FIRE.
    ``f32_anchors``: the tip anchors advance in float32 as in the kernel (kernels/AFM.cl:556) also in
    the FP64 mode (the FP32 scan's exact heights)."""
    g, nn, a, b, c = _grid_args(grid, n, dinvA, dinvB, dinvC)
    pts = _cf(pts).reshape(-1, 4)
    ns = pts.shape[0]
    FEs = np.zeros((ns, nz, 4), dtype=np.float32)
    PPs = np.zeros((ns, nz, 4), dtype=np.float32)
    vs = [_vec(x, 4) for x in (tipA, tipB, tipC, stiffness, dpos0, relax, surfFF)]
    fn = lib().oracle_relax_tilted_f64 if precision == "f64" else lib().oracle_relax_tilted_f32
    fn(_p(g, _f), _p(nn, _i), _p(a, _f), _p(b, _f), _p(c, _f), _p(pts, _f), ns, int(nz),
       *[_p(v, _f) for v in vs], int(max_iter), float(f2conv), int(bool(check_before)), _p(FEs, _f), _p(PPs, _f),
       int(bool(damped)), int(bool(f32_anchors)))
    return FEs, PPs


def relax_strokes(grid, n, dinvA, dinvB, dinvC, pts, nz, dTip, stiffness, dpos0, relax, max_iter=128,
                  f2conv=1e-8, precision="f64", damped=False):
    """Returns (FEs, disps, PPs), each (n_scan,nz,4) float32.  ``damped``: the damped-velocity
    integrator of the OPT_FIRE=0 build (kernels/AFM.cl:378-382) instead of update_FIRE."""
    g, nn, a, b, c = _grid_args(grid, n, dinvA, dinvB, dinvC)
    pts = _cf(pts).reshape(-1, 4)
    ns = pts.shape[0]
    FEs = np.zeros((ns, nz, 4), dtype=np.float32)
    disps = np.zeros((ns, nz, 4), dtype=np.float32)
    PPs = np.zeros((ns, nz, 4), dtype=np.float32)
    vs = [_vec(x, 4) for x in (dTip, stiffness, dpos0, relax)]
    fn = lib().oracle_relax_strokes_f64 if precision == "f64" else lib().oracle_relax_strokes_f32
    fn(_p(g, _f), _p(nn, _i), _p(a, _f), _p(b, _f), _p(c, _f), _p(pts, _f), ns, int(nz), *[_p(v, _f) for v in vs],
       int(max_iter), float(f2conv), _p(FEs, _f), _p(disps, _f), _p(PPs, _f), int(bool(damped)))
    return FEs, disps, PPs


def relax_2d(grid, n, dinvA, dinvB, dinvC, pts, nz, k_lat, dh, dt, damp, max_iter=128, f2conv=1e-8,
             precision="f64"):
    """relaxStrokes2D (kernels/AFM.cl:396-450): returns (FEs, disps, PPs), each (n_scan,nz,4) float32."""
    g, nn, a, b, c = _grid_args(grid, n, dinvA, dinvB, dinvC)
    pts = _cf(pts).reshape(-1, 4)
    ns = pts.shape[0]
    FEs = np.zeros((ns, nz, 4), dtype=np.float32)
    disps = np.zeros((ns, nz, 4), dtype=np.float32)
    PPs = np.zeros((ns, nz, 4), dtype=np.float32)
    fn = lib().oracle_relax_2d_f64 if precision == "f64" else lib().oracle_relax_2d_f32
    fn(_p(g, _f), _p(nn, _i), _p(a, _f), _p(b, _f), _p(c, _f), _p(pts, _f), ns, int(nz), float(k_lat), float(dh),
       float(dt), float(damp), int(max_iter), float(f2conv), _p(FEs, _f), _p(disps, _f), _p(PPs, _f))
    return FEs, disps, PPs


def fe_in_strokes_tilted(grid, n, dinvA, dinvB, dinvC, pts, nz, tipA, tipB, tipC, dTip, dpos0, precision="f64"):
    g, nn, a, b, c = _grid_args(grid, n, dinvA, dinvB, dinvC)
    pts = _cf(pts).reshape(-1, 4)
    ns = pts.shape[0]
    FEs = np.zeros((ns, nz, 4), dtype=np.float32)
    vs = [_vec(x, 4) for x in (tipA, tipB, tipC, dTip, dpos0)]
    fn = lib().oracle_fe_in_strokes_tilted_f64 if precision == "f64" else lib().oracle_fe_in_strokes_tilted_f32
    fn(_p(g, _f), _p(nn, _i), _p(a, _f), _p(b, _f), _p(c, _f), _p(pts, _f), ns, int(nz), *[_p(v, _f) for v in vs],
       _p(FEs, _f))
    return FEs


# ── FP64 end to end (double inputs, double grid) ─────────────────────────────────

def _dvec(a, n):
    out = np.zeros(n, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64).ravel()
    out[: min(n, a.size)] = a[: min(n, a.size)]
    return out


def grid_project_d(model, atoms, cpar, n, p0, dA, dB, dC, Qs, QZs, npbc=(0, 0), lvA=(0, 0, 0), lvB=(0, 0, 0),
                   cap=True):
    """Grid (nz,ny,nx,4) float64 from float64 atoms/params (every operation in double)."""
    atoms = _cd(atoms).reshape(-1, 4)
    cpar = _cd(cpar)
    nn = np.asarray(n, dtype=np.int32).ravel()[:3].copy()
    out = np.zeros((int(nn[2]), int(nn[1]), int(nn[0]), 4), dtype=np.float64)
    g = [_dvec(x, 3) for x in (p0, dA, dB, dC)]
    Qs, QZs, lvA, lvB = _dvec(Qs, 4), _dvec(QZs, 4), _dvec(lvA, 3), _dvec(lvB, 3)
    lib().oracle_grid_project_f64g(int(model), _p(atoms, _d), _p(cpar, _d), cpar.shape[1], atoms.shape[0],
                                   _p(Qs, _d), _p(QZs, _d), int(npbc[0]), int(npbc[1]), _p(lvA, _d), _p(lvB, _d),
                                   _p(nn, _i), *[_p(x, _d) for x in g], int(bool(cap)), _p(out, _d))
    return out


def relax_tilted_d(grid, n, dinvA, dinvB, dinvC, pts, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF,
                   max_iter=128, f2conv=1e-8, check_before=False, f32_anchors=False):
    """relaxStrokesTilted in double on a double grid: (FEs, PPs) float64 (n_scan,nz,4);
    ``f32_anchors`` as in relax_tilted."""
    grid = _cd(grid)
    nn = np.asarray(n, dtype=np.int32).ravel()[:3].copy()
    assert grid.size == 4 * int(np.prod(nn.astype(np.int64)))
    pts = _cd(pts).reshape(-1, 4)
    ns = pts.shape[0]
    FEs = np.zeros((ns, nz, 4))
    PPs = np.zeros((ns, nz, 4))
    vs = [_dvec(x, 4) for x in (dinvA, dinvB, dinvC)]
    ts = [_dvec(x, 4) for x in (tipA, tipB, tipC, stiffness, dpos0, relax, surfFF)]
    lib().oracle_relax_tilted_f64g(_p(grid, _d), _p(nn, _i), *[_p(v, _d) for v in vs], _p(pts, _d), ns, int(nz),
                                   *[_p(t, _d) for t in ts], int(max_iter), float(f2conv), int(bool(check_before)),
                                   _p(FEs, _d), _p(PPs, _d), int(bool(f32_anchors)))
    return FEs, PPs


# ── gridless direct Morse path ────────────────────────────────────────────────

def relax_direct(atoms, cMs, Qs, QZs, pts, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF, max_iter=128,
                 f2conv=1e-8, precision="f64"):
    """relaxStrokesTiltedMorseDirect: (FEs, PPs) float32 (n_scan,nz,4); cMs rows (R0,E0,K,0)."""
    atoms = _cf(atoms).reshape(-1, 4)
    cMs = _cf(cMs).reshape(-1, 4)
    pts = _cf(pts).reshape(-1, 4)
    ns = pts.shape[0]
    FEs = np.zeros((ns, nz, 4), dtype=np.float32)
    PPs = np.zeros((ns, nz, 4), dtype=np.float32)
    q, qz = _vec(Qs, 4), _vec(QZs, 4)
    vs = [_vec(x, 4) for x in (tipA, tipB, tipC, stiffness, dpos0, relax, surfFF)]
    fn = lib().oracle_relax_direct_f64 if precision == "f64" else lib().oracle_relax_direct_f32
    fn(_p(atoms, _f), _p(cMs, _f), atoms.shape[0], _p(q, _f), _p(qz, _f), _p(pts, _f), ns, int(nz),
       *[_p(v, _f) for v in vs], int(max_iter), float(f2conv), _p(FEs, _f), _p(PPs, _f))
    return FEs, PPs


def relax_direct_d(atoms, cMs, Qs, QZs, pts, nz, tipA, tipB, tipC, stiffness, dpos0, relax, surfFF, max_iter=128,
                   f2conv=1e-8):
    """relax_direct in double end to end: (FEs, PPs) float64."""
    atoms = _cd(atoms).reshape(-1, 4)
    cMs = _cd(cMs).reshape(-1, 4)
    pts = _cd(pts).reshape(-1, 4)
    ns = pts.shape[0]
    FEs = np.zeros((ns, nz, 4))
    PPs = np.zeros((ns, nz, 4))
    q, qz = _dvec(Qs, 4), _dvec(QZs, 4)
    vs = [_dvec(x, 4) for x in (tipA, tipB, tipC, stiffness, dpos0, relax, surfFF)]
    lib().oracle_relax_direct_f64d(_p(atoms, _d), _p(cMs, _d), atoms.shape[0], _p(q, _d), _p(qz, _d), _p(pts, _d),
                                   ns, int(nz), *[_p(v, _d) for v in vs], int(max_iter), float(f2conv),
                                   _p(FEs, _d), _p(PPs, _d))
    return FEs, PPs


def direct_vjp(atoms, cMs, Qs, QZs, pts, nz, tipA, tipB, tipC, stiffness, dpos0, surfFF, PPs, dL_dFEs):
    """Implicit-equilibrium VJP of the direct path in double.

    Returns (grad (na,6) in theta order x,y,z,R0,E0,Q; lambdas (n_state,4); diag (n_state,4) =
    (residual, lambda_min, cond, status))."""
    atoms = _cd(atoms).reshape(-1, 4)
    cMs = _cd(cMs).reshape(-1, 4)
    pts = _cd(pts).reshape(-1, 4)
    ns = pts.shape[0]
    PPs = _cd(PPs).reshape(ns * nz, 4)
    U = _cd(dL_dFEs).reshape(ns * nz, 4)
    q, qz = _dvec(Qs, 4), _dvec(QZs, 4)
    vs = [_dvec(x, 4) for x in (tipA, tipB, tipC, stiffness, dpos0, surfFF)]
    grad = np.zeros((atoms.shape[0], 6))
    lam = np.zeros((ns * nz, 4))
    diag = np.zeros((ns * nz, 4))
    lib().oracle_direct_vjp_f64(_p(atoms, _d), _p(cMs, _d), atoms.shape[0], _p(q, _d), _p(qz, _d), _p(pts, _d), ns,
                                int(nz), *[_p(v, _d) for v in vs], _p(PPs, _d), _p(U, _d), _p(lam, _d),
                                _p(diag, _d), _p(grad, _d))
    return grad, lam, diag


def abs_term_sums(atoms, cLJ, q_tip, npbc, lvA, lvB, P):
    """Per query point (k,3), the FP64 sum over atoms and x,y images of |E| and |F| of the individual
    LJ and Coulomb (tip charge q_tip at z offset 0) pair terms: the scale of an FP32 sum's rounding
    where terms cancel (a bound helper for the parity tests, not a restatement of a reference
    function)."""
    P = np.asarray(P, np.float64)
    S = np.zeros(len(P))
    qs = float(q_tip) * 14.3996448915
    c6, c12 = np.asarray(cLJ, np.float64)[:, 0], np.asarray(cLJ, np.float64)[:, 1]
    q = np.asarray(atoms, np.float64)[:, 3]
    for kx in range(-npbc[0], npbc[0] + 1):
        for ky in range(-npbc[1], npbc[1] + 1):
            sh = kx * np.asarray(lvA, np.float64) + ky * np.asarray(lvB, np.float64)
            A = np.asarray(atoms, np.float64)[:, :3] + sh
            d = P[:, None, :] - A[None, :, :]
            r2 = (d * d).sum(2) + 1e-4
            ir2 = 1.0 / r2
            ir6 = ir2 ** 3
            ec = np.abs(q * qs) * np.sqrt(ir2)
            S += (np.abs((c12 * ir6 - c6) * ir6) + np.abs((12 * c12 * ir6 - 6 * c6) * ir6 * ir2) * np.sqrt(r2)
                  + ec + ec * np.sqrt(ir2)).sum(1)
    return S
