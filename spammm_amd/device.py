"""Device-tensor wrappers over the C ABI (include/ppm_api.h).

Every function takes and returns float32 CUDA (HIP) tensors, enqueues on the
current torch stream of the tensor's device and returns without synchronizing.
There is no CPU path: a missing GPU or library raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib


def require_gpu(device=None):
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("spammm_amd needs a HIP GPU (torch.cuda.is_available() is False); "
                           "the AFM path has no CPU fallback")
    _lib.load()
    return torch.device(device if device is not None else "cuda:0")


def _stream(t):
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


class Stream:
    """A non-blocking HIP stream of the library's own (ppm_stream_create) and its torch handle
    ``.torch`` (torch.cuda.ExternalStream).  Each one takes the least used hardware queue when it is
    created, so streams created one after another run side by side; torch's pooled streams can
    share a queue.  Buffers used on it are allocated on the caller's stream and must stay alive
    until that stream has waited for this one (close() makes the current stream wait, then releases
    the stream; so does garbage collection)."""

    def __init__(self, device=None):
        import torch
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(_lib.load().ppm_stream_create(ctypes.byref(h)))
        self.handle = h.value
        self.torch = torch.cuda.ExternalStream(self.handle, device=dev)

    def close(self):
        if self.handle:
            import torch
            torch.cuda.current_stream(self.torch.device).wait_stream(self.torch)
            h, self.handle = self.handle, None
            _lib.check(_lib.load().ppm_stream_destroy(ctypes.c_void_p(h)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32dev(t, name, shape_last=None):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
        raise ValueError(f"{name} must be a float32 CUDA tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape_last is not None and t.shape[-1] != shape_last:
        raise ValueError(f"{name} must have last dimension {shape_last}, got {tuple(t.shape)}")
    return t


@dataclass
class GridSpec:
    """Grid geometry: voxel (ia,ib,ic) at p0 + dA*ia + dB*ib + dC*ic; sampling rows dinvA/B/C."""
    n: tuple
    p0: np.ndarray
    dA: np.ndarray
    dB: np.ndarray
    dC: np.ndarray
    dinvA: np.ndarray
    dinvB: np.ndarray
    dinvC: np.ndarray

    def c(self) -> _lib.PpmGridSpec:
        s = _lib.PpmGridSpec()
        s.nx, s.ny, s.nz = (int(x) for x in self.n)
        s.reserved = 0
        s.p0, s.dA, s.dB, s.dC = (_lib.f4(v) for v in (self.p0, self.dA, self.dB, self.dC))
        s.dinvA, s.dinvB, s.dinvC = (_lib.f4(v) for v in (self.dinvA, self.dinvB, self.dinvC))
        return s

    @property
    def nvox(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])


def dinv_axis_aligned(L, origin, n):
    """(dinvA, dinvB, dinvC) with the half-texel offset (AFMulator._make_dinv_axis_aligned,
    spammm/SPM/AFM.py:275-285): dot((pos,1), dinvA) = (pos.x - origin.x)/L.x + 0.5/nx."""
    nx, ny, nz = n
    dinvA = np.array([1. / L[0], 0., 0., -origin[0] / L[0] + 0.5 / nx], dtype=np.float32)
    dinvB = np.array([0., 1. / L[1], 0., -origin[1] / L[1] + 0.5 / ny], dtype=np.float32)
    dinvC = np.array([0., 0., 1. / L[2], -origin[2] / L[2] + 0.5 / nz], dtype=np.float32)
    return dinvA, dinvB, dinvC


def dinv_lvec(invMT, n):
    """(dinvA, dinvB, dinvC) of a skew cell (AFMulator._make_dinv_lvec, spammm/SPM/AFM.py:287-294)."""
    nx, ny, nz = n
    dinvA = np.array([invMT[0, 0], invMT[0, 1], invMT[0, 2], 0.5 / nx], dtype=np.float32)
    dinvB = np.array([invMT[1, 0], invMT[1, 1], invMT[1, 2], 0.5 / ny], dtype=np.float32)
    dinvC = np.array([invMT[2, 0], invMT[2, 1], invMT[2, 2], 0.5 / nz], dtype=np.float32)
    return dinvA, dinvB, dinvC


def axis_aligned_spec(p0, L, n) -> GridSpec:
    p0 = np.asarray(p0, dtype=np.float32).reshape(3)
    L = np.asarray(L, dtype=np.float32).reshape(3)
    nx, ny, nz = (int(v) for v in n)
    dA = np.array([L[0] / nx, 0., 0.], dtype=np.float32)
    dB = np.array([0., L[1] / ny, 0.], dtype=np.float32)
    dC = np.array([0., 0., L[2] / nz], dtype=np.float32)
    a, b, c = dinv_axis_aligned(L, p0, (nx, ny, nz))
    return GridSpec((nx, ny, nz), p0, dA, dB, dC, a, b, c)


@dataclass
class Tip:
    """Tip / probe parameter record (AFMulator.DEFAULT_*, spammm/SPM/AFM.py:296-304)."""
    tipA: np.ndarray = field(default_factory=lambda: np.array([1., 0., 0., 0.], np.float32))
    tipB: np.ndarray = field(default_factory=lambda: np.array([0., 1., 0., 0.], np.float32))
    tipC: np.ndarray = field(default_factory=lambda: np.array([0., 0., 1., -0.1], np.float32))
    stiffness: np.ndarray = field(default_factory=lambda: np.array([-0.03, -0.03, -0.03, -1.0], np.float32))
    dpos0: np.ndarray = field(default_factory=lambda: np.array([0., 0., -4.0, 4.0], np.float32))
    relax: np.ndarray = field(default_factory=lambda: np.array([0.5, 0.1, 0.02, 0.5], np.float32))
    surfFF: np.ndarray = field(default_factory=lambda: np.zeros(4, np.float32))

    def c(self) -> _lib.PpmTip:
        t = _lib.PpmTip()
        t.tipA, t.tipB, t.tipC = _lib.f4(self.tipA), _lib.f4(self.tipB), _lib.f4(self.tipC)
        t.stiffness, t.dpos0 = _lib.f4(self.stiffness), _lib.f4(self.dpos0)
        t.relax, t.surfFF = _lib.f4(self.relax), _lib.f4(self.surfFF)
        return t


def relax_ctl(max_iter=128, f2conv=1e-8, check_before=False, scan_ny=0, f32_anchors=False,
              damped=False, lanes=0) -> _lib.PpmRelaxCtl:
    """Relaxation control record; ``damped``: the damped-velocity integrator of the reference's
    OPT_FIRE=0 build instead of FIRE; ``lanes``: lanes per scan point (1, 2, 4; 0 = the library's
    choice for a launch that has the GPU to itself; results do not depend on it)."""
    c = _lib.PpmRelaxCtl()
    c.max_iter = int(max_iter)
    c.f2conv = float(f2conv)
    c.check_before = int(bool(check_before))
    c.scan_ny = int(scan_ny)
    c.f32_anchors = int(bool(f32_anchors))
    c.damped = int(bool(damped))
    c.lanes = int(lanes)
    return c


def grid_project(atoms, cpar, model: int, spec: GridSpec, Qs, QZs, npbc=(0, 0), lvA=(0, 0, 0), lvB=(0, 0, 0),
                 cap=True, out=None, base=None, acc64=False, window=None):
    """Force-field grid (nz,ny,nx,4) on the device (evalLJC_QZs_toImg / evalMorseC_QZs_toImg).
    ``base``: a raw (cap=False) grid of other atoms on the same spec, added before the cap.
    ``acc64``: sum each voxel's FP32 pair terms in FP64 (ppm_grid_project_acc64; no base).
    ``window``: (klo, nzw) -> only planes [klo, klo + nzw), returned as (nzw,ny,nx,4)
    (ppm_grid_project_window; no base, no acc64)."""
    import torch
    _f32dev(atoms, "atoms", 4)
    _f32dev(cpar, "cpar", 2 if model == _lib.PPM_MODEL_LJ else 4)
    if atoms.shape[0] != cpar.shape[0]:
        raise ValueError("atoms and cpar must have the same number of rows")
    nx, ny, nz = (int(v) for v in spec.n)
    if window is not None:
        if base is not None or acc64:
            raise ValueError("a z window takes no base grid and no acc64")
        nz = int(window[1])
    if out is None:
        out = torch.empty((nz, ny, nx, 4), dtype=torch.float32, device=atoms.device)
    else:
        _f32dev(out, "out", 4)
        if out.numel() != 4 * nx * ny * nz:
            raise ValueError("out has the wrong size")
    L = _lib.load()
    lva = (ctypes.c_float * 3)(*[float(x) for x in np.asarray(lvA, float).reshape(3)])
    lvb = (ctypes.c_float * 3)(*[float(x) for x in np.asarray(lvB, float).reshape(3)])
    s = spec.c()
    args = (int(model), atoms.data_ptr(), cpar.data_ptr(), int(atoms.shape[0]), ctypes.byref(s), _lib.f4(Qs),
            _lib.f4(QZs), int(npbc[0]), int(npbc[1]), ctypes.byref(lva), ctypes.byref(lvb), int(bool(cap)))
    if window is not None:
        rc = L.ppm_grid_project_window(*args[:5], int(window[0]), int(window[1]), *args[5:], out.data_ptr(),
                                       _stream(atoms))
    elif acc64:
        if base is not None:
            raise ValueError("acc64 takes no base grid")
        rc = L.ppm_grid_project_acc64(*args, out.data_ptr(), _stream(atoms))
    elif base is None:
        rc = L.ppm_grid_project(*args, out.data_ptr(), _stream(atoms))
    else:
        _f32dev(base, "base", 4)
        if base.numel() != out.numel():
            raise ValueError("base has the wrong size")
        rc = L.ppm_grid_project_add(*args, base.data_ptr(), out.data_ptr(), _stream(atoms))
    _lib.check(rc)
    return out


def _f64dev(t, name, last):
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float64 device tensor")
    if t.dim() < 1 or t.shape[-1] != last:
        raise ValueError(f"{name} must have a last dimension of {last}")


def grid_project_f64(atoms, cpar, model: int, spec: GridSpec, Qs, QZs, npbc=(0, 0), lvA=(0, 0, 0), lvB=(0, 0, 0),
                     cap=True, out=None):
    """The force-field grid (nz,ny,nx,4) in FP64 from FP64 atoms / parameters (reference precision,
    e.g. for finite-difference gradient checks); same laws as `grid_project`."""
    import torch
    _f64dev(atoms, "atoms", 4)
    _f64dev(cpar, "cpar", 2 if model == _lib.PPM_MODEL_LJ else 4)
    if atoms.shape[0] != cpar.shape[0]:
        raise ValueError("atoms and cpar must have the same number of rows")
    nx, ny, nz = (int(v) for v in spec.n)
    if out is None:
        out = torch.empty((nz, ny, nx, 4), dtype=torch.float64, device=atoms.device)
    else:
        _f64dev(out, "out", 4)
        if out.numel() != 4 * nx * ny * nz:
            raise ValueError("out has the wrong size")
    L = _lib.load()
    D3, D4 = ctypes.c_double * 3, ctypes.c_double * 4
    d4 = lambda v: D4(*([float(x) for x in list(np.asarray(v, float).ravel())[:4]] + [0.0] * 4)[:4])  # noqa: E731
    lva = D3(*[float(x) for x in np.asarray(lvA, float).reshape(3)])
    lvb = D3(*[float(x) for x in np.asarray(lvB, float).reshape(3)])
    s = spec.c()
    rc = L.ppm_grid_project_f64(int(model), atoms.data_ptr(), cpar.data_ptr(), int(atoms.shape[0]), ctypes.byref(s),
                                d4(Qs), d4(QZs), int(npbc[0]), int(npbc[1]), ctypes.byref(lva), ctypes.byref(lvb),
                                int(bool(cap)), out.data_ptr(), _stream(atoms))
    _lib.check(rc)
    return out


def relax_tilted(grid, spec: GridSpec, points, nz: int, tip: Tip, max_iter=128, f2conv=1e-8, check_before=False,
                 scan_ny=0, FEs=None, PPs=None, want_pps=True, damped=False, window=None, flag=None, df_op=None,
                 df=None, lanes=0):
    """relaxStrokesTilted on the device: returns (FEs (n_scan,nz,4), PPs (n_scan,nz,4) or None).
    ``window``: (klo, nzw) -> ``grid`` holds only planes [klo, klo + nzw) of the spec's grid
    (ppm_relax_strokes_tilted_window); ``flag`` (int32 device tensor, zeroed by the caller) turns
    nonzero if a probe sampled outside the window.  ``df_op`` (a df.AxisOperator over nz slices) with
    ``df`` ((n_scan, n_out) float32 device tensor): the operator applied to each Fz column in the
    same launch (ppm_relax_strokes_tilted_df).  ``lanes``: lanes per scan point (relax_ctl)."""
    import torch
    _f32dev(grid, "grid", 4)
    _f32dev(points, "points", 4)
    nplanes = spec.n[2] if window is None else int(window[1])
    if grid.numel() != 4 * spec.n[0] * spec.n[1] * nplanes:
        raise ValueError("grid size does not match spec")
    if window is not None and (flag is None or flag.dtype != torch.int32 or not flag.is_cuda):
        raise ValueError("a windowed relaxation needs an int32 device flag")
    ns = int(points.shape[0])
    if FEs is None:
        FEs = torch.empty((ns, nz, 4), dtype=torch.float32, device=grid.device)
    if PPs is None and want_pps:
        PPs = torch.empty((ns, nz, 4), dtype=torch.float32, device=grid.device)
    for nm, t in (("FEs", FEs), ("PPs", PPs)):
        if t is not None:
            _f32dev(t, nm, 4)
            if t.numel() != ns * nz * 4:
                raise ValueError(f"{nm} has the wrong size")
    L = _lib.load()
    s, tp, ctl = spec.c(), tip.c(), relax_ctl(max_iter, f2conv, check_before, scan_ny, damped=damped, lanes=lanes)
    if df_op is not None:
        if window is not None:
            raise ValueError("the fused df takes no z window")
        _f32dev(df, "df", None)
        if df.numel() != ns * int(df_op.n_out):
            raise ValueError("df must be (n_scan, n_out)")
        rc = L.ppm_relax_strokes_tilted_df(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz),
                                           ctypes.byref(tp), ctypes.byref(ctl), FEs.data_ptr(),
                                           PPs.data_ptr() if PPs is not None else None, df_op.W.data_ptr(),
                                           df_op.start.data_ptr(), int(df_op.taps), int(df_op.n_out), df.data_ptr(),
                                           _stream(grid))
    elif window is not None:
        rc = L.ppm_relax_strokes_tilted_window(grid.data_ptr(), ctypes.byref(s), int(window[0]), int(window[1]),
                                               points.data_ptr(), ns, int(nz), ctypes.byref(tp), ctypes.byref(ctl),
                                               FEs.data_ptr(), PPs.data_ptr() if PPs is not None else None,
                                               flag.data_ptr(), _stream(grid))
    else:
        rc = L.ppm_relax_strokes_tilted(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz),
                                        ctypes.byref(tp), ctypes.byref(ctl), FEs.data_ptr(),
                                        PPs.data_ptr() if PPs is not None else None, _stream(grid))
    _lib.check(rc)
    return FEs, PPs


def relax_tilted_f64(grid, spec: GridSpec, points, nz: int, tip: Tip, max_iter=128, f2conv=1e-8,
                     check_before=False, want_pps=True, f32_anchors=False, damped=False):
    """`relax_tilted` in FP64 on an FP64 grid and FP64 points (reference precision): returns
    (FEs (n_scan,nz,4), PPs (n_scan,nz,4) or None) as float64 device tensors.  ``f32_anchors``:
    the tip anchors advance in float32 as in the FP32 scan (whose heights drift from p0 + iz*dz by
    up to ~1e-5 A over tens of slices), so the FP64 scan relaxes at the FP32 scan's exact heights."""
    import torch
    _f64dev(grid, "grid", 4)
    _f64dev(points, "points", 4)
    if grid.numel() != 4 * spec.nvox:
        raise ValueError("grid size does not match spec")
    ns = int(points.shape[0])
    FEs = torch.empty((ns, nz, 4), dtype=torch.float64, device=grid.device)
    PPs = torch.empty((ns, nz, 4), dtype=torch.float64, device=grid.device) if want_pps else None
    L = _lib.load()
    s, tp, ctl = spec.c(), tip.c(), relax_ctl(max_iter, f2conv, check_before, 0, f32_anchors, damped)
    rc = L.ppm_relax_strokes_tilted_f64(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz),
                                        ctypes.byref(tp), ctypes.byref(ctl), FEs.data_ptr(),
                                        PPs.data_ptr() if PPs is not None else None, _stream(grid))
    _lib.check(rc)
    return FEs, PPs


def relax_strokes(grid, spec: GridSpec, points, nz: int, dTip, stiffness, dpos0, relax, max_iter=128, f2conv=1e-8,
                  scan_ny=0, want_disps=True, want_pps=True, damped=False):
    """relaxStrokes on the device: returns (FEs, disps or None, PPs or None), each (n_scan,nz,4).
    ``damped``: damped-velocity integrator (the reference's OPT_FIRE=0 build) instead of FIRE."""
    import torch
    _f32dev(grid, "grid", 4)
    _f32dev(points, "points", 4)
    if grid.numel() != 4 * spec.nvox:
        raise ValueError("grid size does not match spec")
    ns = int(points.shape[0])
    dev = grid.device
    FEs = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev)
    disps = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev) if want_disps else None
    PPs = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev) if want_pps else None
    L = _lib.load()
    s, ctl = spec.c(), relax_ctl(max_iter, f2conv, False, scan_ny, damped=damped)
    rc = L.ppm_relax_strokes(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz), _lib.f4(dTip),
                             _lib.f4(stiffness), _lib.f4(dpos0), _lib.f4(relax), ctypes.byref(ctl), FEs.data_ptr(),
                             disps.data_ptr() if disps is not None else None,
                             PPs.data_ptr() if PPs is not None else None, _stream(grid))
    _lib.check(rc)
    return FEs, disps, PPs


def relax_strokes_2d(grid, spec: GridSpec, points, nz: int, k_lat, dh, dt, damp, max_iter=128, f2conv=1e-8,
                     scan_ny=0, want_disps=True, want_pps=False):
    """relaxStrokes2D on the device (lateral-only damped relaxation at fixed heights az - iz*dh):
    returns (FEs, disps or None, PPs or None), each (n_scan,nz,4)."""
    import torch
    _f32dev(grid, "grid", 4)
    _f32dev(points, "points", 4)
    if grid.numel() != 4 * spec.nvox:
        raise ValueError("grid size does not match spec")
    ns = int(points.shape[0])
    dev = grid.device
    FEs = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev)
    disps = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev) if want_disps else None
    PPs = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev) if want_pps else None
    L = _lib.load()
    s, ctl = spec.c(), relax_ctl(max_iter, f2conv, False, scan_ny)
    rc = L.ppm_relax_strokes_2d(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz), float(k_lat),
                                float(dh), float(dt), float(damp), ctypes.byref(ctl), FEs.data_ptr(),
                                disps.data_ptr() if disps is not None else None,
                                PPs.data_ptr() if PPs is not None else None, _stream(grid))
    _lib.check(rc)
    return FEs, disps, PPs


def fe_in_strokes_tilted(grid, spec: GridSpec, points, nz: int, tip: Tip, dTip, dpos0):
    """getFEinStrokesTilted on the device: (n_scan,nz,4)."""
    import torch
    _f32dev(grid, "grid", 4)
    _f32dev(points, "points", 4)
    if grid.numel() != 4 * spec.nvox:
        raise ValueError("grid size does not match spec")
    ns = int(points.shape[0])
    FEs = torch.empty((ns, nz, 4), dtype=torch.float32, device=grid.device)
    L = _lib.load()
    s, tp = spec.c(), tip.c()
    rc = L.ppm_fe_in_strokes_tilted(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz),
                                    ctypes.byref(tp), _lib.f4(dTip), _lib.f4(dpos0), FEs.data_ptr(), _stream(grid))
    _lib.check(rc)
    return FEs


def relax_adjoint(grid, spec: GridSpec, points, nz: int, tip: Tip, PPs, dL_dFEs, grid_cot=None, lam=None):
    """Implicit-function adjoint of relaxed states: accumulates dL/dgrid into ``grid_cot`` (zeros if None).
    Returns (grid_cot, lambdas (n_scan,nz,4) with the status word in .w; into ``lam`` if given)."""
    import torch
    _f32dev(grid, "grid", 4)
    _f32dev(points, "points", 4)
    _f32dev(PPs, "PPs", 4)
    _f32dev(dL_dFEs, "dL_dFEs", 4)
    ns = int(points.shape[0])
    if PPs.numel() != ns * nz * 4 or dL_dFEs.numel() != ns * nz * 4:
        raise ValueError("PPs / dL_dFEs must be (n_scan, nz, 4)")
    if grid_cot is None:
        grid_cot = torch.zeros_like(grid)
    else:
        _f32dev(grid_cot, "grid_cot", 4)
        if grid_cot.numel() != grid.numel():
            raise ValueError("grid_cot must have the grid's shape")
    if lam is None:
        lam = torch.empty((ns, nz, 4), dtype=torch.float32, device=grid.device)
    else:
        _f32dev(lam, "lam", 4)
        if lam.numel() != ns * nz * 4:
            raise ValueError("lam must be (n_scan, nz, 4)")
    L = _lib.load()
    s, tp = spec.c(), tip.c()
    rc = L.ppm_relax_adjoint(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz), ctypes.byref(tp),
                             PPs.data_ptr(), dL_dFEs.data_ptr(), grid_cot.data_ptr(), lam.data_ptr(), _stream(grid))
    _lib.check(rc)
    return grid_cot, lam


def relax_state_diag(grid, spec: GridSpec, points, nz: int, tip: Tip, PPs, out=None):
    """Per-state (|G|, smallest eigenvalue of sym(-J), condition number, status bits) of relaxed
    states, (n_scan, nz, 4) float32 (into ``out`` if given); status bits as relax_adjoint's."""
    import torch
    _f32dev(grid, "grid", 4)
    _f32dev(points, "points", 4)
    _f32dev(PPs, "PPs", 4)
    ns = int(points.shape[0])
    if PPs.numel() != ns * nz * 4:
        raise ValueError("PPs must be (n_scan, nz, 4)")
    if out is None:
        diag = torch.empty((ns, nz, 4), dtype=torch.float32, device=grid.device)
    else:
        _f32dev(out, "out", 4)
        if out.numel() != ns * nz * 4:
            raise ValueError("out must be (n_scan, nz, 4)")
        diag = out
    L = _lib.load()
    s, tp = spec.c(), tip.c()
    rc = L.ppm_relax_state_diag(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz), ctypes.byref(tp),
                                PPs.data_ptr(), diag.data_ptr(), _stream(grid))
    _lib.check(rc)
    return diag


def relax_polish(grid, spec: GridSpec, points, nz: int, tip: Tip, PPs, FEs=None, steps=3):
    """FP64 Newton polish of converged probe positions in place (and FEs resampled there)."""
    _f32dev(grid, "grid", 4)
    _f32dev(points, "points", 4)
    _f32dev(PPs, "PPs", 4)
    ns = int(points.shape[0])
    if PPs.numel() != ns * nz * 4 or (FEs is not None and FEs.numel() != ns * nz * 4):
        raise ValueError("PPs / FEs must be (n_scan, nz, 4)")
    if FEs is not None:
        _f32dev(FEs, "FEs", 4)
    L = _lib.load()
    s, tp = spec.c(), tip.c()
    rc = L.ppm_relax_polish(grid.data_ptr(), ctypes.byref(s), points.data_ptr(), ns, int(nz), ctypes.byref(tp),
                            int(steps), PPs.data_ptr(), FEs.data_ptr() if FEs is not None else None, _stream(grid))
    _lib.check(rc)
    return PPs, FEs


def grid_adjoint(atoms, cpar, model: int, spec: GridSpec, Qs, QZs, grid, grid_cot, npbc=(0, 0), lvA=(0, 0, 0),
                 lvB=(0, 0, 0), cap=True, workspace=None, base=None, params=True, out=None):
    """dL/d(x,y,z,q,p0,p1) per atom (na,6) float64 from dL/dgrid (p0,p1 = c6,c12 for LJ, R0,E0 for Morse),
    into ``out`` if given.  ``base``: the base grid the forward was projected over
    (grid_project(..., base=)).  ``params=False``: positions and charges only (columns 4, 5 are zero),
    less arithmetic."""
    import torch
    _f32dev(atoms, "atoms", 4)
    _f32dev(cpar, "cpar", 2 if model == _lib.PPM_MODEL_LJ else 4)
    _f32dev(grid, "grid", 4)
    _f32dev(grid_cot, "grid_cot", 4)
    if grid.numel() != 4 * spec.nvox or grid_cot.numel() != grid.numel():
        raise ValueError("grid / grid_cot size does not match spec")
    na = int(atoms.shape[0])
    L = _lib.load()
    s = spec.c()
    need = int(L.ppm_grid_adjoint_workspace_size(na, ctypes.byref(s)))
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=atoms.device)
    if out is None:
        grad = torch.empty((na, 6), dtype=torch.float64, device=atoms.device)
    else:
        if out.dtype != torch.float64 or not out.is_cuda or not out.is_contiguous() or out.numel() != na * 6:
            raise ValueError("out must be a contiguous (na, 6) float64 device tensor")
        grad = out
    lva = (ctypes.c_float * 3)(*[float(x) for x in np.asarray(lvA, float).reshape(3)])
    lvb = (ctypes.c_float * 3)(*[float(x) for x in np.asarray(lvB, float).reshape(3)])
    head = (int(model), atoms.data_ptr(), cpar.data_ptr(), na, ctypes.byref(s), _lib.f4(Qs), _lib.f4(QZs),
            int(npbc[0]), int(npbc[1]), ctypes.byref(lva), ctypes.byref(lvb), int(bool(cap)))
    tail = (grid.data_ptr(), grid_cot.data_ptr(), workspace.data_ptr(), workspace.numel(), int(bool(params)),
            grad.data_ptr(), _stream(atoms))
    if base is None:
        rc = L.ppm_grid_adjoint(*head, *tail)
    else:
        _f32dev(base, "base", 4)
        if base.numel() != grid.numel():
            raise ValueError("base has the wrong size")
        rc = L.ppm_grid_adjoint_add(*head, base.data_ptr(), *tail)
    _lib.check(rc)
    return grad


def relax_direct_morse(atoms, cMs, points, nz: int, tip: Tip, Qs, QZs, max_iter=128, f2conv=1e-8, scan_ny=0,
                       FEs=None, PPs=None, want_pps=True, damped=False):
    """Gridless direct Morse relaxed scan (relaxStrokesTiltedMorseDirect): (FEs, PPs or None), (n_scan,nz,4)."""
    import torch
    _f32dev(atoms, "atoms", 4)
    _f32dev(cMs, "cMs", 4)
    _f32dev(points, "points", 4)
    na, ns = int(atoms.shape[0]), int(points.shape[0])
    if cMs.shape[0] != na:
        raise ValueError("cMs must have one row per atom")
    if FEs is None:
        FEs = torch.empty((ns, nz, 4), dtype=torch.float32, device=atoms.device)
    if PPs is None and want_pps:
        PPs = torch.empty((ns, nz, 4), dtype=torch.float32, device=atoms.device)
    for nm, t in (("FEs", FEs), ("PPs", PPs)):
        if t is not None:
            _f32dev(t, nm, 4)
            if t.numel() != ns * nz * 4:
                raise ValueError(f"{nm} has the wrong size")
    L = _lib.load()
    tp, ctl = tip.c(), relax_ctl(max_iter, f2conv, True, scan_ny, damped=damped)
    rc = L.ppm_relax_direct_morse(atoms.data_ptr(), cMs.data_ptr(), na, points.data_ptr(), ns, int(nz),
                                  ctypes.byref(tp), _lib.f4(Qs), _lib.f4(QZs), ctypes.byref(ctl), FEs.data_ptr(),
                                  PPs.data_ptr() if PPs is not None else None, _stream(atoms))
    _lib.check(rc)
    return FEs, PPs


def direct_morse_vjp(atoms, cMs, points, nz: int, tip: Tip, Qs, QZs, PPs, dL_dFEs):
    """Implicit-equilibrium VJP of relax_direct_morse: (grad (na,6) float64 in theta order
    (x,y,z,R0,E0,Q), lambdas (n_scan,nz,4), diag (n_scan,nz,4) = (residual, lambda_min, cond, status))."""
    import torch
    for nm, t in (("atoms", atoms), ("cMs", cMs), ("points", points), ("PPs", PPs), ("dL_dFEs", dL_dFEs)):
        _f32dev(t, nm, 4)
    na, ns = int(atoms.shape[0]), int(points.shape[0])
    if PPs.numel() != ns * nz * 4 or dL_dFEs.numel() != ns * nz * 4:
        raise ValueError("PPs / dL_dFEs must have n_scan*nz float4 entries")
    L = _lib.load()
    ws_bytes = int(L.ppm_direct_morse_vjp_workspace_size(na, ns, int(nz)))
    if ws_bytes == 0:
        raise ValueError(f"invalid direct VJP sizes (na={na}, n_scan={ns}, nz={nz})")
    dev = atoms.device
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    lam = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev)
    diag = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev)
    grad = torch.empty((na, 6), dtype=torch.float64, device=dev)
    tp = tip.c()
    rc = L.ppm_direct_morse_vjp(atoms.data_ptr(), cMs.data_ptr(), na, points.data_ptr(), ns, int(nz),
                                ctypes.byref(tp), _lib.f4(Qs), _lib.f4(QZs), PPs.data_ptr(), dL_dFEs.data_ptr(),
                                lam.data_ptr(), diag.data_ptr(), ws.data_ptr(), ws_bytes, grad.data_ptr(),
                                _stream(atoms))
    _lib.check(rc)
    return grad, lam, diag


def scan_points(scan_p0, scan_da, scan_db, nx: int, ny: int) -> np.ndarray:
    """(nx*ny,4) float32 ix-major scan starts: p0 + da*ix + db*iy in float32, the arithmetic of the
    reference loop (spammm/SPM/AFM.py:977-980) and build_scan_xy_points_vectorized (:82-96)."""
    p0 = np.asarray(scan_p0, dtype=np.float32).reshape(3)
    da = np.asarray(scan_da, dtype=np.float32).reshape(3)
    db = np.asarray(scan_db, dtype=np.float32).reshape(3)
    ix = np.arange(int(nx), dtype=np.float32)[:, None, None]
    iy = np.arange(int(ny), dtype=np.float32)[None, :, None]
    xyz = (p0 + da * ix) + db * iy
    pts = np.zeros((int(nx) * int(ny), 4), dtype=np.float32)
    pts[:, :3] = xyz.reshape(-1, 3)
    return pts
