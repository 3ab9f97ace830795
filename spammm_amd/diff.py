"""Differentiable df image: a torch.autograd.Function over the HIP forward and adjoint.

Forward  (atoms xyz, charges q, per-atom LJ R0/E0) -> LJ (c6,c12) -> grid projection ->
         relaxed scan -> df operator along z  (all on the GPU).
Backward dL/ddf -> transposed df operator -> dL/dFEs (Fz channel) ->
         ppm_relax_adjoint (implicit-function theorem at each converged probe position,
         per-state 3x3 solve in FP64, trilinear-adjoint scatter into dL/dgrid) ->
         ppm_grid_adjoint (grid -> atom pair-law adjoint, FP64 reduction) ->
         chain (c6,c12) -> (R0,E0):  c6 = 2 E0 R0^6, c12 = E0 R0^12.
The relaxation trajectory is never stored: the backward needs only the converged probe
positions, which it either keeps from the forward (16 B per state) or recomputes by
re-running the deterministic relaxation (``recompute=True``).

There is no reference counterpart for the grid path (the reference differentiates only its
gridless direct-Morse path, spammm/SPM/AFM.py:2241-2519); the loss convention follows
df_loss_seed (spammm/SPM/AFM.py:3125-3200).

``direct_df_image`` is the same Function over the gridless direct Morse path:
ppm_relax_direct_morse forward, ppm_direct_morse_vjp backward (kernels/AFM.cl:1419-2029).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib, df as DF
from . import device as D


@dataclass
class Problem:
    """Everything of the forward that is not differentiated."""
    spec: D.GridSpec
    pts: object                 # (n_scan,4) float32 device tensor
    nxy: tuple
    nz: int
    tip: D.Tip
    Qs: np.ndarray
    QZs: np.ndarray
    npbc: tuple = (0, 0)
    lvA: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    lvB: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    cap: bool = True
    max_iter: int = 1000
    f2conv: float = 1e-12
    check_before: bool = True
    polish: int = 3             # FP64 Newton steps on each converged probe position (0: FIRE result as is)
    df_kind: str = "giessibl"
    amp: float = 1.0
    f0: float = 30000.0
    k_cant: float = 1800.0
    grid_acc64: bool = False    # grid voxel sums in FP64 (ppm_grid_project_acc64): the accuracy form
    recompute: bool = False     # backward re-runs the relaxation instead of keeping probe positions
    strict: bool = True         # raise if any state is non-finite in the adjoint (gated states are counted)
    # Column gating (off when gate_eig_frac <= 0): the forward flags each pixel column whose probe
    # descent comes near a fold of its equilibrium branch -- a state with smallest Hessian
    # eigenvalue below gate_eig_frac * k_lat, a jump of the probe (relative to the tip) of
    # gate_max_jump A or more between consecutive slices, or a state the adjoint gates -- where a
    # small parameter change moves the probe to another branch and the df column jumps.
    # last_columns: (nx, ny) bool device tensor, True = column kept; callers weight the loss by it.
    gate_eig_frac: float = 0.0
    gate_max_jump: float = 0.3
    last_columns: object = None
    last_gate: dict = field(default_factory=dict)
    last_status: dict = field(default_factory=dict)
    # Scan window (off when window_margin is None): the forward projects the grid only on the
    # sub-box of its voxels that holds every probe rest position widened by (lateral, vertical)
    # margins in A (poses.scan_window); a forward whose probes come within 0.5 A of a cut face is
    # recomputed on the whole grid (window_fallbacks counts them).  last_spec: the grid the last
    # forward ran on (its adjoint runs there too).
    window_margin: tuple = None
    window_fallbacks: int = 0
    last_spec: object = None
    _window: object = None

    @property
    def n_scan(self):
        return int(self.nxy[0]) * int(self.nxy[1])

    @property
    def dz(self):
        return float(self.tip.tipC[3])

    def run_spec(self):
        """(spec, faces) of the grid the forward runs on first: the scan window and its
        (lo, hi) shared-face flags, or (spec, None) without a window."""
        if self._window is None:
            self._window = False
            if self.window_margin is not None:
                from .poses import scan_window
                w = scan_window(self.spec, self.pts.cpu().numpy(), self.nz, self.tip, *self.window_margin)
                if w is not None:
                    self._window = (w[0], (w[1], w[2]))
        return self._window if self._window else (self.spec, None)

    @classmethod
    def from_workload(cls, w, device, **kw):
        import torch
        spec = D.axis_aligned_spec(w.p0, w.L, w.n)
        pts = torch.as_tensor(D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy), device=device)
        tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
        args = dict(spec=spec, pts=pts, nxy=tuple(w.nxy), nz=int(w.nz), tip=tip, Qs=w.Qs, QZs=w.QZs,
                    npbc=tuple(w.npbc), lvA=w.lvA, lvB=w.lvB, amp=w.amp, f0=w.f0, k_cant=w.k_cant)
        args.update(kw)
        return cls(**args)


def lj_cpar(R0, E0):
    """(na,2) float32 (c6, c12) = (2 E0 R0^6, E0 R0^12), evaluated in FP64."""
    import torch
    R0d, E0d = R0.double(), E0.double()
    r6 = R0d ** 6
    return torch.stack([2.0 * E0d * r6, E0d * r6 * r6], dim=1).float().contiguous()


def _relax(grid, prob: Problem, spec=None):
    spec = prob.spec if spec is None else spec
    FEs, PPs = D.relax_tilted(grid, spec, prob.pts, prob.nz, prob.tip, prob.max_iter, prob.f2conv,
                              prob.check_before, scan_ny=int(prob.nxy[1]))
    if prob.polish > 0:
        D.relax_polish(grid, spec, prob.pts, prob.nz, prob.tip, PPs, FEs, prob.polish)
    return FEs, PPs


def column_gate(grid, spec, pts, nxy, nz, tip, PPs, eig_frac, max_jump):
    """(nx, ny) bool device tensor of the pixel columns that stay on one equilibrium branch under
    small parameter changes -- every state passes the adjoint's gates, has a smallest Hessian
    eigenvalue (of the symmetric part of -J, ppm_relax_state_diag) of at least eig_frac * k_lat, and
    the probe moves by less than max_jump A relative to the tip between consecutive slices -- and a
    dict of counts."""
    keep, parts = column_gate_parts(grid, spec, pts, nxy, nz, tip, PPs, eig_frac, max_jump)
    return keep, gate_stats(parts, eig_frac, max_jump)


def column_gate_parts(grid, spec, pts, nxy, nz, tip, PPs, eig_frac, max_jump):
    """column_gate without host synchronisation, all on the current stream: ((nx, ny) bool keep,
    (n_scan,) bool device tensors (adjoint_gated, near_fold, jumped))."""
    dg = D.relax_state_diag(grid, spec, pts, nz, tip, PPs)
    keep, parts = gate_from_diag(dg, PPs, pts, nz, tip, eig_frac, max_jump)
    return keep.reshape(int(nxy[0]), int(nxy[1])), parts


def gate_from_diag(dg, PPs, pts, nz, tip, eig_frac, max_jump):
    """The column tests of column_gate from the state diagnostics dg and probe positions PPs, both
    (..., n_scan, nz, 4) (any leading image axes): ((..., n_scan) bool keep, (adjoint_gated,
    near_fold, jumped))."""
    import torch
    k_lat = float(np.min(np.abs(np.asarray(tip.stiffness[:3], np.float64))))
    dev = PPs.device
    iz = torch.arange(nz, device=dev, dtype=torch.float32)
    C = torch.as_tensor(np.asarray(tip.tipC[:3], np.float32) * np.float32(tip.tipC[3]), device=dev)
    rel = PPs[..., :3] - (pts[:, None, :3] + iz[None, :, None] * C)       # probe - tip anchor
    if nz > 1:
        jump = (rel[..., 1:, :] - rel[..., :-1, :]).norm(dim=-1).amax(dim=-1)
    else:
        jump = torch.zeros(PPs.shape[:-2], device=dev)
    gated = (dg[..., 3] != 0).any(dim=-1)
    soft = dg[..., 1].amin(dim=-1) < eig_frac * k_lat
    jumped = jump >= max_jump
    keep = ~(gated | soft | jumped)
    return keep, (gated, soft, jumped)


def gate_stats(parts, eig_frac, max_jump):
    """Counts of column_gate_parts (one or several images' parts stacked along a leading axis)."""
    gated, soft, jumped = parts
    keep = ~(gated | soft | jumped)
    return {"columns": int(keep.numel()), "kept": int(keep.sum()), "adjoint_gated": int(gated.sum()),
            "near_fold": int(soft.sum()), "jumped": int(jumped.sum()), "eig_frac": eig_frac, "max_jump_A": max_jump}


def stable_columns(grid, PPs, prob: Problem, spec=None):
    """column_gate for a Problem (its gate_eig_frac, gate_max_jump) on `grid` of `spec` (default:
    the grid of the last forward)."""
    if spec is None:
        spec = prob.last_spec if prob.last_spec is not None else prob.spec
    return column_gate(grid, spec, prob.pts, prob.nxy, prob.nz, prob.tip, PPs, prob.gate_eig_frac,
                       prob.gate_max_jump)


def masked_l2(df, target, columns=None):
    """1/2 sum w (df - target)^2 (float64) over pixel columns weighted by `columns` ((nx, ny) bool or
    float; None: all): the L2 loss of the differentiable image restricted to stable columns (the
    intersection of the forward's and the target's Problem.last_columns)."""
    r = df.double() - target.double().reshape(df.shape)
    if columns is not None:
        r = r * columns.reshape(*columns.shape, 1).to(r.dtype)
    return 0.5 * (r * r).sum()


def _forward_parts(atoms, cpar, prob: Problem):
    """(grid, FEs, PPs, df (nx, ny, nz), spec): the forward on the scan window, or on the whole
    grid without one or when the probes leave it; spec is the grid it ran on."""
    def run(spec):
        grid = D.grid_project(atoms, cpar, _lib.PPM_MODEL_LJ, spec, prob.Qs, prob.QZs, prob.npbc, prob.lvA,
                              prob.lvB, cap=prob.cap, acc64=prob.grid_acc64)
        return (grid,) + _relax(grid, prob, spec)

    spec, faces = prob.run_spec()
    grid, FEs, PPs = run(spec)
    if faces is not None:
        from .poses import window_inside
        if not window_inside(spec, faces, PPs):
            prob.window_fallbacks += 1
            spec = prob.spec
            grid, FEs, PPs = run(spec)
    prob.last_spec = spec
    if prob.gate_eig_frac > 0:
        prob.last_columns, prob.last_gate = stable_columns(grid, PPs, prob, spec)
    op = DF.z_operator(prob.df_kind, prob.nz, prob.dz, atoms.device, prob.amp, prob.f0, prob.k_cant)
    df = op.apply(FEs, (prob.nz * 4, 4, 0), prob.n_scan, 1, offset=2)
    return grid, FEs, PPs, df.reshape(int(prob.nxy[0]), int(prob.nxy[1]), prob.nz), spec


class DfImage:
    """Namespace for the autograd Function (built lazily so that importing needs no torch)."""
    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn


def _make_function():
    import torch

    class _DfImage(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xyz, q, R0, E0, prob):
            atoms = torch.cat([xyz.detach().float(), q.detach().float()[:, None]], dim=1).contiguous()
            cpar = lj_cpar(R0.detach(), E0.detach())
            grid, FEs, PPs, df, spec = _forward_parts(atoms, cpar, prob)
            ctx.prob = prob
            ctx.spec = spec
            ctx.dtypes = (xyz.dtype, q.dtype, R0.dtype, E0.dtype)
            ctx.save_for_backward(atoms, cpar, grid, R0.detach(), E0.detach(),
                                  PPs if not prob.recompute else torch.empty(0, device=xyz.device))
            return df

        @staticmethod
        def backward(ctx, gdf):
            prob = ctx.prob
            atoms, cpar, grid, R0, E0, PPs = ctx.saved_tensors
            dev = atoms.device
            spec = ctx.spec
            if prob.recompute:
                _, PPs = _relax(grid, prob, spec)
            g = gdf.detach().float().contiguous().reshape(prob.n_scan, prob.nz)
            opT = DF.z_operator(prob.df_kind, prob.nz, prob.dz, dev, prob.amp, prob.f0, prob.k_cant, transpose=True)
            U = torch.zeros((prob.n_scan, prob.nz, 4), dtype=torch.float32, device=dev)
            u_z = opT.apply(g, (prob.nz, 1, 0), prob.n_scan, 1)
            U[..., 2] = u_z[..., 0]
            gG, lam = D.relax_adjoint(grid, spec, prob.pts, prob.nz, prob.tip, PPs, U)
            st = lam[..., 3].to(torch.int32)
            prob.last_status = {"states": int(st.numel()), "gated": int((st != 0).sum()),
                                "residual": int(((st & 1) != 0).sum()), "nonfinite": int(((st & 2) != 0).sum()),
                                "singular": int(((st & 4) != 0).sum()), "unstable": int(((st & 8) != 0).sum())}
            if prob.strict and prob.last_status["nonfinite"]:
                raise RuntimeError(f"relax adjoint: {prob.last_status['nonfinite']} states non-finite")
            # the (c6, c12) columns only when R0 or E0 wants a gradient
            params = bool(ctx.needs_input_grad[2] or ctx.needs_input_grad[3])
            g6 = D.grid_adjoint(atoms, cpar, _lib.PPM_MODEL_LJ, spec, prob.Qs, prob.QZs, grid, gG, prob.npbc,
                                prob.lvA, prob.lvB, cap=prob.cap, params=params)
            R0d, E0d = R0.double(), E0.double()
            r5 = R0d ** 5
            r6 = r5 * R0d
            dR0 = g6[:, 4] * 12.0 * E0d * r5 + g6[:, 5] * 12.0 * E0d * r6 * r5
            dE0 = g6[:, 4] * 2.0 * r6 + g6[:, 5] * r6 * r6
            tx, tq, tr, te = ctx.dtypes
            return g6[:, :3].to(tx), g6[:, 3].to(tq), dR0.to(tr), dE0.to(te), None

    return _DfImage


def df_image(xyz, q, R0, E0, prob: Problem):
    """Differentiable df image (nx, ny, nz) on the GPU."""
    return DfImage.fn().apply(xyz, q, R0, E0, prob)


# ── gridless direct Morse backend ────────────────────────────────────────────────

@dataclass
class DirectProblem:
    """Non-differentiated part of the direct Morse forward (run_scan_morse_direct semantics)."""
    pts: object                 # (n_scan,4) float32 device tensor
    nxy: tuple
    nz: int
    tip: D.Tip
    Qs: np.ndarray
    QZs: np.ndarray
    K: float = -1.8             # Morse exponent per atom row (cMs.z)
    max_iter: int = 1000
    f2conv: float = 1e-12
    df_kind: str = "giessibl"
    amp: float = 1.0
    f0: float = 30000.0
    k_cant: float = 1800.0
    strict: bool = True         # raise if any adjoint state is non-finite (gated states are counted)
    last_status: dict = field(default_factory=dict)

    @property
    def n_scan(self):
        return int(self.nxy[0]) * int(self.nxy[1])

    @property
    def dz(self):
        return float(self.tip.tipC[3])


def _direct_inputs(xyz, q, R0, E0, prob: DirectProblem):
    import torch
    atoms = torch.cat([xyz.detach().float(), q.detach().float()[:, None]], dim=1).contiguous()
    na = atoms.shape[0]
    cMs = torch.zeros((na, 4), dtype=torch.float32, device=atoms.device)
    cMs[:, 0] = R0.detach().float()
    cMs[:, 1] = E0.detach().float()
    cMs[:, 2] = float(prob.K)
    return atoms, cMs


def _make_direct_function():
    import torch

    class _DirectDf(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xyz, q, R0, E0, prob):
            atoms, cMs = _direct_inputs(xyz, q, R0, E0, prob)
            FEs, PPs = D.relax_direct_morse(atoms, cMs, prob.pts, prob.nz, prob.tip, prob.Qs, prob.QZs, prob.max_iter,
                                            prob.f2conv, scan_ny=int(prob.nxy[1]))
            op = DF.z_operator(prob.df_kind, prob.nz, prob.dz, atoms.device, prob.amp, prob.f0, prob.k_cant)
            df = op.apply(FEs, (prob.nz * 4, 4, 0), prob.n_scan, 1, offset=2)
            ctx.prob = prob
            ctx.dtypes = (xyz.dtype, q.dtype, R0.dtype, E0.dtype)
            ctx.save_for_backward(atoms, cMs, PPs)
            return df.reshape(int(prob.nxy[0]), int(prob.nxy[1]), prob.nz)

        @staticmethod
        def backward(ctx, gdf):
            prob = ctx.prob
            atoms, cMs, PPs = ctx.saved_tensors
            dev = atoms.device
            g = gdf.detach().float().contiguous().reshape(prob.n_scan, prob.nz)
            opT = DF.z_operator(prob.df_kind, prob.nz, prob.dz, dev, prob.amp, prob.f0, prob.k_cant, transpose=True)
            U = torch.zeros((prob.n_scan, prob.nz, 4), dtype=torch.float32, device=dev)
            U[..., 2] = opT.apply(g, (prob.nz, 1, 0), prob.n_scan, 1)[..., 0]
            grad, _, diag = D.direct_morse_vjp(atoms, cMs, prob.pts, prob.nz, prob.tip, prob.Qs, prob.QZs, PPs, U)
            st = diag[..., 3].to(torch.int32)
            prob.last_status = {"states": int(st.numel()), "gated": int((st != 0).sum()),
                                **{name: int((st == c).sum()) for c, name in
                                   ((1, "nonconverged"), (2, "nonfinite"), (3, "unstable"), (4, "singular"))}}
            if prob.strict and prob.last_status["nonfinite"]:
                raise RuntimeError(f"direct adjoint: {prob.last_status['nonfinite']} states non-finite")
            tx, tq, tr, te = ctx.dtypes
            return grad[:, :3].to(tx), grad[:, 5].to(tq), grad[:, 3].to(tr), grad[:, 4].to(te), None

    return _DirectDf


class DirectDfImage:
    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            cls._fn = _make_direct_function()
        return cls._fn


def direct_df_image(xyz, q, R0, E0, prob: DirectProblem):
    """Differentiable df image (nx, ny, nz) of the gridless direct Morse path on the GPU;
    gradients w.r.t. atom positions, charges and Morse R0, E0 (the reference's theta,
    spammm/SPM/AFM.py:2241-2519)."""
    return DirectDfImage.fn().apply(xyz, q, R0, E0, prob)
