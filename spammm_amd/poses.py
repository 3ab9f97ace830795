"""Rigid placements of a molecule over a substrate: df images and their adjoint w.r.t. the
placement (shift, rotation about z) and the molecule's charges.

Every grid of a placement lives on the scan window: the sub-box of the grid (same voxels, same
spacing) that holds every probe position the scan can reach (the scan's tip positions moved by
the rest offset, plus lateral and vertical margins).  A placement whose probes leave the window
(checked after each relaxation) is recomputed on the full grid.  The substrate's raw force
field (no |E|>100 rescale) is projected once (its Coulomb part either
as the direct sum over npbc x,y images or, sub_coulomb = "ewald", as the 2D Ewald series of
the periodic slab, spammm_amd.ewald); every placement adds
the molecule's pair sum to it and applies the rescale to the total (ppm_grid_project_add), then
relaxes the scan and applies the df operator.  The backward takes the probe positions the
forward kept (16 B per state; with keep_states = False it recomputes the relaxation instead)
and, unless the shared grid buffer still holds this placement's grid (no other placement's
forward ran in between), recomputes the grid (one grid buffer is shared by all placements);
relaxation trajectories are never stored.  It then runs the implicit-function relax adjoint, the grid->atom adjoint of the molecule's atoms over
the substrate base (ppm_grid_adjoint_add) and the rigid-motion chain rule:
    x_i = R(theta) (x0_i - c) + c + s
    dL/ds = sum_i dL/dx_i,   dL/dtheta = sum_i dL/dx_i . R'(theta) (x0_i - c).
There is no reference counterpart (the reference relaxes one molecule per image without a
substrate grid); the df, loss and adjoint conventions are those of spammm_amd.diff.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib, df as DF
from . import device as D


@dataclass
class Scene:
    """Substrate + molecule template + scan; everything of the forward that is not a placement."""
    spec: D.GridSpec
    sub_atoms: object           # (ns,4) float32 device tensor (x,y,z,q)
    sub_cpar: object            # (ns,2) float32 device tensor (c6,c12)
    mol_xyz0: np.ndarray        # (na,3) float64 template positions
    mol_cpar: object            # (na,2) float32 device tensor
    pts: object                 # (n_scan,4) float32 device tensor
    nxy: tuple
    nz: int
    tip: D.Tip
    Qs: np.ndarray
    QZs: np.ndarray
    npbc: tuple = (1, 1)
    lvA: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    lvB: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    max_iter: int = 1000
    f2conv: float = 1e-12
    polish: int = 3
    df_kind: str = "giessibl"
    amp: float = 1.0
    f0: float = 30000.0
    k_cant: float = 1800.0
    strict: bool = True
    last_status: dict = field(default_factory=dict)
    base: object = None         # raw substrate grid, filled by prepare()
    grid: object = None         # placement grid buffer
    grid_cot: object = None     # adjoint buffer
    grid_serial: int = 0        # forward count; identifies the placement whose grid `grid` holds
    sub_coulomb: str = "direct"  # "direct": npbc image sum; "ewald": 2D Ewald series over lvA, lvB
    ewald_n_harm: int = 40
    ewald_drop_tol: float = 1e-6  # G with |W| below this fraction of the largest: FP32 position noise
    window_margin: tuple = (2.0, 1.0)  # lateral, vertical margins [A] of the scan window; None: full grid
    spec_run: object = None     # the grid the placements run on (the scan window or `spec`)
    window_fallbacks: int = 0   # placements recomputed on the full grid
    window_faces: tuple = None  # (lo, hi) bool[3]: window faces that are faces of the full grid
    full: object = None         # (base, grid, grid_cot, serial) of the full grid, built on first fallback
    keep_states: bool = True    # the backward reuses the forward's probe positions (else recomputes them)
    # column gating (diff.column_gate; off when gate_eig_frac <= 0): last_columns (nx, ny) bool, True =
    # the column's probes stay on one equilibrium branch; callers weight the loss by it
    gate_eig_frac: float = 0.0
    gate_max_jump: float = 0.3
    last_columns: object = None
    last_gate: dict = field(default_factory=dict)

    @property
    def n_scan(self):
        return int(self.nxy[0]) * int(self.nxy[1])

    @property
    def dz(self):
        return float(self.tip.tipC[3])

    @property
    def center(self):
        return self.mol_xyz0.mean(axis=0)

    def prepare(self):
        """Project the substrate once (raw, no rescale) on the grid the placements run on and
        allocate the shared buffers."""
        import torch
        if self.sub_coulomb not in ("direct", "ewald"):
            raise ValueError(f"sub_coulomb must be 'direct' or 'ewald', got {self.sub_coulomb!r}")
        self.spec_run = self.spec
        if self.window_margin is not None:
            w = scan_window(self.spec, self.pts.cpu().numpy(), self.nz, self.tip, *self.window_margin)
            if w is not None:
                self.spec_run, lo, hi = w
                self.window_faces = (lo, hi)
        self.base = self._substrate(self.spec_run)
        self.grid = torch.empty_like(self.base)
        self.grid_cot = torch.empty_like(self.base)
        return self

    @property
    def windowed(self):
        return self.spec_run is not self.spec

    def _substrate(self, spec):
        if self.sub_coulomb == "direct":
            return D.grid_project(self.sub_atoms, self.sub_cpar, _lib.PPM_MODEL_LJ, spec, self.Qs, self.QZs,
                                  self.npbc, self.lvA, self.lvB, cap=False)
        from .ewald import SurfaceEwald
        base = D.grid_project(self.sub_atoms, self.sub_cpar, _lib.PPM_MODEL_LJ, spec, np.zeros(4, np.float32),
                              np.zeros(4, np.float32), self.npbc, self.lvA, self.lvB, cap=False)
        ew = SurfaceEwald(self.sub_atoms.device)
        ew.prepare_system(self.sub_atoms.cpu().numpy(), self.lvA[:2], self.lvB[:2], self.ewald_n_harm,
                          self.ewald_drop_tol)
        return ew.tip_grid(spec, self.Qs, self.QZs, out=base, accumulate=True)

    def buffers(self, full=False):
        """(spec, base, grid, grid_cot) of the scan window (or of the full grid)."""
        import torch
        if not full or not self.windowed:
            return self.spec_run, self.base, self.grid, self.grid_cot
        if self.full is None:
            b = self._substrate(self.spec)
            self.full = [b, torch.empty_like(b), torch.empty_like(b), 0]
        return self.spec, self.full[0], self.full[1], self.full[2]

    def inside(self, PPs, guard=0.5, guard_shared=2) -> bool:
        """window_inside for the scene's window (True without one)."""
        if not self.windowed:
            return True
        return window_inside(self.spec_run, self.window_faces, PPs, guard, guard_shared)


def window_inside(spec_run, faces, PPs, guard=0.5, guard_shared=2) -> bool:
    """True if every probe position of PPs samples the window `spec_run` exactly as it samples the
    full grid: at least `guard` A inside each window face that cuts the full grid (so the FIRE steps
    between stored positions stay on the window's voxels), and `guard_shared` voxels inside a face
    the window shares with the full grid (where both would wrap, each to its own opposite face).
    An axis whose two faces are both shared is the full grid's axis: both wrap alike, no guard.
    faces: (lo, hi) bool[3], True where the window face is a face of the full grid."""
    import torch
    sp = spec_run
    d = np.array([sp.dA[0], sp.dB[1], sp.dC[2]], np.float64)
    shared_lo, shared_hi = (np.asarray(f, bool) for f in faces)
    whole = shared_lo & shared_hi
    g_lo = np.where(shared_lo, float(guard_shared), np.ceil(guard / d))
    g_hi = np.where(shared_hi, float(guard_shared), np.ceil(guard / d))
    mn, mx = torch.aminmax(PPs.reshape(-1, 4), dim=0)
    lo = mn[:3].cpu().numpy().astype(np.float64)
    hi = mx[:3].cpu().numpy().astype(np.float64)
    ulo = (lo - sp.p0.astype(np.float64)) / d
    uhi = (hi - sp.p0.astype(np.float64)) / d
    ok = whole | ((ulo >= g_lo) & (uhi < np.asarray(sp.n, np.float64) - 1 - g_hi))
    return bool(np.isfinite(ulo).all() and np.isfinite(uhi).all() and ok.all())


def scan_window(spec: D.GridSpec, pts, nz: int, tip: D.Tip, margin_xy: float, margin_z: float):
    """The sub-box of an axis-aligned grid that holds every probe rest position of the scan
    (tip positions pts + k dTip, k < nz, plus the rest offset dpos0 in the tip frame) widened by
    the margins, clamped to the grid's faces: (GridSpec on the same voxels, lo, hi) with lo / hi
    bool[3] marking the window faces that are faces of the grid.  None if the grid is not
    axis-aligned or the window is the whole grid."""
    dA, dB, dC = (np.asarray(v, np.float64)[:3] for v in (spec.dA, spec.dB, spec.dC))
    if dA[1] or dA[2] or dB[0] or dB[2] or dC[0] or dC[1]:
        return None
    d = np.array([dA[0], dB[1], dC[2]])
    R = np.stack([np.asarray(tip.tipA, np.float64)[:3], np.asarray(tip.tipB, np.float64)[:3],
                  np.asarray(tip.tipC, np.float64)[:3]])
    rest = R.T @ np.asarray(tip.dpos0, np.float64)[:3]
    dtip = np.asarray(tip.tipC, np.float64)[:3] * float(tip.tipC[3])
    P = np.asarray(pts, np.float64)[:, :3] + rest
    ends = np.concatenate([P, P + (nz - 1) * dtip])
    m = np.array([margin_xy, margin_xy, margin_z])
    lo, hi = ends.min(0) - m, ends.max(0) + m
    p0 = np.asarray(spec.p0, np.float64)[:3]
    n = np.asarray(spec.n, np.int64)
    i0 = np.maximum(np.floor((lo - p0) / d).astype(np.int64), 0)
    i1 = np.minimum(np.ceil((hi - p0) / d).astype(np.int64) + 1, n - 1)
    shared_lo, shared_hi = i0 == 0, i1 == n - 1
    if (shared_lo & shared_hi).all() or (i1 <= i0).any():
        return None
    nw = i1 - i0 + 1
    p0w = (p0 + i0 * d).astype(np.float32)
    L = nw * d
    dinvA, dinvB, dinvC = D.dinv_axis_aligned(L, p0w.astype(np.float64), tuple(int(v) for v in nw))
    w = D.GridSpec(tuple(int(v) for v in nw), p0w, np.asarray(spec.dA, np.float32).copy(),
                   np.asarray(spec.dB, np.float32).copy(), np.asarray(spec.dC, np.float32).copy(),
                   dinvA, dinvB, dinvC)
    return w, shared_lo, shared_hi


def place(xyz0, center, pose):
    """Positions of the template under pose (sx, sy, sz, theta): rotation about z through
    `center`, then the shift.  Works on numpy or torch float64 arrays."""
    sx, sy, sz, th = pose[0], pose[1], pose[2], pose[3]
    lib = np
    try:
        import torch
        if isinstance(pose, torch.Tensor):
            lib = torch
    except ImportError:
        pass
    c, s = lib.cos(th), lib.sin(th)
    dx = xyz0[:, 0] - center[0]
    dy = xyz0[:, 1] - center[1]
    x = c * dx - s * dy + center[0] + sx
    y = s * dx + c * dy + center[1] + sy
    z = xyz0[:, 2] + sz
    return lib.stack([x, y, z], 1)


def _serial(scene: Scene, full: bool):
    return scene.full[3] if (full and scene.windowed) else scene.grid_serial


def _project(scene: Scene, atoms, full=False):
    spec, base, grid, _ = scene.buffers(full)
    D.grid_project(atoms, scene.mol_cpar, _lib.PPM_MODEL_LJ, spec, scene.Qs, scene.QZs, scene.npbc,
                   scene.lvA, scene.lvB, cap=True, out=grid, base=base)
    if full and scene.windowed:
        scene.full[3] += 1
    else:
        scene.grid_serial += 1


def _forward(scene: Scene, atoms, grid_valid=False, full=False):
    spec, base, grid, _ = scene.buffers(full)
    if not grid_valid:
        _project(scene, atoms, full)
    FEs, PPs = D.relax_tilted(grid, spec, scene.pts, scene.nz, scene.tip, scene.max_iter, scene.f2conv,
                              True, scan_ny=int(scene.nxy[1]))
    if scene.polish > 0:
        D.relax_polish(grid, spec, scene.pts, scene.nz, scene.tip, PPs, FEs, scene.polish)
    return FEs, PPs


def _make_function():
    import torch

    class _PoseDf(torch.autograd.Function):
        @staticmethod
        def forward(ctx, pose, q, scene):
            xyz0 = torch.as_tensor(scene.mol_xyz0, dtype=torch.float64, device=pose.device)
            xyz = place(xyz0, torch.as_tensor(scene.center, device=pose.device), pose.detach())
            atoms = torch.cat([xyz.float(), q.detach().float()[:, None]], 1).contiguous()
            FEs, PPs = _forward(scene, atoms)
            full = False
            if not scene.inside(PPs):
                scene.window_fallbacks += 1
                full = True
                FEs, PPs = _forward(scene, atoms, full=True)
            if scene.gate_eig_frac > 0:
                from .diff import column_gate
                spec, _, grid, _ = scene.buffers(full)
                scene.last_columns, scene.last_gate = column_gate(grid, spec, scene.pts, scene.nxy, scene.nz,
                                                                  scene.tip, PPs, scene.gate_eig_frac,
                                                                  scene.gate_max_jump)
            ctx.full = full
            ctx.PPs = PPs if scene.keep_states else None
            ctx.grid_serial = _serial(scene, full)
            op = DF.z_operator(scene.df_kind, scene.nz, scene.dz, pose.device, scene.amp, scene.f0, scene.k_cant)
            df = op.apply(FEs, (scene.nz * 4, 4, 0), scene.n_scan, 1, offset=2)
            ctx.scene = scene
            ctx.save_for_backward(pose.detach(), q.detach())
            return df.reshape(int(scene.nxy[0]), int(scene.nxy[1]), scene.nz)

        @staticmethod
        def backward(ctx, gdf):
            scene = ctx.scene
            pose, q = ctx.saved_tensors
            dev = pose.device
            xyz0 = torch.as_tensor(scene.mol_xyz0, dtype=torch.float64, device=dev)
            ctr = torch.as_tensor(scene.center, device=dev)
            xyz = place(xyz0, ctr, pose)
            atoms = torch.cat([xyz.float(), q.float()[:, None]], 1).contiguous()
            # the grid if another placement has overwritten it; the probe positions unless kept
            full = ctx.full
            spec, base, grid, grid_cot = scene.buffers(full)
            grid_valid = _serial(scene, full) == ctx.grid_serial
            if ctx.PPs is not None:
                PPs = ctx.PPs
                ctx.PPs = None
                if not grid_valid:
                    _project(scene, atoms, full)
            else:
                _, PPs = _forward(scene, atoms, grid_valid=grid_valid, full=full)
            g = gdf.detach().float().contiguous().reshape(scene.n_scan, scene.nz)
            opT = DF.z_operator(scene.df_kind, scene.nz, scene.dz, dev, scene.amp, scene.f0, scene.k_cant,
                                transpose=True)
            U = torch.zeros((scene.n_scan, scene.nz, 4), dtype=torch.float32, device=dev)
            U[..., 2] = opT.apply(g, (scene.nz, 1, 0), scene.n_scan, 1)[..., 0]
            grid_cot.zero_()
            _, lam = D.relax_adjoint(grid, spec, scene.pts, scene.nz, scene.tip, PPs, U, grid_cot=grid_cot)
            st = lam[..., 3].to(torch.int32)
            scene.last_status = {"states": int(st.numel()), "gated": int((st != 0).sum()),
                                 "nonfinite": int(((st & 2) != 0).sum())}
            if scene.strict and scene.last_status["nonfinite"]:
                raise RuntimeError(f"pose adjoint: {scene.last_status['nonfinite']} states non-finite")
            g6 = D.grid_adjoint(atoms, scene.mol_cpar, _lib.PPM_MODEL_LJ, spec, scene.Qs, scene.QZs,
                                grid, grid_cot, scene.npbc, scene.lvA, scene.lvB, cap=True, base=base, params=False)
            gx = g6[:, :3]
            th = pose[3]
            dx = xyz0[:, 0] - ctr[0]
            dy = xyz0[:, 1] - ctr[1]
            dxr = -torch.sin(th) * dx - torch.cos(th) * dy       # d x_i / d theta
            dyr = torch.cos(th) * dx - torch.sin(th) * dy        # d y_i / d theta
            gpose = torch.stack([gx[:, 0].sum(), gx[:, 1].sum(), gx[:, 2].sum(),
                                 (gx[:, 0] * dxr + gx[:, 1] * dyr).sum()])
            return gpose, g6[:, 3].to(q.dtype), None

    return _PoseDf


class PoseDf:
    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn


def pose_df(pose, q, scene: Scene):
    """Differentiable df image (nx, ny, nz) of the molecule placed by pose = (sx, sy, sz, theta)
    (float64 tensor) with charges q over the scene's substrate."""
    return PoseDf.fn().apply(pose, q, scene)
