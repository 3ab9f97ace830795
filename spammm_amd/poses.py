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
relaxation trajectories are never stored.  It then runs the implicit-function relax adjoint, the
grid->atom adjoint of the molecule's atoms over the substrate base (ppm_grid_adjoint_add) and the
rigid-motion chain rule:
    x_i = R(theta) (x0_i - c) + c + s
    dL/ds = sum_i dL/dx_i,   dL/dtheta = sum_i dL/dx_i . R'(theta) (x0_i - c).
pose_df_batch runs a batch of placements the same way with several in flight at once, each on a
stream of its own, and keeps their grids for the backward within a memory budget.
There is no reference counterpart (the reference relaxes one molecule per image without a
substrate grid); the df, loss and adjoint conventions are those of spammm_amd.diff.
"""
from __future__ import annotations

import ctypes
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
    # pose_df_batch: placements in flight at once (one stream and one set of scratch buffers each),
    # and the bytes of placement grids a batched forward may keep for its backward (others are
    # recomputed there)
    batch_depth: int = 2
    batch_grid_budget: float = 64 * 2 ** 30
    # lanes per scan point of the batch's relaxations (device.relax_ctl): 1, the fewest instructions,
    # since they run beside another placement's grid projection (0: the choice for a lone launch)
    batch_lanes: int = 1
    batch_slots: list = field(default=None, repr=False)

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
    mn, mx = torch.aminmax(PPs.reshape(-1, 4), dim=0)
    return _inside_bounds(spec_run, faces, mn[:3].cpu().numpy(), mx[:3].cpu().numpy(), guard, guard_shared)


def _inside_bounds(spec_run, faces, lo, hi, guard=0.5, guard_shared=2) -> bool:
    """window_inside from the smallest and largest probe coordinates lo, hi (3,)."""
    sp = spec_run
    d = np.array([sp.dA[0], sp.dB[1], sp.dC[2]], np.float64)
    shared_lo, shared_hi = (np.asarray(f, bool) for f in faces)
    whole = shared_lo & shared_hi
    g_lo = np.where(shared_lo, float(guard_shared), np.ceil(guard / d))
    g_hi = np.where(shared_hi, float(guard_shared), np.ceil(guard / d))
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
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


def _atoms(xyz0, ctr, pose, q):
    import torch
    return torch.cat([place(xyz0, ctr, pose).float(), q.float()[:, None]], 1).contiguous()


def _pose_grad(g6, xyz0, ctr, pose):
    """(4,) gradient w.r.t. (sx, sy, sz, theta) from the atoms' position gradients g6[:, :3]."""
    import torch
    gx = g6[:, :3]
    th = pose[3]
    dx = xyz0[:, 0] - ctr[0]
    dy = xyz0[:, 1] - ctr[1]
    dxr = -torch.sin(th) * dx - torch.cos(th) * dy       # d x_i / d theta
    dyr = torch.cos(th) * dx - torch.sin(th) * dy        # d y_i / d theta
    return torch.stack([gx[:, 0].sum(), gx[:, 1].sum(), gx[:, 2].sum(), (gx[:, 0] * dxr + gx[:, 1] * dyr).sum()])


def _make_function():
    import torch

    class _PoseDf(torch.autograd.Function):
        @staticmethod
        def forward(ctx, pose, q, scene):
            xyz0 = torch.as_tensor(scene.mol_xyz0, dtype=torch.float64, device=pose.device)
            atoms = _atoms(xyz0, torch.as_tensor(scene.center, device=pose.device), pose.detach(), q.detach())
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
            atoms = _atoms(xyz0, ctr, pose, q)
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
            return _pose_grad(g6, xyz0, ctr, pose), g6[:, 3].to(q.dtype), None

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


# ---------------------------------------------------------------------------------------------
# Batched placements: several placements in flight on their own streams.  One placement's
# relaxation (one chain of dependent FIRE steps per scan point, ~1 wave per SIMD on a 192x192
# scan) and implicit-function adjoint leave most VALU issue slots idle; the next placement's grid
# projection or grid->atom adjoint fills them.  Each placement runs the same launches on the same
# inputs as pose_df, so its df image is the same bits.  The per-placement work on a stream is
# only the library's launches: a small torch kernel queued behind another stream's large launch
# waits for free compute units, so the atom positions, the df cotangents, the column gate and the
# pose gradients are computed for the whole batch before or after the streams run.

@dataclass
class _Slot:
    """One placement in flight: a stream of the library's own (device.Stream) and buffers allocated
    on the caller's stream, which waits for this stream at the end of every batched call (and when
    the slot is dropped) before any of them can be freed."""
    stream: object
    FEs: object                 # (n_scan, nz, 4) relaxation output of the placement in flight
    grid_cot: object            # adjoint accumulation buffer
    lam: object                 # (n_scan, nz, 4) relax adjoint multipliers + status word
    grid: object = None         # grid of a placement whose grid the forward does not keep
    workspace: object = None    # grid->atom adjoint workspace

    def __del__(self):
        try:
            self.stream.close()
        except Exception:
            pass


def _batch_slots(scene: Scene, depth: int):
    """The scene's `depth` slots, made again when the depth or the grid / scan shapes changed (a
    later prepare() with other window margins)."""
    import torch
    old = scene.batch_slots
    if (old is None or len(old) != depth or old[0].grid_cot.shape != scene.base.shape
            or old[0].FEs.shape != (scene.n_scan, scene.nz, 4)):
        dev = scene.base.device
        slots = []
        for sl in scene.batch_slots or ():
            sl.stream.close()
        for _ in range(depth):
            f32 = dict(dtype=torch.float32, device=dev)
            sl = _Slot(stream=D.Stream(dev), FEs=torch.empty((scene.n_scan, scene.nz, 4), **f32),
                       grid_cot=torch.empty_like(scene.base), lam=torch.empty((scene.n_scan, scene.nz, 4), **f32))
            slots.append(sl)
        scene.batch_slots = slots
    return scene.batch_slots


def _slot_grid(sl: _Slot, base):
    import torch
    if sl.grid is None:
        sl.grid = torch.empty_like(base)
    return sl.grid


class _Chain:
    """Runs the VALU-bound launches of a batch (grid projections, grid->atom adjoints) one at a
    time across the streams: each waits for the previous one, so a placement's latency-bound
    relaxation or relax adjoint runs beside the next placement's VALU-bound kernel rather than
    beside another relaxation."""

    def __init__(self):
        self.last = None

    def __call__(self, stream):
        import contextlib
        import torch

        @contextlib.contextmanager
        def ctx():
            if self.last is not None:
                stream.wait_event(self.last)
            yield
            self.last = torch.cuda.Event()
            self.last.record(stream)
        return ctx()


def _atoms_batch(xyz0, ctr, poses, qs):
    """(B, na, 4) float32 atoms of B placements: row b equals _atoms(xyz0, ctr, poses[b], qs[b])
    bit for bit (the same elementwise operations)."""
    import torch
    th = poses[:, 3:4]
    c, s = torch.cos(th), torch.sin(th)
    dx = (xyz0[:, 0] - ctr[0])[None, :]
    dy = (xyz0[:, 1] - ctr[1])[None, :]
    x = c * dx - s * dy + ctr[0] + poses[:, 0:1]
    y = s * dx + c * dy + ctr[1] + poses[:, 1:2]
    z = xyz0[None, :, 2] + poses[:, 2:3]
    return torch.stack([x.float(), y.float(), z.float(), qs.float()], 2).contiguous()


def _pose_grads(g6, xyz0, ctr, poses):
    """(B, 4) gradients w.r.t. (sx, sy, sz, theta) from the atoms' position gradients g6 (B, na, 6)."""
    import torch
    gx = g6[..., :3]
    th = poses[:, 3:4]
    dx = (xyz0[:, 0] - ctr[0])[None, :]
    dy = (xyz0[:, 1] - ctr[1])[None, :]
    dxr = -torch.sin(th) * dx - torch.cos(th) * dy       # d x_i / d theta
    dyr = torch.cos(th) * dx - torch.sin(th) * dy        # d y_i / d theta
    return torch.stack([gx[..., 0].sum(1), gx[..., 1].sum(1), gx[..., 2].sum(1),
                        (gx[..., 0] * dxr + gx[..., 1] * dyr).sum(1)], 1)


def _make_batch_function():
    import torch
    from .diff import gate_from_diag, gate_stats

    class _PoseDfBatch(torch.autograd.Function):
        @staticmethod
        def forward(ctx, poses, q, scene):
            if scene.base is None:
                raise RuntimeError("Scene.prepare() must run before pose_df_batch")
            na = int(scene.mol_xyz0.shape[0])
            if poses.dim() != 2 or poses.shape[1] != 4 or poses.shape[0] < 1:
                raise ValueError(f"poses must be (B >= 1, 4), got {tuple(poses.shape)}")
            if tuple(q.shape) not in ((na,), (int(poses.shape[0]), na)):
                raise ValueError(f"q must be ({na},) or ({int(poses.shape[0])}, {na}), got {tuple(q.shape)}")
            dev = poses.device
            B = int(poses.shape[0])
            poses_d, q_d = poses.detach(), q.detach()
            qs = q_d if q_d.dim() == 2 else q_d[None].expand(B, -1)
            xyz0 = torch.as_tensor(scene.mol_xyz0, dtype=torch.float64, device=dev)
            ctr = torch.as_tensor(scene.center, device=dev)
            spec, base, pts, tip = scene.spec_run, scene.base, scene.pts, scene.tip
            nx, ny = (int(v) for v in scene.nxy)
            ns, nz = scene.n_scan, scene.nz
            depth = max(1, int(scene.batch_depth))
            slots = _batch_slots(scene, depth)
            keep = bool(ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
            n_keep = min(B, int(scene.batch_grid_budget // (base.numel() * 4))) if keep else 0
            gate = scene.gate_eig_frac > 0
            op = DF.z_operator(scene.df_kind, nz, scene.dz, dev, scene.amp, scene.f0, scene.k_cant)
            cur = torch.cuda.current_stream(dev)
            atoms = _atoms_batch(xyz0, ctr, poses_d, qs)
            f32 = dict(dtype=torch.float32, device=dev)
            PPs = torch.empty((B, ns, nz, 4), **f32)
            df = torch.empty((B, nx, ny, nz), **f32)
            dg = torch.empty((B, ns, nz, 4), **f32) if gate else None
            grids = torch.empty((n_keep, *base.shape), dtype=base.dtype, device=dev)   # one allocation
            for sl in slots:
                if n_keep < B:
                    _slot_grid(sl, base)               # on the caller's stream, before the loop
                sl.stream.torch.wait_stream(cur)
            chain = _Chain()
            for b in range(B):
                sl = slots[b % depth]
                with torch.cuda.stream(sl.stream.torch):
                    g = grids[b] if b < n_keep else _slot_grid(sl, base)
                    with chain(sl.stream.torch):
                        D.grid_project(atoms[b], scene.mol_cpar, _lib.PPM_MODEL_LJ, spec, scene.Qs, scene.QZs,
                                       scene.npbc, scene.lvA, scene.lvB, cap=True, out=g, base=base)
                    D.relax_tilted(g, spec, pts, nz, tip, scene.max_iter, scene.f2conv, True, scan_ny=ny, FEs=sl.FEs,
                                   PPs=PPs[b], lanes=scene.batch_lanes)
                    if scene.polish > 0:
                        D.relax_polish(g, spec, pts, nz, tip, PPs[b], sl.FEs, scene.polish)
                    if gate:
                        D.relax_state_diag(g, spec, pts, nz, tip, PPs[b], out=dg[b])
                    op.apply(sl.FEs, (nz * 4, 4, 0), ns, 1, offset=2, out=df[b])
            for sl in slots:
                cur.wait_stream(sl.stream.torch)
            # probes that left the scan window: that placement again on the full grid
            full = np.zeros(B, bool)
            if scene.windowed:
                mn, mx = torch.aminmax(PPs.reshape(B, -1, 4), dim=1)
                mn, mx = mn[:, :3].cpu().numpy(), mx[:, :3].cpu().numpy()
                for b in range(B):
                    if _inside_bounds(spec, scene.window_faces, mn[b], mx[b]):
                        continue
                    scene.window_fallbacks += 1
                    full[b] = True
                    FEs_b, PPs_b = _forward(scene, atoms[b], full=True)
                    PPs[b] = PPs_b
                    op.apply(FEs_b, (nz * 4, 4, 0), ns, 1, offset=2, out=df[b])
                    if gate:
                        spec_f, _, grid_f, _ = scene.buffers(True)
                        D.relax_state_diag(grid_f, spec_f, pts, nz, tip, PPs[b], out=dg[b])
            if gate:
                keep_cols, parts = gate_from_diag(dg, PPs, pts, nz, tip, scene.gate_eig_frac, scene.gate_max_jump)
                del dg
                scene.last_columns = keep_cols.reshape(B, nx, ny)
                scene.last_gate = gate_stats(parts, scene.gate_eig_frac, scene.gate_max_jump)
            ctx.scene, ctx.grids, ctx.full, ctx.depth, ctx.PPs, ctx.atoms = scene, grids, full, depth, PPs, atoms
            ctx.q_shared = q.dim() == 1
            ctx.save_for_backward(poses_d, q_d)
            return df

        @staticmethod
        def backward(ctx, gdf):
            scene = ctx.scene
            poses, q = ctx.saved_tensors
            dev = poses.device
            B = int(poses.shape[0])
            xyz0 = torch.as_tensor(scene.mol_xyz0, dtype=torch.float64, device=dev)
            ctr = torch.as_tensor(scene.center, device=dev)
            spec, base, pts, tip = scene.spec_run, scene.base, scene.pts, scene.tip
            ns, nz = scene.n_scan, scene.nz
            slots = _batch_slots(scene, ctx.depth)
            opT = DF.z_operator(scene.df_kind, nz, scene.dz, dev, scene.amp, scene.f0, scene.k_cant, transpose=True)
            cur = torch.cuda.current_stream(dev)
            grids, PPs, atoms = ctx.grids, ctx.PPs, ctx.atoms
            # dL/dFEs of every placement (Fz channel only)
            g_all = gdf.detach().float().contiguous().reshape(B * ns, nz)
            U = torch.zeros((B, ns, nz, 4), dtype=torch.float32, device=dev)
            U[..., 2] = opT.apply(g_all, (nz, 1, 0), B * ns, 1).reshape(B, ns, nz)
            g6 = torch.empty((B, atoms.shape[1], 6), dtype=torch.float64, device=dev)
            status = torch.empty((B, ns, nz), dtype=torch.float32, device=dev)
            chain = _Chain()

            def adjoint(b, g, spec_b, base_b, grid_cot, lam, workspace, stream):
                grid_cot.zero_()
                D.relax_adjoint(g, spec_b, pts, nz, tip, PPs[b], U[b], grid_cot=grid_cot, lam=lam)
                status[b].copy_(lam[..., 3])
                with chain(stream):
                    D.grid_adjoint(atoms[b], scene.mol_cpar, _lib.PPM_MODEL_LJ, spec_b, scene.Qs, scene.QZs, g,
                                   grid_cot, scene.npbc, scene.lvA, scene.lvB, cap=True, workspace=workspace,
                                   base=base_b, params=False, out=g6[b])

            need = int(_lib.load().ppm_grid_adjoint_workspace_size(int(atoms.shape[1]), ctypes.byref(spec.c())))
            for sl in slots:
                if sl.workspace is None or sl.workspace.numel() < need:
                    sl.workspace = torch.empty(need, dtype=torch.uint8, device=dev)
                if len(grids) < B:
                    _slot_grid(sl, base)
                sl.stream.torch.wait_stream(cur)
            for b in range(B):
                if ctx.full[b]:
                    continue
                sl = slots[b % ctx.depth]
                with torch.cuda.stream(sl.stream.torch):
                    if b < len(grids):
                        g = grids[b]
                    else:
                        g = _slot_grid(sl, base)
                        with chain(sl.stream.torch):
                            D.grid_project(atoms[b], scene.mol_cpar, _lib.PPM_MODEL_LJ, spec, scene.Qs, scene.QZs,
                                           scene.npbc, scene.lvA, scene.lvB, cap=True, out=g, base=base)
                    adjoint(b, g, spec, base, sl.grid_cot, sl.lam, sl.workspace, sl.stream.torch)
            for sl in slots:
                cur.wait_stream(sl.stream.torch)
            for b in np.nonzero(ctx.full)[0]:        # full-grid placements, one at a time
                _project(scene, atoms[b], full=True)
                spec_f, base_f, grid_f, cot_f = scene.buffers(True)
                adjoint(int(b), grid_f, spec_f, base_f, cot_f, slots[0].lam.new_empty((ns, nz, 4)), None, cur)
            ctx.grids = ctx.PPs = ctx.atoms = None
            st = status.to(torch.int32)
            scene.last_status = {"states": B * ns * nz, "gated": int((st != 0).sum()),
                                 "nonfinite": int(((st & 2) != 0).sum())}
            if scene.strict and scene.last_status["nonfinite"]:
                raise RuntimeError(f"pose adjoint: {scene.last_status['nonfinite']} states non-finite")
            gpose = _pose_grads(g6, xyz0, ctr, poses)
            gq = g6[..., 3]
            return gpose.to(poses.dtype), (gq.sum(0) if ctx.q_shared else gq).to(q.dtype), None

    return _PoseDfBatch


class PoseDfBatch:
    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            cls._fn = _make_batch_function()
        return cls._fn


def pose_df_batch(poses, q, scene: Scene):
    """Differentiable df images (B, nx, ny, nz) of B placements, poses (B, 4) = (sx, sy, sz, theta)
    (float64 tensor), charges q (na,) shared by all placements or (B, na).  Placement b gives the
    bits of pose_df(poses[b], q, scene); up to scene.batch_depth placements run at once, each on a
    stream of its own.  The forward keeps the placements' grids for the backward while they fit in
    scene.batch_grid_budget bytes (the backward recomputes the others) and their probe positions
    (16 B per state).  With column gating on, scene.last_columns is (B, nx, ny)."""
    return PoseDfBatch.fn().apply(poses, q, scene)
