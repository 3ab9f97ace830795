"""Rigid placements of a molecule over a substrate: df images and their adjoint w.r.t. the
placement (shift, rotation about z) and the molecule's charges.

The substrate's raw force field (no |E|>100 rescale) is projected once (its Coulomb part either
as the direct sum over npbc x,y images or, sub_coulomb = "ewald", as the 2D Ewald series of
the periodic slab, spammm_amd.ewald); every placement adds
the molecule's pair sum to it and applies the rescale to the total (ppm_grid_project_add), then
relaxes the scan and applies the df operator.  The backward recomputes the probe positions
and, unless the shared grid buffer still holds this placement's grid (no other placement's
forward ran in between), the grid (one grid buffer is shared by all placements),
runs the implicit-function relax adjoint, the grid->atom adjoint of the molecule's atoms over
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
        """Project the substrate once (raw, no rescale) and allocate the shared buffers."""
        import torch
        if self.sub_coulomb == "direct":
            self.base = D.grid_project(self.sub_atoms, self.sub_cpar, _lib.PPM_MODEL_LJ, self.spec, self.Qs,
                                       self.QZs, self.npbc, self.lvA, self.lvB, cap=False)
        elif self.sub_coulomb == "ewald":
            from .ewald import SurfaceEwald
            self.base = D.grid_project(self.sub_atoms, self.sub_cpar, _lib.PPM_MODEL_LJ, self.spec,
                                       np.zeros(4, np.float32), np.zeros(4, np.float32), self.npbc, self.lvA,
                                       self.lvB, cap=False)
            ew = SurfaceEwald(self.sub_atoms.device)
            ew.prepare_system(self.sub_atoms.cpu().numpy(), self.lvA[:2], self.lvB[:2], self.ewald_n_harm,
                              self.ewald_drop_tol)
            ew.tip_grid(self.spec, self.Qs, self.QZs, out=self.base, accumulate=True)
        else:
            raise ValueError(f"sub_coulomb must be 'direct' or 'ewald', got {self.sub_coulomb!r}")
        self.grid = torch.empty_like(self.base)
        self.grid_cot = torch.empty_like(self.base)
        return self


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


def _forward(scene: Scene, atoms, grid_valid=False):
    if not grid_valid:
        D.grid_project(atoms, scene.mol_cpar, _lib.PPM_MODEL_LJ, scene.spec, scene.Qs, scene.QZs, scene.npbc,
                       scene.lvA, scene.lvB, cap=True, out=scene.grid, base=scene.base)
        scene.grid_serial += 1
    FEs, PPs = D.relax_tilted(scene.grid, scene.spec, scene.pts, scene.nz, scene.tip, scene.max_iter, scene.f2conv,
                              True, scan_ny=int(scene.nxy[1]))
    if scene.polish > 0:
        D.relax_polish(scene.grid, scene.spec, scene.pts, scene.nz, scene.tip, PPs, FEs, scene.polish)
    return FEs, PPs


def _make_function():
    import torch

    class _PoseDf(torch.autograd.Function):
        @staticmethod
        def forward(ctx, pose, q, scene):
            xyz0 = torch.as_tensor(scene.mol_xyz0, dtype=torch.float64, device=pose.device)
            xyz = place(xyz0, torch.as_tensor(scene.center, device=pose.device), pose.detach())
            atoms = torch.cat([xyz.float(), q.detach().float()[:, None]], 1).contiguous()
            FEs, _ = _forward(scene, atoms)
            ctx.grid_serial = scene.grid_serial
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
            # recompute the probe positions, and the grid if another placement has overwritten it
            _, PPs = _forward(scene, atoms, grid_valid=scene.grid_serial == ctx.grid_serial)
            g = gdf.detach().float().contiguous().reshape(scene.n_scan, scene.nz)
            opT = DF.z_operator(scene.df_kind, scene.nz, scene.dz, dev, scene.amp, scene.f0, scene.k_cant,
                                transpose=True)
            U = torch.zeros((scene.n_scan, scene.nz, 4), dtype=torch.float32, device=dev)
            U[..., 2] = opT.apply(g, (scene.nz, 1, 0), scene.n_scan, 1)[..., 0]
            scene.grid_cot.zero_()
            _, lam = D.relax_adjoint(scene.grid, scene.spec, scene.pts, scene.nz, scene.tip, PPs, U,
                                     grid_cot=scene.grid_cot)
            st = lam[..., 3].to(torch.int32)
            scene.last_status = {"states": int(st.numel()), "gated": int((st != 0).sum()),
                                 "nonfinite": int(((st & 2) != 0).sum())}
            if scene.strict and scene.last_status["nonfinite"]:
                raise RuntimeError(f"pose adjoint: {scene.last_status['nonfinite']} states non-finite")
            g6 = D.grid_adjoint(atoms, scene.mol_cpar, _lib.PPM_MODEL_LJ, scene.spec, scene.Qs, scene.QZs,
                                scene.grid, scene.grid_cot, scene.npbc, scene.lvA, scene.lvB, cap=True,
                                base=scene.base)
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
