"""Finite-difference check of the differentiable df image against the FP64 device forward.

The differentiable path (diff.df_image) evaluates the forward in FP32 and its adjoint with FP64
solves.  This module evaluates the same loss with every stage in FP64 on the device -- FP64 grid
projection (ppm_grid_project_f64), FP64 relaxation to |F| < 1e-12 eV/A (ppm_relax_strokes_tilted_f64,
tip anchors advanced in float32 as the FP32 scan advances them) and the FP64 df operator -- and
takes central differences of it, at reference precision (the role of the reference's CPU64 oracle
in its VJP tests, tests/SPM/test_afm_morse.py:1472-1628).

The loss is L = 1/2 sum_columns w (df - target)^2 over pixel columns with weight w (the
column mask of diff.stable_columns, or all ones), the L2 convention of df_loss_seed
(spammm/SPM/AFM.py:3125-3200) without its normalisation.  Differences are taken at several steps
h; the reference's own plateau rule (tests/SPM/test_afm_morse.py:1670-1716) accepts the
best-agreeing step, and the per-step agreement is reported beside it.
"""
from __future__ import annotations

import numpy as np

from . import df as DF
from . import device as D


class Fp64Forward:
    """FP64 df image of a diff.Problem on the device, for fixed R0, E0 (LJ, c6 = 2 E0 R0^6,
    c12 = E0 R0^12 in FP64)."""

    def __init__(self, prob, R0, E0, f32_anchors=True, max_iter=40000, f2conv=1e-24):
        import torch
        self.prob = prob
        self.pts64 = prob.pts.double().contiguous()
        self.M = torch.as_tensor(DF.giessibl_matrix(prob.nz, prob.dz, prob.amp, prob.f0, prob.k_cant),
                                 device=prob.pts.device) if prob.df_kind == "giessibl" else None
        if self.M is None:
            raise ValueError("the FP64 forward implements the Giessibl df operator only")
        r6 = R0.double() ** 6
        self.cpar = torch.stack([2.0 * E0.double() * r6, E0.double() * r6 * r6], 1).contiguous()
        self.f32_anchors, self.max_iter, self.f2conv = bool(f32_anchors), int(max_iter), float(f2conv)
        # the grid the FP32 forward ran on (its scan window, if it has one)
        self.spec = prob.last_spec if prob.last_spec is not None else prob.run_spec()[0]
        self._grid = None

    def df(self, xyz, q):
        """(n_scan, nz) float64 df at positions xyz (na,3) and charges q (na,) (float64 device)."""
        import torch
        p = self.prob
        atoms = torch.cat([xyz.double(), q.double()[:, None]], 1).contiguous()
        self._grid = D.grid_project_f64(atoms, self.cpar, 0, self.spec, p.Qs, p.QZs, p.npbc, p.lvA, p.lvB, cap=p.cap,
                                        out=self._grid)
        FEs, _ = D.relax_tilted_f64(self._grid, self.spec, self.pts64, p.nz, p.tip, max_iter=self.max_iter,
                                    f2conv=self.f2conv, check_before=True, want_pps=False,
                                    f32_anchors=self.f32_anchors)
        return FEs[..., 2] @ self.M.T


def l2_loss(df, target, columns=None):
    """1/2 sum w (df - target)^2 in float64; df, target (n_scan, nz) or (nx, ny, nz); columns: (nx, ny)
    bool/float weights or None."""
    n = df.numel() // df.shape[-1]
    r = df.double().reshape(n, -1) - target.double().reshape(n, -1)
    if columns is not None:
        r = r * columns.reshape(-1, 1).double()
    return 0.5 * float((r * r).sum())


def fd_check(fwd: Fp64Forward, target, xyz0, q0, grad_xyz, grad_q, sel, steps=((1e-3, 1e-4), (1e-4, 1e-5),
             (1e-5, 1e-6)), columns=None, tol=1e-3, order4_step=None):
    """Central differences of the FP64 loss on parameters `sel` (indices into [xyz.flatten(), q]),
    one pair (position step A, charge step e) per entry of `steps`; each compared with the adjoint's
    gradient (grad_xyz (na,3), grad_q (na,)).  order4_step = (A, e): also the fourth-order central
    difference (8 (L(+h) - L(-h)) - (L(+2h) - L(-2h))) / 12h at that step.  Returns a dict: per step
    the count within `tol` (relative to |FD|), median and max relative error; the plateau count
    (best-agreeing second-order step per parameter); and the rows (name, adjoint, FD per step)."""
    na = xyz0.shape[0]

    def loss(x, q):
        return l2_loss(fwd.df(x, q), target, columns)

    memo = {}

    def shifted(k, hx, hq, m):
        key = (k, hx, hq, m)
        if key not in memo:
            x, q = xyz0.clone(), q0.clone()
            if k < 3 * na:
                i, c = divmod(k, 3)
                x[i, c] += m * hx
            else:
                q[k - 3 * na] += m * hq
            memo[key] = loss(x, q)
        return memo[key]

    rows, rows4 = [], []
    for k in sel:
        k = int(k)
        if k < 3 * na:
            i, c = divmod(k, 3)
            ad, name = float(grad_xyz[i, c]), f"xyz[{i},{c}]"
        else:
            i = k - 3 * na
            ad, name = float(grad_q[i]), f"q[{i}]"
        fds = []
        for hx, hq in steps:
            h = hx if k < 3 * na else hq
            fds.append((shifted(k, hx, hq, 1) - shifted(k, hx, hq, -1)) / (2 * h))
        rows.append((name, ad, fds))
        if order4_step is not None:
            hx, hq = order4_step
            h = hx if k < 3 * na else hq
            d1 = shifted(k, hx, hq, 1) - shifted(k, hx, hq, -1)
            d2 = shifted(k, hx, hq, 2) - shifted(k, hx, hq, -2)
            rows4.append(abs(ad - (8 * d1 - d2) / (12 * h)) / max(abs((8 * d1 - d2) / (12 * h)), 1e-30))
    rel = np.array([[abs(ad - f) / max(abs(f), 1e-30) for f in fds] for _, ad, fds in rows])   # (n, n_steps)
    out = {"tol": tol, "n": len(rows), "per_step": []}
    for j, (hx, hq) in enumerate(steps):
        out["per_step"].append({"h": f"{hx:g} A / {hq:g} e", "within_tol": int((rel[:, j] <= tol).sum()),
                                "median_rel_err": float(np.median(rel[:, j])), "max_rel_err": float(rel[:, j].max())})
    if rows4:
        r4 = np.array(rows4)
        out["order4"] = {"h": f"{order4_step[0]:g} A / {order4_step[1]:g} e", "within_tol": int((r4 <= tol).sum()),
                         "median_rel_err": float(np.median(r4)), "max_rel_err": float(r4.max())}
    best = rel.min(axis=1)
    out["plateau_within_tol"] = int((best <= tol).sum())
    out["plateau_median_rel_err"] = float(np.median(best))
    out["plateau_max_rel_err"] = float(best.max())
    out["rows"] = [(n, round(a, 6), [round(f, 6) for f in fs]) for n, a, fs in rows]
    return out
