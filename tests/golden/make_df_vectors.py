"""Writes tests/golden/df_vectors.npz: inputs and expected outputs of the df functions, from the
CPU restatement in oracle/df_ref.py (which stays in the build container).  The GPU tests compare the
device operators against these vectors.

    python tests/golden/make_df_vectors.py

Cases (seeded; the names prefix the arrays in the file):
  grad   compute_df of Fz (5,6,41) at dz 0.1
  amp1   compute_df_amp of the same Fz, dz 0.1, A 1.0
  amp2   compute_df_amp, dz -0.1, A 0.37
  loss   df_loss_seed of FEs (4,3,9,4) with weights on every slice, dtip -0.1
  lossd  df_loss_seed with the default weights
  dir    compute_df_dir / compute_df_amp_dir (oblique and vertical) / compute_dissipation of FEs (6,5,11,4)
  post   compute_df_amp_dir and compute_dissipation on the heights of a ScanSpec ladder (6,5,nz,4)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import df_ref as R                      # noqa: E402
from spammm_amd.postprocess import afm_df_height_stacks   # noqa: E402

OUT = os.path.join(HERE, "df_vectors.npz")


def vectors():
    v = {}
    rng = np.random.default_rng(42)
    Fz = rng.normal(size=(5, 6, 41)).astype(np.float32)
    v["grad__Fz"] = Fz
    v["grad__df"] = R.compute_df(Fz.astype(np.float64), 0.1)
    v["amp1__df"] = R.compute_df_amp(Fz, 0.1, 1.0)
    v["amp2__df"] = R.compute_df_amp(Fz, -0.1, 0.37)

    rng = np.random.default_rng(7)
    FEs = rng.normal(size=(4, 3, 9, 4)).astype(np.float32)
    tgt = rng.normal(size=(4, 3, 9))
    w = rng.uniform(0.5, 1.5, size=(4, 3, 9))
    L, df, dL = R.df_loss_seed(FEs.astype(np.float64), tgt, -0.1, w)
    v.update({"loss__FEs": FEs, "loss__target": tgt, "loss__w": w, "loss__L": np.float64(L), "loss__df": df,
              "loss__dL": dL})
    L, df, dL = R.df_loss_seed(FEs.astype(np.float64), tgt, -0.1)
    v.update({"lossd__L": np.float64(L), "lossd__df": df, "lossd__dL": dL})

    rng = np.random.default_rng(9)
    FEs = rng.normal(size=(6, 5, 11, 4)).astype(np.float32)
    n = np.array([1., 2., 3.]) / np.linalg.norm([1., 2., 3.])
    sp = np.array([0.4, 0.3, 0.2])
    h = 3.0 + 0.2 * np.arange(11)
    bwd = FEs + rng.normal(scale=0.1, size=FEs.shape).astype(np.float32)
    v.update({"dir__FEs": FEs, "dir__bwd": bwd, "dir__n": n, "dir__spacing": sp, "dir__h_scan": h,
              "dir__h_df": h[2:9:3],
              "dir__df_dir": R.compute_df_dir(FEs.astype(np.float64), sp, n),
              "dir__df_amp_dir": R.compute_df_amp_dir(FEs, sp, n, 0.5),
              "dir__df_amp_vertical": R.compute_df_amp_dir(FEs, sp, (0., 0., 1.), 0.5),
              "dir__E_diss": R.compute_dissipation(FEs, bwd, h, h[2:9:3], 0.4)})

    rng = np.random.default_rng(17)
    h_df, h_Fz, h_scan = afm_df_height_stacks(3.0, 4.0, 0.1, amp=0.3, osc_dir=(1., 2., 3.))
    FEs = rng.normal(size=(6, 5, h_scan.size, 4)).astype(np.float32)
    bwd = FEs + rng.normal(scale=0.1, size=FEs.shape).astype(np.float32)
    n = np.array([1., 2., 3.]) / np.linalg.norm([1., 2., 3.])
    spacing = (0.2, 0.3, float(h_scan[1] - h_scan[0]))
    v.update({"post__FEs": FEs, "post__bwd": bwd,
              "post__df_amp_dir": R.compute_df_amp_dir(FEs.astype(np.float64), spacing, n, 0.3),
              "post__E_diss": R.compute_dissipation(FEs, bwd, h_scan, h_df, 0.3, n)})
    return v


if __name__ == "__main__":
    np.savez_compressed(OUT, **vectors())
    print(f"wrote {OUT}")
