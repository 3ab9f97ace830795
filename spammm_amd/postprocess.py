"""Height ladders and the backend-independent scan / result contract.

`afm_df_height_stacks` follows spammm/SPM/AFM_utils.py:653-681 and `ScanSpec`, `ScanResult`,
`shared_postprocess` follow spammm/SPM/AFM_utils.py:3307-3414: the same names, fields,
argument meaning, return types and exceptions.  The Δf and dissipation volumes are computed
by the device operators of `spammm_amd.afm` (`compute_df_amp_dir`, `compute_dissipation`);
the ladders and the index selection are host bookkeeping, as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import afm


def _ladder(h0: float, h1: float, dh: float) -> np.ndarray:
    """Equally spaced heights h0, h0+dh, ... through h1 (count rounded), float32."""
    count = int(round((float(h1) - float(h0)) / float(dh))) + 1
    return (float(h0) + float(dh) * np.arange(count, dtype=np.float64)).astype(np.float32)


def afm_df_height_stacks(h_min, h_max, h_step, amp=1.0, amp_align=True, osc_dir=(0., 0., 1.)):
    """Height ladders of one AFM image series (spammm/SPM/AFM_utils.py:653-681).

    Heights are probe-atom centre z above the molecule plane.  Returns
      h_df   -- oscillation-centre heights where Δf is read, h_min..h_max step h_step;
      h_Fz   -- heights where Fz is read: h_df lowered by the vertical amplitude
                amp*|n_z| when `amp_align`, else h_df;
      h_scan -- the dense approach covering h_df ± amp*|n_z| at the same step, which
                `compute_df_amp_dir` needs.
    A purely lateral oscillation (n_z = 0) leaves all three on the same heights.
    Raises ValueError for a zero / non-finite / non-3-vector `osc_dir` and when h_scan has
    fewer than 3 heights.
    """
    h_df = _ladder(h_min, h_max, h_step)
    n = np.asarray(osc_dir, dtype=np.float64)
    norm = float(np.linalg.norm(n)) if n.shape == (3,) else float("nan")
    if n.shape != (3,) or not np.isfinite(norm) or norm <= 0.0:
        raise ValueError(f"afm_df_height_stacks: osc_dir must be a finite non-zero vector, got {osc_dir!r}")
    amp_z = float(amp) * abs(float(n[2]) / norm)
    if amp_align:
        h_Fz = (h_df.astype(np.float64) - amp_z).astype(np.float32)
    else:
        h_Fz = h_df.copy()
    h_scan = _ladder(float(h_df[0]) - amp_z, float(h_df[-1]) + amp_z, h_step)
    if h_scan.size < 3:
        raise ValueError(f"afm_df_height_stacks: need denser heights for amp={amp}: n_scan={h_scan.size}")
    return h_df, h_Fz, h_scan


@dataclass
class ScanSpec:
    """Scan / grid contract shared by the force-field backends (spammm/SPM/AFM_utils.py:3307-3323)."""
    scan_xs: np.ndarray        # (nx,) x scan axis [A]
    scan_ys: np.ndarray        # (ny,) y scan axis [A]
    h_df: np.ndarray           # (nz_df,) Δf heights [A]
    h_Fz: np.ndarray           # (nz_Fz,) Fz heights [A]
    h_scan: np.ndarray         # (nz_scan,) dense approach [A]
    amplitude: float           # oscillation amplitude [A, peak]
    osc_dir: tuple             # (3,) oscillation direction (any length)
    K_LAT: float               # lateral stiffness [eV/A^2]
    K_RAD: float               # radial stiffness [eV/A^2]
    bond_length: float         # probe bond length [A]
    scan_margin: float = 2.0   # xy margin around the molecule [A]


@dataclass
class ScanResult:
    """Backend-independent result of `shared_postprocess` (spammm/SPM/AFM_utils.py:3326-3359);
    readable as a dict too (`r['df']`, `'df' in r`, `r.get('Fz')`)."""
    df: np.ndarray             # (nx, ny, nz_df)
    Fz: np.ndarray             # (nx, ny, nz_Fz)
    heights: np.ndarray        # (nz_df,)
    heights_Fz: np.ndarray     # (nz_Fz,)
    scan_xs: np.ndarray        # (nx,)
    scan_ys: np.ndarray        # (ny,)
    amp_align: bool            # h_Fz lowered by the vertical amplitude
    FEs: np.ndarray            # (nx, ny, nz_scan, 4)
    tip_disp: dict             # {'dx','dy','dz'} each (nx, ny, nz_scan)
    E_diss: np.ndarray         # (nx, ny, nz_df) dissipated energy per cycle [eV]
    backend_name: str = "unknown"
    fft_path: str = "none"

    def __getitem__(self, key):
        return getattr(self, key)

    def __contains__(self, key):
        return hasattr(self, key)

    def get(self, key, default=None):
        return getattr(self, key, default)


def _step(axis) -> float:
    axis = np.asarray(axis)
    return float(axis[1] - axis[0]) if axis.size > 1 else 0.0


def _nearest(ladder: np.ndarray, heights) -> list:
    """Index of the ladder entry nearest to each height (the first one on a tie)."""
    ladder = np.asarray(ladder)
    return [int(np.argmin(np.abs(ladder - h))) for h in np.asarray(heights)]


def shared_postprocess(FEs_full, scan_spec: ScanSpec, *, FEs_bwd=None, tip_disp=None, backend_name="unknown",
                       fft_path="none"):
    """Δf, Fz and dissipation read from a relaxed force volume (spammm/SPM/AFM_utils.py:3362-3414).

    FEs_full (nx, ny, nz_scan, 4) covers scan_xs x scan_ys x h_scan.  Δf is the directional
    amplitude-averaged Δf along the unit oscillation direction over the whole volume, read at the
    h_scan entries nearest to h_df; Fz is read at those nearest to h_Fz; E_diss integrates the
    forward-minus-backward stroke force over each oscillation (zeros without `FEs_bwd`);
    `tip_disp` defaults to zero volumes.
    """
    FEs_full = np.asarray(FEs_full, dtype=np.float32)
    nx, ny, nz_scan = FEs_full.shape[:3]
    spacing = (_step(scan_spec.scan_xs), _step(scan_spec.scan_ys), _step(scan_spec.h_scan))
    n = np.asarray(scan_spec.osc_dir, dtype=np.float64)
    n = n / np.linalg.norm(n)
    amp = float(scan_spec.amplitude)
    df_vol = afm.compute_df_amp_dir(FEs_full, spacing, osc_dir=n, amp=amp)
    df = df_vol[:, :, _nearest(scan_spec.h_scan, scan_spec.h_df)]
    Fz = FEs_full[..., 2][:, :, _nearest(scan_spec.h_scan, scan_spec.h_Fz)]
    if FEs_bwd is None:
        E_diss = np.zeros((nx, ny, len(scan_spec.h_df)), dtype=np.float32)
    else:
        E_diss = afm.compute_dissipation(FEs_full, FEs_bwd, scan_spec.h_scan, scan_spec.h_df, amp, osc_dir=n)
    if tip_disp is None:
        tip_disp = {c: np.zeros((nx, ny, nz_scan), dtype=np.float32) for c in ("dx", "dy", "dz")}
    return ScanResult(df=df, Fz=Fz, heights=np.asarray(scan_spec.h_df, dtype=np.float32),
                      heights_Fz=np.asarray(scan_spec.h_Fz, dtype=np.float32),
                      scan_xs=np.asarray(scan_spec.scan_xs, dtype=np.float32),
                      scan_ys=np.asarray(scan_spec.scan_ys, dtype=np.float32),
                      amp_align=bool(abs(float(n[2])) > 1e-12), FEs=FEs_full, tip_disp=tip_disp, E_diss=E_diss,
                      backend_name=backend_name, fft_path=fft_path)
