"""Fz -> df conversion on the GPU (stage c of the AFM path).

Every df definition used by the reference is a linear map along one axis of the
(nx, ny, nz) force volume.  The host builds that map once, in FP64, as a banded
matrix (per output row: first input index ``start[j]`` and ``taps`` weights); the
GPU applies it with ``ppm_axis_operator`` in FP32.

Definitions (z axis; ``dz`` is the signed z step between consecutive slices):

* ``gradient`` — ``-np.gradient(Fz, |dz|)`` of compute_df (spammm/SPM/AFM.py:3048-3050):
  central differences inside, one-sided at both ends.
* ``amp9`` — compute_df_amp (spammm/SPM/AFM.py:3052-3089): mean over the 9
  Gauss-Chebyshev nodes u_k = cos((2k+1)pi/18) (equal weights) of -dFz/dz at
  z + u_k*A, linear interpolation with the coordinate clamped to the data
  ('nearest'), then np.gradient.  This is the arcsine (time-averaged) mean of
  -dFz/dz, which is what the reference computes.
* ``central`` — the df of df_loss_seed (spammm/SPM/AFM.py:3125-3200): interior
  central differences, zero at both boundary slices.
* ``giessibl`` — the Giessibl frequency shift of an oscillating tip,
    df(z_c) = -(f0 / (pi k A)) * int_{-1}^{1} Fz(z_c + A u) u / sqrt(1-u^2) du,
  integrated exactly over the piecewise-linear interpolant of the samples
  (clamped to the end values outside the scanned range), with k in N/m, Fz in
  eV/A, A in A (1 eV/A^2 = 16.02176634 N/m).  Output in Hz, row j = oscillation
  centre at slice j.  For A -> 0 it reduces to -(f0/2k) dFz/dz; for a force
  linear in z it equals -(f0/2k) dFz/dz for every A.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from . import _lib

EV_A2_TO_N_M = 16.02176634  # spammm/SPM/AFM.py:67-72


# ----------------------------------------------------------------------------------
# dense FP64 operators along one axis (n_out, n_in)
# ----------------------------------------------------------------------------------

def gradient_matrix(n: int, h: float) -> np.ndarray:
    """np.gradient(., h) along an axis of length n (edge_order=1) as an (n,n) matrix."""
    if n < 2:
        raise ValueError("gradient needs at least 2 samples")
    G = np.zeros((n, n))
    G[0, 0], G[0, 1] = -1.0 / h, 1.0 / h
    G[-1, -2], G[-1, -1] = -1.0 / h, 1.0 / h
    for i in range(1, n - 1):
        G[i, i - 1], G[i, i + 1] = -0.5 / h, 0.5 / h
    return G


def shift_matrix(n: int, s: float) -> np.ndarray:
    """Linear interpolation at coordinate i+s, coordinate clamped to [0, n-1] (mode 'nearest')."""
    S = np.zeros((n, n))
    for i in range(n):
        c = min(max(i + s, 0.0), n - 1.0)
        j = int(math.floor(c))
        t = c - j
        if j >= n - 1:
            S[i, n - 1] += 1.0
        else:
            S[i, j] += 1.0 - t
            S[i, j + 1] += t
    return S


def gc9():
    k = np.arange(9)
    return np.cos((2 * k + 1) / 18.0 * np.pi), np.full(9, 1.0 / 9.0)


def amp9_matrix(n: int, dz: float, amp: float) -> np.ndarray:
    h = abs(dz)
    u, w = gc9()
    G = gradient_matrix(n, h)
    M = np.zeros((n, n))
    for uk, wk in zip(u, w):
        M -= wk * G @ shift_matrix(n, uk * amp / h)
    return M


def central_matrix(n: int, dz: float) -> np.ndarray:
    h = abs(dz)
    M = np.zeros((n, n))
    for i in range(1, n - 1):
        M[i, i - 1], M[i, i + 1] = 0.5 / h, -0.5 / h
    return M


def central_seed_matrix(n: int, dz: float) -> np.ndarray:
    """dL/dFz = S @ dL/ddf of df_loss_seed (spammm/SPM/AFM.py:3185-3194), as written there:
    S[m, m+1] = +1/(2|dz|) for m in [0, n-2], S[m, m-1] = -1/(2|dz|) for m in [1, n-1].
    It equals central_matrix(n, dz).T except at S[1, 0] and S[n-2, n-1], which multiply
    the boundary residuals (zero under the default weights)."""
    c = 1.0 / (2.0 * abs(dz))
    S = np.zeros((n, n))
    for m in range(n - 1):
        S[m, m + 1] = c
    for m in range(1, n):
        S[m, m - 1] = -c
    return S


def _J1(u):
    return -math.sqrt(max(0.0, 1.0 - u * u))


def _J2(u):
    return 0.5 * (math.asin(max(-1.0, min(1.0, u))) - u * math.sqrt(max(0.0, 1.0 - u * u)))


def giessibl_matrix(n: int, dz: float, amp: float, f0: float, k: float) -> np.ndarray:
    """Exact Giessibl integral of the clamped piecewise-linear interpolant, as an (n,n) matrix [Hz per eV/A]."""
    if amp <= 0:
        raise ValueError("amplitude must be > 0")
    if dz == 0:
        raise ValueError("dz must be non-zero")
    beta = amp / dz                       # index units per unit u (signed)
    pref = -(f0 / (math.pi * k * amp)) * EV_A2_TO_N_M
    M = np.zeros((n, n))
    for i in range(n):
        # breakpoints in u where c(u) = i + beta*u crosses an integer or a clamp bound
        cs = [i - abs(beta), i + abs(beta)]
        lo, hi = min(cs), max(cs)
        pts = [-1.0, 1.0]
        for m in range(int(math.ceil(lo)), int(math.floor(hi)) + 1):
            pts.append((m - i) / beta)
        pts = sorted(set(p for p in pts if -1.0 <= p <= 1.0))
        for a, b in zip(pts[:-1], pts[1:]):
            if b - a <= 0:
                continue
            um = 0.5 * (a + b)
            cm = i + beta * um
            dJ1 = _J1(b) - _J1(a)
            dJ2 = _J2(b) - _J2(a)
            if cm <= 0.0:
                M[i, 0] += dJ1
            elif cm >= n - 1:
                M[i, n - 1] += dJ1
            else:
                j = int(math.floor(cm))
                # F(c) = F_j (1 - t) + F_{j+1} t,  t = c - j = (i - j) + beta*u
                M[i, j] += (1.0 - (i - j)) * dJ1 - beta * dJ2
                M[i, j + 1] += (i - j) * dJ1 + beta * dJ2
    return pref * M


# ----------------------------------------------------------------------------------
# banded form + device upload
# ----------------------------------------------------------------------------------

def band(M: np.ndarray):
    """Compress a dense (n_out, n_in) operator into (start (n_out,), W (n_out, taps))."""
    n_out, n_in = M.shape
    nz = np.abs(M) > 0
    lo = np.array([np.argmax(r) if r.any() else 0 for r in nz])
    hi = np.array([n_in - 1 - np.argmax(r[::-1]) if r.any() else 0 for r in nz])
    taps = int(max(1, (hi - lo + 1).max()))
    start = np.minimum(lo, n_in - taps).astype(np.int32)
    W = np.zeros((n_out, taps))
    for j in range(n_out):
        W[j] = M[j, start[j]:start[j] + taps]
    return start, W


class AxisOperator:
    """A banded FP64-built operator resident on the device (float32 weights, int32 starts)."""

    def __init__(self, M: np.ndarray, device):
        import torch
        self.M = M
        self.n_out, self.n_in = M.shape
        start, W = band(M)
        self.taps = W.shape[1]
        self.start = torch.as_tensor(start, dtype=torch.int32, device=device)
        self.W = torch.as_tensor(W.astype(np.float32), device=device).contiguous()

    def apply(self, x, axis_strides, n_outer, n_inner, chw=(1.0, 0, 0, 0), nch=1, offset=0, out=None,
              alpha=1.0, accumulate=False):
        """Apply to a float32 device tensor ``x`` (any layout given by strides in floats)."""
        import torch
        so, sa, si = axis_strides
        if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
            raise ValueError("AxisOperator.apply expects a contiguous float32 CUDA tensor")
        if min(so, sa, si) < 0 or not 1 <= nch <= 4 or offset < 0:
            raise ValueError("invalid strides / channels / offset")
        last = offset + (n_outer - 1) * so + (self.n_in - 1) * sa + (n_inner - 1) * si + (nch - 1)
        if last >= x.numel():
            raise ValueError(f"operator would read element {last} of a tensor with {x.numel()}")
        if out is None:
            out = torch.empty((n_outer, self.n_out, n_inner), dtype=torch.float32, device=x.device)
        elif (out.dtype != torch.float32 or not out.is_contiguous() or out.device != x.device
              or out.numel() != n_outer * self.n_out * n_inner):
            raise ValueError("out must be a contiguous float32 tensor of n_outer*n_out*n_inner elements")
        L = _lib.load()
        stream = torch.cuda.current_stream(x.device).cuda_stream
        rc = L.ppm_axis_operator(x.data_ptr() + 4 * offset, so, sa, si, _lib.f4(chw), nch, n_outer, self.n_in,
                                 n_inner, self.W.data_ptr(), self.start.data_ptr(), self.taps, self.n_out,
                                 float(alpha), int(bool(accumulate)), out.data_ptr(), stream)
        _lib.check(rc)
        return out


@lru_cache(maxsize=64)
def _dense(kind: str, n: int, dz: float, amp: float, f0: float, k: float) -> np.ndarray:
    if kind == "gradient":
        return -gradient_matrix(n, abs(dz))
    if kind == "amp9":
        return amp9_matrix(n, dz, amp)
    if kind == "central":
        return central_matrix(n, dz)
    if kind == "central_seed":
        return central_seed_matrix(n, dz)
    if kind == "giessibl":
        return giessibl_matrix(n, dz, amp, f0, k)
    raise ValueError(f"unknown df kind {kind!r}")


_ops: dict = {}


def z_operator(kind: str, n: int, dz: float, device, amp: float = 1.0, f0: float = 30000.0,
               k: float = 1800.0, transpose: bool = False) -> AxisOperator:
    key = (kind, n, float(dz), float(amp), float(f0), float(k), str(device), bool(transpose))
    op = _ops.get(key)
    if op is None:
        M = _dense(kind, int(n), float(dz), float(amp), float(f0), float(k))
        op = AxisOperator(M.T.copy() if transpose else M, device)
        _ops[key] = op
    return op


def apply_z(kind: str, FEs, dz: float, amp: float = 1.0, f0: float = 30000.0, k: float = 1800.0, channel: int = 2):
    """df volume (…, nz) from a device tensor FEs (…, nz, 4) (channel 2 = Fz) or Fz (…, nz)."""
    import torch
    if FEs.dtype != torch.float32 or not FEs.is_cuda:
        raise ValueError("apply_z expects a float32 CUDA tensor")
    FEs = FEs.contiguous()
    if FEs.dim() >= 2 and FEs.shape[-1] == 4 and channel is not None:
        nz = FEs.shape[-2]
        lead = FEs.shape[:-2]
        stride, off = 4, channel
    else:
        nz = FEs.shape[-1]
        lead = FEs.shape[:-1]
        stride, off = 1, 0
    npix = int(np.prod(lead)) if len(lead) else 1
    op = z_operator(kind, nz, dz, FEs.device, amp, f0, k)
    out = op.apply(FEs, (nz * stride, stride, 0), npix, 1, offset=off)
    return out.reshape(*lead, nz)
