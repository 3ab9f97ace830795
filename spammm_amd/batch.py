"""Batched forward: hundreds of molecules through grid -> relax -> df in one persistent launch.

``BatchForward`` packs molecules of uneven atom counts and cell sizes into one set of device
buffers and one descriptor array (``PpmMolDesc``, include/ppm_api.h), then runs
``ppm_batch_forward``: persistent workgroups take grid tiles and relaxation blocks from one
atomic work queue (molecule by molecule), each lane relaxes one pixel through its z descent
and applies the df operator to its own Fz column.  The launch can be captured in a hipGraph
(``capture``/``replay``).  Launch shape (voxels per thread, pixel tile per wavefront,
workgroups per CU, fused or two-launch schedule) is chosen by ``spammm_amd.autotune``.

Only the z planes a probe can reach are projected: the planes covering the probe rest heights
of the scan (tip anchor + dpos0) widened by ``z_margin`` A below and above (default 0.3, 0.5:
over 48 molecules of the batched configuration the probes reached at most 0.34 A above the
highest and 0.03 A below the lowest rest height).  A probe that nevertheless samples outside
them flags its molecule, and ``forward`` recomputes flagged molecules on the plain per-image
path with the whole grid, so every output equals the plain path's.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib, df as DF
from . import device as D

_F4 = ctypes.c_float * 4


class PpmMolDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "atom_off", "na", "nx", "ny", "nz", "klo", "nzw", "npbc_a", "npbc_b", "pix_off", "scan_nx", "scan_ny",
        "item_off", "n_grid_items", "n_relax_items", "relax_item_off")] + [("grid_off", ctypes.c_longlong)] + \
        [(n, _F4) for n in ("p0", "dA", "dB", "dC", "dinvA", "dinvB", "dinvC", "lvA", "lvB", "scan_p0", "scan_da",
                            "scan_db")]


class PpmBatchParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("model", "cap", "nz", "nmol", "nitems", "vz", "tile_w", "mode",
                                             "df_taps", "df_nz_out")] + \
        [("reserved", ctypes.c_int * 2), ("Qs", _F4), ("QZs", _F4), ("tip", _lib.PpmTip),
         ("ctl", _lib.PpmRelaxCtl)]


MODES = {"fused": 0, "grid_only": 1, "relax_only": 2}


@dataclass
class LaunchConfig:
    vz: int = 8                 # grid voxels per thread along z
    tile_w: int = 8             # pixels along iy of a wavefront's 64-pixel tile
    workers_per_cu: int = 3     # persistent workgroups per compute unit
    schedule: str = "fused"     # "fused" (one launch) or "two_phase" (grid launch, then relax launch)
    lag: int = 0                # queue order: the tiles of molecule m + lag come before the blocks of m

    def key(self):
        k = f"vz{self.vz}_tw{self.tile_w}_w{self.workers_per_cu}_{self.schedule}"
        return k + (f"_lag{self.lag}" if self.lag else "")


def z_window(w, sp, tip, nz, margin):
    """Stored z planes [klo, klo+nzw) covering every probe rest height of the scan widened by
    ``margin`` A (a number, or (below, above)); rest height = tip anchor + tip-frame dpos0; one
    extra plane below and two above for the trilinear neighbours.  The whole grid when the
    window would not be smaller."""
    m_lo, m_hi = (float(margin), float(margin)) if np.isscalar(margin) else (float(margin[0]), float(margin[1]))
    A, B, C = (np.asarray(v, np.float64)[:3] for v in (tip.tipA, tip.tipB, tip.tipC))
    d0 = np.asarray(tip.dpos0, np.float64)[:3]
    d0r = A * d0[0] + B * d0[1] + C * d0[2]
    dTip = C * float(tip.tipC[3])
    z0 = float(w.scan_p0[2]) + d0r[2]
    z1 = z0 + (nz - 1) * dTip[2]
    zlo, zhi = min(z0, z1) - m_lo, max(z0, z1) + m_hi
    dz = float(sp.dC[2])
    nzg = int(sp.n[2])
    if zhi < zlo:
        zlo = zhi = 0.5 * (zlo + zhi)
    klo = int(math.floor((zlo - float(sp.p0[2])) / dz)) - 1
    khi = int(math.ceil((zhi - float(sp.p0[2])) / dz)) + 2
    klo, khi = max(0, klo), min(nzg, khi)
    if khi - klo >= nzg or khi <= klo:
        return 0, nzg
    return klo, khi - klo


class BatchForward:
    """Grid projection + relaxed scan + df for a list of molecules (``spammm_amd.workloads.Workload``)."""

    def __init__(self, molecules, device=None, model=_lib.PPM_MODEL_LJ, cpars=None, tip=None, max_iter=100,
                 f2conv=1e-8, check_before=False, df_kind="giessibl", amp=1.0, f0=30000.0, k_cant=1800.0,
                 z_margin=(0.3, 0.5), cap=True, want_pps=True):
        import torch
        self.device = D.require_gpu(device)
        mols = list(molecules)
        if not mols:
            raise ValueError("empty batch")
        w0 = mols[0]
        self.nz = int(w0.nz)
        self.dtip = float(w0.dtip)
        for w in mols:
            if int(w.nz) != self.nz or float(w.dtip) != self.dtip:
                raise ValueError("every molecule of a batch must use the same z descent (nz, dtip)")
            if not (np.array_equal(w.Qs, w0.Qs) and np.array_equal(w.QZs, w0.QZs)):
                raise ValueError("every molecule of a batch must use the same probe charges")
        self.molecules = mols
        self.z_margin = (float(z_margin), float(z_margin)) if np.isscalar(z_margin) else tuple(map(float, z_margin))
        self.model = int(model)
        self.cap = bool(cap)
        self.tip = tip if tip is not None else D.Tip(stiffness=w0.stiffness, dpos0=w0.dpos0,
                                                     tipC=np.array([0., 0., 1., self.dtip], np.float32))
        self.max_iter, self.f2conv, self.check_before = int(max_iter), float(f2conv), bool(check_before)
        self.Qs, self.QZs = np.asarray(w0.Qs, np.float32), np.asarray(w0.QZs, np.float32)
        cs = 2 if self.model == _lib.PPM_MODEL_LJ else 4
        if cpars is None:
            if self.model != _lib.PPM_MODEL_LJ:
                raise ValueError("Morse batches need explicit cpars")
            cpars = [w.cLJ for w in mols]
        self.specs = [D.axis_aligned_spec(w.p0, w.L, w.n) for w in mols]
        descs = (PpmMolDesc * len(mols))()
        atom_off = pix_off = item_off = 0
        grid_off = 0
        self.windows = []
        for i, (w, sp, cp) in enumerate(zip(mols, self.specs, cpars)):
            cp = np.asarray(cp, np.float32)
            if cp.shape != (w.atoms.shape[0], cs):
                raise ValueError(f"molecule {i}: parameter rows must be (na,{cs})")
            if sp.dC[0] != 0 or sp.dC[1] != 0:
                raise ValueError("batched grids need dC along z")
            nx, ny, nzg = (int(v) for v in sp.n)
            klo, nzw = self._window(w, sp, z_margin)
            self.windows.append((klo, nzw))
            d = descs[i]
            d.atom_off, d.na = atom_off, int(w.atoms.shape[0])
            d.nx, d.ny, d.nz = nx, ny, nzg
            d.klo, d.nzw = klo, nzw
            d.npbc_a, d.npbc_b = int(w.npbc[0]), int(w.npbc[1])
            d.pix_off, d.scan_nx, d.scan_ny = pix_off, int(w.nxy[0]), int(w.nxy[1])
            d.item_off = item_off
            vz_max = 4   # grid items are counted for the smallest vz so every launch shape fits
            d.n_grid_items = ((nx * ny + 255) // 256) * ((nzw + vz_max - 1) // vz_max)
            d.n_relax_items = 0          # set per launch shape in _set_items
            d.grid_off = grid_off
            d.p0, d.dA, d.dB, d.dC = (_lib.f4(v) for v in (sp.p0, sp.dA, sp.dB, sp.dC))
            d.dinvA, d.dinvB, d.dinvC = (_lib.f4(v) for v in (sp.dinvA, sp.dinvB, sp.dinvC))
            d.lvA, d.lvB = _lib.f4(w.lvA), _lib.f4(w.lvB)
            d.scan_p0, d.scan_da, d.scan_db = _lib.f4(w.scan_p0), _lib.f4(w.scan_da), _lib.f4(w.scan_db)
            atom_off += d.na
            pix_off += d.scan_nx * d.scan_ny
            grid_off += nx * ny * nzw
        self.descs_host = descs
        self.n_pix = pix_off
        self.grid_floats = 4 * grid_off
        dev = self.device
        self.atoms = torch.as_tensor(np.concatenate([w.atoms for w in mols]).astype(np.float32), device=dev)
        self.cpar = torch.as_tensor(np.concatenate([np.asarray(c, np.float32) for c in cpars]), device=dev)
        self.grids = torch.empty(self.grid_floats, dtype=torch.float32, device=dev)
        self.FEs = torch.empty((self.n_pix, self.nz, 4), dtype=torch.float32, device=dev)
        self.PPs = torch.empty((self.n_pix, self.nz, 4), dtype=torch.float32, device=dev) if want_pps else None
        self.df_kind, self.amp, self.f0, self.k_cant = df_kind, float(amp), float(f0), float(k_cant)
        self.dfop = DF.z_operator(df_kind, self.nz, self.dtip, dev, amp, f0, k_cant) if df_kind else None
        self.df = (torch.empty((self.n_pix, self.dfop.n_out), dtype=torch.float32, device=dev)
                   if self.dfop is not None else None)
        L = _lib.load()
        self.ws_bytes = int(L.ppm_batch_workspace_size(len(mols)))
        self.workspace = torch.zeros(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.status_buf = torch.zeros(1 + len(mols), dtype=torch.int32, device=dev)   # allocation-free launches
        self._last_noalloc = False
        self.descs_dev = torch.empty(ctypes.sizeof(descs), dtype=torch.uint8, device=dev)
        self.n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
        self._items_for = None
        self.graph = None

    # ── geometry ──────────────────────────────────────────────────────────────

    def _window(self, w, sp, margin):
        return z_window(w, sp, self.tip, self.nz, margin)

    @property
    def pair_evals(self):
        """Pair evaluations of the projected windows."""
        tot = 0
        for w, (klo, nzw) in zip(self.molecules, self.windows):
            nimg = (2 * w.npbc[0] + 1) * (2 * w.npbc[1] + 1)
            tot += int(w.n[0]) * int(w.n[1]) * nzw * w.atoms.shape[0] * nimg
        return tot

    @property
    def states(self):
        return self.n_pix * self.nz

    def _set_items(self, cfg: LaunchConfig):
        """Work items of every molecule and the queue order of their segments: the grid tiles of
        molecule m, then (lag = k) the relaxation blocks of molecule m - k, so that a block of m is
        dequeued k molecules' tiles after the last tile of m (it rarely waits)."""
        key = (cfg.vz, cfg.tile_w, cfg.lag)
        if self._items_for == key:
            return
        if cfg.lag < 0:
            raise ValueError("lag must be >= 0")
        th = 64 // cfg.tile_w
        for d in self.descs_host:
            d.n_grid_items = ((d.nx * d.ny + 255) // 256) * ((d.nzw + cfg.vz - 1) // cfg.vz)
            tiles = ((d.scan_nx + th - 1) // th) * ((d.scan_ny + cfg.tile_w - 1) // cfg.tile_w)
            d.n_relax_items = (tiles + 3) // 4
        n = len(self.descs_host)
        order = []
        for m in range(n):
            order.append(("grid", m))
            if m - cfg.lag >= 0:
                order.append(("relax", m - cfg.lag))
        order += [("relax", m) for m in range(max(0, n - cfg.lag), n)]
        item = 0
        for kind, m in order:
            d = self.descs_host[m]
            if kind == "grid":
                d.item_off = item
                item += d.n_grid_items
            else:
                d.relax_item_off = item
                item += d.n_relax_items
        self.nitems = item
        raw = np.frombuffer(ctypes.string_at(ctypes.addressof(self.descs_host), ctypes.sizeof(self.descs_host)),
                            dtype=np.uint8)
        import torch
        self.descs_dev.copy_(torch.as_tensor(raw.copy()))
        self._items_for = key

    def _params(self, cfg: LaunchConfig, mode: int) -> PpmBatchParams:
        P = PpmBatchParams()
        P.model, P.cap, P.nz = self.model, int(self.cap), self.nz
        P.nmol, P.nitems = len(self.molecules), self.nitems
        P.vz, P.tile_w, P.mode = cfg.vz, cfg.tile_w, mode
        if self.dfop is not None:
            P.df_taps, P.df_nz_out = self.dfop.taps, self.dfop.n_out
        P.Qs, P.QZs = _lib.f4(self.Qs), _lib.f4(self.QZs)
        P.tip = self.tip.c()
        P.ctl = D.relax_ctl(self.max_iter, self.f2conv, self.check_before, 0)
        return P

    # ── launches ──────────────────────────────────────────────────────────────

    def launch(self, cfg: LaunchConfig = LaunchConfig(), noalloc: bool = False):
        """Enqueue the batch on the current stream (no synchronization).  noalloc: through
        ppm_batch_forward_noalloc (the library's scratch instead of this batch's workspace; the
        status lands in status_buf)."""
        import torch
        self._set_items(cfg)
        L = _lib.load()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        workers = max(1, int(cfg.workers_per_cu) * self.n_cu)
        df_ptr = self.df.data_ptr() if self.df is not None else None
        W = self.dfop.W.data_ptr() if self.dfop is not None else None
        S = self.dfop.start.data_ptr() if self.dfop is not None else None
        pp = self.PPs.data_ptr() if self.PPs is not None else None
        bufs = (self.atoms.data_ptr(), self.cpar.data_ptr(), self.grids.data_ptr(), self.FEs.data_ptr(), pp, df_ptr,
                W, S)
        if noalloc:
            def call(P, reset):
                return L.ppm_batch_forward_noalloc(self.descs_dev.data_ptr(), ctypes.addressof(P), *bufs,
                                                   self.status_buf.data_ptr(), workers, reset, stream)
        else:
            def call(P, reset):
                return L.ppm_batch_forward(self.descs_dev.data_ptr(), ctypes.addressof(P), *bufs,
                                           self.workspace.data_ptr(), self.ws_bytes, workers, reset, stream)
        if cfg.schedule == "fused":
            _lib.check(call(self._params(cfg, MODES["fused"]), 1))
        elif cfg.schedule == "two_phase":
            _lib.check(call(self._params(cfg, MODES["grid_only"]), 1))
            _lib.check(call(self._params(cfg, MODES["relax_only"]), 2))
        else:
            raise ValueError(f"unknown schedule {cfg.schedule!r}")
        self._last_noalloc = bool(noalloc)

    def status(self):
        """(error word, per-molecule window flags) of the last launch (synchronizes)."""
        import torch
        n = len(self.molecules)
        if self._last_noalloc:
            st = self.status_buf.cpu().numpy()
            return int(st[0]), st[1:1 + n].copy()
        ws = self.workspace.view(torch.int32).cpu().numpy()
        return int(ws[1]), ws[4 + n:4 + 2 * n].copy()

    def recompute_plain(self, indices, FEs=None, df=None, PPs=None, with_pps=True):
        """Plain per-image path (whole grid) for the given molecules, written into the batch outputs
        (or into FEs / df / PPs of the same layout; with_pps=False writes no PPs)."""
        import torch
        FEs = self.FEs if FEs is None else FEs
        df = self.df if df is None else df
        PPs = (self.PPs if PPs is None else PPs) if with_pps else None
        for i in indices:
            w, sp, d = self.molecules[i], self.specs[i], self.descs_host[i]
            a0, a1 = d.atom_off, d.atom_off + d.na
            G = D.grid_project(self.atoms[a0:a1].contiguous(), self.cpar[a0:a1].contiguous(), self.model, sp,
                               self.Qs, self.QZs, w.npbc, w.lvA, w.lvB, cap=self.cap)
            pts = torch.as_tensor(D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy), device=self.device)
            p0, p1 = d.pix_off, d.pix_off + d.scan_nx * d.scan_ny
            D.relax_tilted(G, sp, pts, self.nz, self.tip, self.max_iter, self.f2conv, self.check_before,
                           scan_ny=d.scan_ny, FEs=FEs[p0:p1], PPs=None if PPs is None else PPs[p0:p1],
                           want_pps=PPs is not None)
            if self.dfop is not None and df is not None:
                self.dfop.apply(FEs, (self.nz * 4, 4, 0), p1 - p0, 1, offset=2 + p0 * self.nz * 4,
                                out=df[p0:p1].view(p1 - p0, self.dfop.n_out, 1))

    def forward(self, cfg: LaunchConfig = LaunchConfig()):
        """Run the batch; recompute molecules whose probes left the projected window; return df."""
        import torch
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch(cfg)
        err, flags = self.status()
        if err:
            raise RuntimeError("batched forward: a relaxation block timed out waiting for its grid")
        bad = np.flatnonzero(flags)
        if bad.size:
            self.recompute_plain(bad)
            torch.cuda.synchronize(self.device)
        self.last_fallback = bad
        return self.df

    def capture(self, cfg: LaunchConfig = LaunchConfig()):
        """Capture the batch launch in a hipGraph (torch.cuda.CUDAGraph); forward() then replays it."""
        import torch
        self._set_items(cfg)
        self.launch(cfg)                      # warm-up outside the capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch(cfg)
        self.graph = g
        return g

    # ── the plain per-image path (baseline) ───────────────────────────────────

    def plain_loop(self, FEs_out=None, df_out=None, windowed=False, streams=1):
        """The same work as a loop of plain per-image launches: returns (FEs, df).  windowed=False:
        whole grids; True: each molecule's grid projected and relaxed on the z window of its batch
        descriptor (ppm_grid_project_window + ppm_relax_strokes_tilted_window), the same pair
        evaluations as the batched launch; a molecule whose probes leave the window is recomputed
        with its whole grid (as the batched path does).  streams > 1: molecule i on stream
        i % streams (device.Stream) with a grid buffer of its own, the grid projections one at a
        time across the streams, so a molecule's relaxation runs beside the next one's projection
        (the two-stream form of pipeline.ScanPipeline for molecules of different grids)."""
        import contextlib
        import torch
        FEs = FEs_out if FEs_out is not None else torch.empty_like(self.FEs)
        df = df_out if df_out is not None else (torch.empty_like(self.df) if self.df is not None else None)
        if getattr(self, "_pts", None) is None:
            self._pts = [torch.as_tensor(D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy), device=self.device)
                         for w in self.molecules]
        if windowed:
            flags = torch.zeros(len(self.molecules), dtype=torch.int32, device=self.device)
        shapes = [(int(d.nzw) if windowed else int(sp.n[2]), int(sp.n[1]), int(sp.n[0]), 4)
                  for sp, d in zip(self.specs, self.descs_host)]
        cur = torch.cuda.current_stream(self.device)
        if streams > 1:
            lanes = [D.Stream(self.device) for _ in range(streams)]
            big = max(int(np.prod(s)) for s in shapes)
            bufs = [torch.empty(big, dtype=torch.float32, device=self.device) for _ in range(streams)]
            for ln in lanes:
                ln.torch.wait_stream(cur)
        last = None
        cache = {}
        for i, (w, sp, d) in enumerate(zip(self.molecules, self.specs, self.descs_host)):
            if streams > 1:
                s = lanes[i % streams].torch
                G = bufs[i % streams][:int(np.prod(shapes[i]))].view(shapes[i])
                ctx = torch.cuda.stream(s)
            else:
                if shapes[i] not in cache:
                    cache.clear()
                    cache[shapes[i]] = torch.empty(shapes[i], dtype=torch.float32, device=self.device)
                G = cache[shapes[i]]
                ctx = contextlib.nullcontext()
            a0, a1 = d.atom_off, d.atom_off + d.na
            win = (int(d.klo), int(d.nzw)) if windowed else None
            with ctx:
                if streams > 1 and last is not None:
                    s.wait_event(last)                 # projections one at a time
                D.grid_project(self.atoms[a0:a1], self.cpar[a0:a1], self.model, sp, self.Qs, self.QZs, w.npbc,
                               w.lvA, w.lvB, cap=self.cap, out=G, window=win)
                if streams > 1:
                    last = torch.cuda.Event()
                    last.record(s)
                pts = self._pts[i]
                p0, p1 = d.pix_off, d.pix_off + d.scan_nx * d.scan_ny
                D.relax_tilted(G, sp, pts, self.nz, self.tip, self.max_iter, self.f2conv, self.check_before,
                               scan_ny=d.scan_ny, FEs=FEs[p0:p1], want_pps=False, window=win,
                               flag=flags[i:i + 1] if windowed else None)
                if df is not None:
                    self.dfop.apply(FEs, (self.nz * 4, 4, 0), p1 - p0, 1, offset=2 + p0 * self.nz * 4,
                                    out=df[p0:p1].view(p1 - p0, self.dfop.n_out, 1))
        if streams > 1:
            for ln in lanes:
                ln.close()                             # the current stream waits for it
            del bufs
        if windowed:
            bad = np.flatnonzero(flags.cpu().numpy())
            self.last_plain_fallback = bad
            if bad.size:
                self.recompute_plain(bad, FEs=FEs, df=df, with_pps=False)
        return FEs, df
