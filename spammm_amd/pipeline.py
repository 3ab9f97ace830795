"""Consecutive plain scans on two HIP streams.

A plain scan of one image is a grid projection (evalLJC_QZs_toImg, kernels/AFM.cl:916-991) followed
by the relaxed scan over that grid (relaxStrokesTilted, kernels/AFM.cl:478-562) with the df operator
applied to each Fz column in the same launch.  On one stream the relaxation runs alone on the
machine: one scan point is a serial chain of FIRE steps, so a 256x256 raster (two lanes per point)
fills about two waves per SIMD, and most VALU issue slots sit idle while those chains wait on their
own results.  The grid projection of the next image has no such wait.  ``ScanPipeline`` runs the images of a sequence
on ``depth`` streams, image k on stream k % depth with its own grid and output buffers: the grid
projections run one after another (each waits for the previous image's projection), and image k's
relaxation runs beside image k+1's projection.  Every image gives the same bits as a scan of it
alone (the kernels and their inputs are the same; only the interleaving differs).

The reference runs one image per blocking call (AFMulator.run_scan, spammm/SPM/AFM.py:950-1010,
ends in queue.finish()); this is the stream-ordered form of a loop of such calls.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import _lib
from . import device as D


@dataclass
class ScanSlot:
    """The stream (device.Stream), buffers and events of one in-flight image: ``grid`` (nz,ny,nx,4),
    ``FEs``/``PPs`` (n_scan,nz,4), ``df`` (n_scan,n_out) or None.  Valid once ``done`` has completed;
    reused by the image submitted ``depth`` images later."""
    stream: object
    grid: object
    FEs: object
    PPs: object
    df: object
    grid_done: object = None
    done: object = None
    events: tuple = field(default_factory=tuple)
    index: int = -1

    def wait(self, stream=None):
        """Make ``stream`` (default: the current stream) wait for this image's outputs."""
        import torch
        (stream or torch.cuda.current_stream(self.grid.device)).wait_event(self.done)


class ScanPipeline:
    """Plain scans of a sequence of atom sets that share one grid geometry, scan raster and tip.

    spec: device.GridSpec; points: (n_scan,4) float32 device tensor; nz: z-steps per point;
    tip: device.Tip; model: _lib.PPM_MODEL_LJ or _lib.PPM_MODEL_MORSE; Qs, QZs, npbc, lvA, lvB, cap:
    as device.grid_project; max_iter, f2conv, check_before, scan_ny, damped: as device.relax_tilted;
    df_op: a df.AxisOperator over nz slices (or None: no df); depth: buffer sets and streams (>= 2);
    lanes: lanes per scan point of the relaxation (default 1: beside a grid projection the
    relaxation's latency is hidden and only its instruction count matters, and one lane per point
    issues the fewest; 0 = the library's choice for a launch alone on the GPU).
    """

    def __init__(self, spec, points, nz, tip, model=_lib.PPM_MODEL_LJ, Qs=(0, 0, 0, 0), QZs=(0, 0, 0, 0),
                 npbc=(0, 0), lvA=(0, 0, 0), lvB=(0, 0, 0), cap=True, max_iter=128, f2conv=1e-8, check_before=False,
                 scan_ny=0, damped=False, df_op=None, depth=2, want_pps=True, lanes=1):
        import torch
        if depth < 2:
            raise ValueError("a pipeline needs at least two buffer sets")
        D._f32dev(points, "points", 4)
        self.dev = points.device
        self.spec, self.points, self.nz, self.tip = spec, points, int(nz), tip
        self.model, self.Qs, self.QZs, self.npbc, self.lvA, self.lvB, self.cap = model, Qs, QZs, npbc, lvA, lvB, cap
        self.relax = dict(max_iter=max_iter, f2conv=f2conv, check_before=check_before, scan_ny=scan_ny, damped=damped,
                          lanes=lanes)
        self.df_op = df_op
        if df_op is not None and int(df_op.n_in) != self.nz:
            raise ValueError("df_op must act on nz slices")
        ns = int(points.shape[0])
        nx, ny, ng = (int(v) for v in spec.n)
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.slots = []
        for i in range(depth):
            # streams of the library's own (device.Stream: each takes the least used hardware queue;
            # default priority); a slot on the caller's stream would make the next image's inputs
            # wait for this image's relaxation
            self.slots.append(ScanSlot(
                stream=D.Stream(self.dev), grid=torch.empty((ng, ny, nx, 4), **f32),
                FEs=torch.empty((ns, self.nz, 4), **f32),
                PPs=torch.empty((ns, self.nz, 4), **f32) if want_pps else None,
                df=torch.empty((ns, int(df_op.n_out)), **f32) if df_op is not None else None))
        self.count = 0
        self._last_grid = None
        self._inputs = [None] * depth

    def submit(self, atoms, cpar, record=False) -> ScanSlot:
        """Queue the scan of one atom set (atoms (na,4) x,y,z,q and cpar on the device, produced on the
        current stream); returns its slot.  record=True adds timing events (grid start, grid end,
        relaxation end) in ``slot.events``."""
        import torch
        slot = self.slots[self.count % len(self.slots)]
        if slot.stream.handle is None:
            raise RuntimeError("the pipeline is closed")
        cur = torch.cuda.current_stream(self.dev)
        s = slot.stream.torch
        i = self.count % len(self.slots)
        if self._inputs[i] is not None:
            cur.wait_event(slot.done)              # the previous inputs of this slot may be freed after it
        self._inputs[i] = (atoms, cpar)            # kept alive while their image is in flight
        s.wait_stream(cur)                         # inputs were written on the caller's stream
        with torch.cuda.stream(s):
            if self._last_grid is not None:
                s.wait_event(self._last_grid)      # grid projections run one at a time
            ev = tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) if record else ()
            if record:
                ev[0].record(s)
            D.grid_project(atoms, cpar, self.model, self.spec, self.Qs, self.QZs, self.npbc, self.lvA, self.lvB,
                           self.cap, out=slot.grid)
            slot.grid_done = ev[1] if record else torch.cuda.Event()
            slot.grid_done.record(s)
            D.relax_tilted(slot.grid, self.spec, self.points, self.nz, self.tip, FEs=slot.FEs, PPs=slot.PPs,
                           want_pps=slot.PPs is not None, df_op=self.df_op, df=slot.df, **self.relax)
            slot.done = ev[2] if record else torch.cuda.Event()
            slot.done.record(s)
        self._last_grid = slot.grid_done
        slot.events, slot.index = ev, self.count
        self.count += 1
        return slot

    def synchronize(self):
        """Make the current stream wait for every queued image."""
        import torch
        cur = torch.cuda.current_stream(self.dev)
        for sl in self.slots:
            if sl.done is not None:
                cur.wait_event(sl.done)

    def close(self):
        """Make the current stream wait for every queued image and release the streams (the slots'
        buffers stay valid)."""
        for sl in self.slots:
            sl.stream.close()
        self._inputs = [None] * len(self.slots)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
