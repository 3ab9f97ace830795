"""Autotuner of the batched forward's launch shape.

Candidates: grid voxels per thread along z (4, 8), the pixel tile of a wavefront (16x4, 8x8,
4x16 pixels), persistent workgroups per compute unit (2, 3, 4), the schedule (one fused
launch, or a grid launch followed by a relaxation launch) and, for the fused launch, the queue
lag (0: each molecule's relaxation blocks right after its tiles; k: after the tiles of molecule
m + k).  Each candidate runs the real batch
(or a prefix of it) and is timed with HIP events on the launch stream; every candidate's df
must equal the first candidate's bit for bit (they compute the same arithmetic, only the
work decomposition differs).  The winner is cached in a JSON file keyed by device name and a
signature of the batch (number of molecules, pixels, z-steps, projected pair evaluations).

Cache file: $SPAMMM_AMD_TUNE_CACHE or ~/.cache/spammm_amd/autotune.json.
"""
from __future__ import annotations

import itertools
import json
import os

from .batch import BatchForward, LaunchConfig

CANDIDATES = {
    "vz": (4, 8),
    "tile_w": (4, 8, 16),
    "workers_per_cu": (2, 3, 4),
    "schedule": ("fused", "two_phase"),
    "lag": (0, 2, 8),
}


def cache_path() -> str:
    p = os.environ.get("SPAMMM_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "spammm_amd", "autotune.json")


def signature(batch: BatchForward) -> str:
    import torch
    name = torch.cuda.get_device_properties(batch.device).name
    pairs = batch.pair_evals
    return f"{name}|nmol{len(batch.molecules)}|pix{batch.n_pix}|nz{batch.nz}|pairs{pairs:.3e}"


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(data):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def time_config(batch: BatchForward, cfg: LaunchConfig, reps: int = 2) -> float:
    """Best-of-reps milliseconds of one batch launch with ``cfg``."""
    import torch
    stream = torch.cuda.current_stream(batch.device)
    best = float("inf")
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        batch.launch(cfg)
        b.record(stream)
        torch.cuda.synchronize(batch.device)
        best = min(best, a.elapsed_time(b))
    return best


def tune(batch: BatchForward, candidates=None, reps: int = 2, use_cache: bool = True, log=None):
    """Return (best LaunchConfig, {config key: ms}) for this batch, using / filling the cache."""
    import torch
    sig = signature(batch)
    data = _load() if use_cache else {}
    if use_cache and sig in data:
        c = data[sig]["best"]
        return LaunchConfig(**c), data[sig].get("timings", {})
    cand = candidates or CANDIDATES
    keys = list(cand)
    results = {}
    ref = None
    best, best_ms = None, float("inf")
    for vals in itertools.product(*(cand[k] for k in keys)):
        cfg = LaunchConfig(**dict(zip(keys, vals)))
        if cfg.schedule == "two_phase" and cfg.lag:
            continue        # the lag orders fused work only (two phases separate the kinds anyway)
        ms = time_config(batch, cfg, reps)
        err, _ = batch.status()
        if err:
            raise RuntimeError(f"autotune: launch {cfg.key()} reported a wait timeout")
        out = batch.df.clone() if batch.df is not None else batch.FEs.clone()
        if ref is None:
            ref = out
        elif not torch.equal(ref, out):
            raise RuntimeError(f"autotune: launch {cfg.key()} changed the results")
        results[cfg.key()] = ms
        if log:
            log(f"[autotune] {cfg.key()}: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = cfg, ms
    if use_cache:
        data[sig] = {"best": best.__dict__, "best_ms": best_ms, "timings": results}
        _store(data)
    return best, results


def refine(batch: BatchForward, cfg: LaunchConfig, lags=(32, 64), reps: int = 2, log=None):
    """The tuned shape ``cfg`` checked on the whole ``batch``: the same vz, tile and workgroups per
    CU as two-phase and as fused with each queue lag in ``lags`` (lags longer than the tuning
    sample's molecule count; a relaxation block then finds its tiles done instead of waiting), each
    timed on the whole batch; returns (the fastest, {key: ms}).  Every variant must give the first
    one's results bit for bit."""
    import torch
    variants = [cfg, LaunchConfig(cfg.vz, cfg.tile_w, cfg.workers_per_cu, "two_phase", 0)]
    variants += [LaunchConfig(cfg.vz, cfg.tile_w, cfg.workers_per_cu, "fused", int(lag)) for lag in lags
                 if int(lag) < len(batch.molecules)]
    seen, results, ref = set(), {}, None
    best, best_ms = cfg, float("inf")
    for v in variants:
        if v.key() in seen:
            continue
        seen.add(v.key())
        ms = time_config(batch, v, reps)
        err, _ = batch.status()
        if err:
            raise RuntimeError(f"autotune: launch {v.key()} reported a wait timeout")
        out = batch.df.clone() if batch.df is not None else batch.FEs.clone()
        if ref is None:
            ref = out
        elif not torch.equal(ref, out):
            raise RuntimeError(f"autotune: launch {v.key()} changed the results")
        results[v.key()] = ms
        if log:
            log(f"[autotune] whole batch {v.key()}: {ms:.3f} ms")
        if ms < best_ms:
            best, best_ms = v, ms
    return best, results
