"""Benchmark of the probe-particle AFM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload flake|naphthalene]

One step = one pass of the path over one synthetic image, inputs resident in HBM:
force-field grid projection (LJ + point charge, 9 periodic x,y images) -> relaxed
z-descent scan (FIRE, trilinear sampling, lateral + radial tip springs) -> Giessibl
df along z.  ``value`` = relaxed pixel*z-steps per second over K consecutive images run
on two streams (spammm_amd.pipeline: image k's relaxation beside image k+1's grid
projection, the projections one at a time); the same K steps on one stream are timed
first and reported under ``stages.one_stream``, and both give the same output bits.

Default workload (flake): 200-atom flat polycyclic flake, 30x30x24 A cell on a
0.1 A grid (300x300x240 float4), 256x256 pixels x 60 z-steps.

Prints ONE compact JSON line (under 4 KB: the headline, roofline, CPU baseline and a summary
per leg) and writes the full per-leg record to --detail (default bench_detail.json).
Roofline of the dominant kernel (grid projection) is
measured live with HIP events on the stream the kernels run on; the CPU baseline
times the C++ oracle (FP32 mode, OpenMP) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X published FP32
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="flake", choices=["flake", "naphthalene"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the accuracy leg")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable leg")
    ap.add_argument("--no-fd", action="store_true",
                    help="skip the finite-difference checks of the differentiable and substrate legs (timing only)")
    ap.add_argument("--adam-steps", type=int, default=50)
    ap.add_argument("--batch-mols", type=int, default=512, help="molecules of the batched leg (0: skip)")
    ap.add_argument("--no-direct", action="store_true", help="skip the gridless direct Morse leg")
    ap.add_argument("--poses", type=int, default=64, help="placements of the substrate leg (0: skip)")
    ap.add_argument("--no-fdbm", action="store_true", help="skip the FDBM force-field producer leg")
    ap.add_argument("--no-ewald", action="store_true", help="skip the 2D Ewald substrate electrostatics leg")
    ap.add_argument("--z-margin", default=None,
                    help="z window margins of the batched leg in A, 'below,above' (default: the library's)")
    ap.add_argument("--detail", default="bench_detail.json",
                    help="file for the full per-leg record (the printed line carries summaries)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device-to-device copy (scale of the HBM byte counters under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (every box has one MI355X); "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import torch
    from spammm_amd import device as D, df as DF, workloads as WL
    from spammm_amd import _lib

    dev = D.require_gpu()
    w = WL.flake() if args.workload == "flake" else WL.naphthalene()
    sp = D.axis_aligned_spec(w.p0, w.L, w.n)
    atoms = torch.as_tensor(w.atoms, device=dev)
    cpar = torch.as_tensor(w.cLJ, device=dev)
    pts_h = D.scan_points(w.scan_p0, w.scan_da, w.scan_db, *w.nxy)
    pts = torch.as_tensor(pts_h, device=dev)
    tip = D.Tip(stiffness=w.stiffness, dpos0=w.dpos0, tipC=np.array([0., 0., 1., w.dtip], np.float32))
    nz = w.nz
    grid = torch.empty((w.n[2], w.n[1], w.n[0], 4), dtype=torch.float32, device=dev)
    FEs = torch.empty((w.n_scan, nz, 4), dtype=torch.float32, device=dev)
    PPs = torch.empty((w.n_scan, nz, 4), dtype=torch.float32, device=dev)
    dfop = DF.z_operator("giessibl", nz, w.dtip, dev, w.amp, w.f0, w.k_cant)
    df = torch.empty((w.n_scan, nz, 1), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    ev = {k: [] for k in ("grid", "relax")}
    df2 = df.view(w.n_scan, dfop.n_out)

    def step(record=False):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)] if record else None
        if record:
            e[0].record(stream)
        D.grid_project(atoms, cpar, _lib.PPM_MODEL_LJ, sp, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB, True, out=grid)
        if record:
            e[1].record(stream)
        # relaxation with the df operator applied to each Fz column in the same launch
        D.relax_tilted(grid, sp, pts, nz, tip, w.max_iter, w.f2conv, False, w.nxy[1], FEs=FEs, PPs=PPs,
                       df_op=dfop, df=df2)
        if record:
            e[2].record(stream)
            ev["grid"].append((e[0], e[1]))
            ev["relax"].append((e[1], e[2]))

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] copy calibration: 3 x 1 GiB device copies")
    log(f"[bench] workload={w.name} grid={w.n} atoms={w.atoms.shape[0]} scan={w.nxy}x{nz} warmup={args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(record=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps

    def avg_ms(k):
        return float(np.mean([a.elapsed_time(b) for a, b in ev[k]]))

    t_grid, t_relax = avg_ms("grid"), avg_ms("relax")
    states = w.n_scan * nz

    # the headline: consecutive images on two streams (image k's relaxation beside image k+1's grid
    # projection; the projections run one at a time), the same kernels on the same inputs
    from spammm_amd.pipeline import ScanPipeline
    pipe = ScanPipeline(sp, pts, nz, tip, _lib.PPM_MODEL_LJ, w.Qs, w.QZs, w.npbc, w.lvA, w.lvB, True, w.max_iter,
                        w.f2conv, False, w.nxy[1], df_op=dfop)
    for _ in range(args.warmup):
        pipe.submit(atoms, cpar)
    torch.cuda.synchronize()
    pev = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pev.append(pipe.submit(atoms, cpar, record=True).events)
    torch.cuda.synchronize()
    wall2 = time.perf_counter() - t0
    p_grid = float(np.mean([a.elapsed_time(b) for a, b, _ in pev]))
    p_relax = float(np.mean([b.elapsed_time(c) for _, b, c in pev]))
    p_same = all(torch.equal(sl.df, df2) and torch.equal(sl.PPs, PPs) and torch.equal(sl.FEs, FEs)
                 for sl in pipe.slots)
    del pipe
    torch.cuda.empty_cache()
    log(f"[bench] two streams: {1e3 * wall2 / args.steps:.3f} ms per image (grid {p_grid:.3f} ms, relax+df "
        f"{p_relax:.3f} ms per launch); outputs equal to the one-stream step: {p_same}")
    value = states / (wall2 / args.steps)
    ms_one = ms_step
    ms_step = 1e3 * wall2 / args.steps
    pp = PPs.cpu().numpy()
    mean_iter = float(np.abs(pp[..., 3]).mean())
    conv_frac = float((pp[..., 3] > 0).mean())
    pairs = w.pair_evals
    flop = pairs * 52
    grid_tflops = flop / (p_grid * 1e-3) / 1e12
    grid_tflops_one = flop / (t_grid * 1e-3) / 1e12
    relax_bytes = states * (128.0 * (mean_iter + 1.0) + 32.0)
    relax_gbs = relax_bytes / (t_relax * 1e-3) / 1e9
    log(f"[bench] one stream: step {ms_one:.3f} ms (grid {t_grid:.3f} relax+df {t_relax:.3f}); "
        f"mean iters {mean_iter:.2f} converged {conv_frac:.4f}")

    out = {
        "metric": "relaxed pixel*z-steps/s on 1 MI355X (batched vs plain); adjoint/forward ratio",
        "value_is": "flake grid + relaxed scan + Giessibl df per image, consecutive images on two streams",
        "value": value,
        "unit": "pixel*z-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"{w.name}: {w.atoms.shape[0]} atoms LJ+charge, {w.n[0]}x{w.n[1]}x{w.n[2]} float4 grid "
                         f"({2 * w.npbc[0] + 1}x{2 * w.npbc[1] + 1} periodic images), "
                         f"{w.nxy[0]}x{w.nxy[1]} pixels x {nz} z-steps, FIRE |F|<1e-4 <= {w.max_iter} it, "
                         f"Giessibl df A={w.amp} A"),
            "grid": list(w.n), "atoms": int(w.atoms.shape[0]), "pixels": list(w.nxy), "z_steps": nz,
        },
        "roofline": {
            "kernel": "k_grid_project (grid projection)",
            "bound": "valu",
            "compute_unit": "FP32 VALU, no matrix instructions; peak = published MI355X FP32",
            "achieved": grid_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": grid_tflops / PEAK_FP32_TFLOPS,
            "traffic": _pmc_traffic("k_grid_project", w.name),
            "work_unit": f"{pairs} pair evaluations x 52 FLOP per launch",
            "avg_launch_ms": p_grid,
            # the same kernel with the machine to itself (one stream: no relaxation beside it)
            "avg_launch_ms_one_stream": t_grid,
            "frac_one_stream": grid_tflops_one / PEAK_FP32_TFLOPS,
        },
        "stages": {
            "one_stream": {"ms_per_step": ms_one, "value": states / (ms_one * 1e-3)},
            "two_streams": {"grid_ms": p_grid, "relax_ms": p_relax, "outputs_equal_one_stream": p_same},
            "grid_points_per_s": w.nvox / (t_grid * 1e-3),
            "grid_ms": t_grid,
            "relax_ms": t_relax,
            "relax_pixel_z_steps_per_s": states / (t_relax * 1e-3),
            "relax_mean_fire_iters": mean_iter,
            "relax_converged_fraction": conv_frac,
            # algorithmic bytes: 8 float4 corners per force evaluation + FE/PP write; the corners are
            # served mostly from registers and L2/MALL, so this is not HBM traffic (relax_hbm_bytes)
            "relax_algorithmic_GBps": relax_gbs,
            "relax_algorithmic_frac_of_hbm_peak": relax_gbs / PEAK_HBM_GBS,
            "relax_hbm_bytes_per_launch": _pmc_traffic("k_relax_tilted", w.name),
            "df": "fused into the relaxation launch (ppm_relax_strokes_tilted_df)",
        },
    }

    legs = {}
    if not args.no_diff:
        legs["differentiable"] = diff_leg(dev, args.adam_steps, fd=not args.no_fd)
    if args.batch_mols > 0:
        legs["batched"] = batch_leg(dev, args.batch_mols, args.z_margin)
    if not args.no_direct:
        legs["direct_morse"] = direct_leg(dev)
    if args.poses > 0:
        legs["substrate_poses"] = substrate_leg(dev, args.poses, fd=not args.no_fd)
    if not args.no_fdbm:
        legs["fdbm"] = fdbm_leg(dev)
    if not args.no_ewald:
        legs["ewald"] = ewald_leg(dev)
    if not args.no_check:
        legs["accuracy"] = accuracy_leg(w, sp, grid, pts_h, FEs, PPs, df, tip)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(w, sp, grid, pts_h, tip, states)
    detail = dict(out, **legs)
    out["detail_file"] = write_detail(detail, args.detail)
    out.update(summarize(legs))
    line = compact_line(out)
    log(f"[bench] result line: {len(line)} bytes")
    print(line, flush=True)


SUMMARY_KEYS = {
    "batched": ("batched_ms", "plain_loop_ms", "plain_windowed_loop_ms", "speedup", "speedup_vs_windowed_plain",
                "speedup_vs_plain_two_streams", "speedup_vs_windowed_plain_two_streams",
                "batched_pixel_z_steps_per_s", "launch", "autotune_fused_vs_two_phase_ms", "bitexact_vs_plain",
                "bitexact_windowed_plain", "window_fallback_molecules", "wait_timeout"),
    "differentiable": ("forward_ms", "adjoint_ms", "adjoint_over_forward", "fd_fp64_configured_heights",
                       "fd_fp64_smooth_heights", "gated_columns_configured_heights", "adam_configured_heights",
                       "adam_smooth_heights"),
    "substrate_poses": ("forward_ms_median", "adjoint_ms_median", "adjoint_over_forward",
                        "placements_per_s_forward_plus_adjoint", "batched_placements_per_s", "window_fallbacks",
                        "gated_states_total", "gated_columns_total", "directional_fd_rel_err"),
    "direct_morse": ("forward_ms", "vjp_ms", "vjp_over_forward", "pixel_z_steps_per_s"),
    "fdbm": ("stage3_device_ms", "grid_points_per_s_stage3"),
    "ewald": ("tip_grid_ms", "direct_3x3_coulomb_ms", "kept_g_pairs"),
    "accuracy": ("relax_bitexact_vs_fp32_oracle", "pp_err_vs_fp64_max_A", "pp_err_vs_fp64_rms_A",
                 "fz_err_vs_fp64_max", "fz_err_vs_fp64_rms", "df_max_rel_err_vs_fp64", "grid_err_over_term_bar_max"),
}


def summarize(legs):
    """One small object per leg: the figures the result line carries (the full record of every leg
    goes to the detail file)."""
    out = {}
    for name, res in legs.items():
        keys = SUMMARY_KEYS.get(name, ())
        out[name] = {k: res[k] for k in keys if k in res}
    return out


def _round_floats(x, sig=5):
    if isinstance(x, float):
        return float(f"{x:.{sig}g}")
    if isinstance(x, dict):
        return {k: _round_floats(v, sig) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_round_floats(v, sig) for v in x]
    return x


LINE_LIMIT = 4000
# leg summaries dropped first when the line would exceed LINE_LIMIT (the detail file keeps them)
DROP_ORDER = ("accuracy", "ewald", "fdbm", "direct_morse", "substrate_poses", "differentiable", "batched")


def compact_line(out):
    """The result line: floats to 5 significant digits, no whitespace, at most LINE_LIMIT bytes.
    The headline keys always stay; leg summaries are dropped in DROP_ORDER if the line would be
    longer (the record of every leg is in the detail file)."""
    out = _round_floats(out)
    line = json.dumps(out, separators=(",", ":"))
    dropped = []
    for k in DROP_ORDER:
        if len(line) <= LINE_LIMIT:
            break
        if k in out:
            out = {kk: v for kk, v in out.items() if kk != k}
            dropped.append(k)
            out["dropped_from_line"] = dropped
            line = json.dumps(out, separators=(",", ":"))
    if dropped:
        log(f"[bench] result line over {LINE_LIMIT} bytes: dropped {dropped} (kept in the detail file)")
    return line


def write_detail(detail, path):
    """Writes the full per-leg record as JSON; returns the path written (or None)."""
    if not path:
        return None
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(detail, f, indent=1, default=str)
        return path
    except OSError as e:
        log(f"[bench] detail file not written: {e}")
        return None


def batch_leg(dev, n_mol, z_margin=None):
    """Batched forward of n_mol molecules (U{10..150} atoms, square cells U(16,32) A x 20 A, 0.1 A grids,
    128x128 pixels, 40 z-steps) in one persistent launch replayed from a hipGraph, against the loop of
    plain per-image launches over the same molecules; outputs compared bit for bit."""
    import torch
    from spammm_amd import autotune, batch as B, workloads as WL
    mols = WL.batch(n_mol)
    t0 = time.perf_counter()
    kw = {} if z_margin is None else {"z_margin": tuple(float(v) for v in str(z_margin).split(","))}
    bf = B.BatchForward(mols, dev, want_pps=False, **kw)
    tune_sub = B.BatchForward(mols[:min(32, n_mol)], dev, want_pps=False, **kw)
    best, timings = autotune.tune(tune_sub, use_cache=False, reps=1)
    del tune_sub
    best, whole = autotune.refine(bf, best, reps=1)          # longer queue lags, on the whole batch
    log(f"[bench] batch setup {time.perf_counter() - t0:.1f}s, autotuned {best.key()} (whole batch: {whole})")
    stream = torch.cuda.current_stream(dev)

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize(dev)
            ts.append(a.elapsed_time(b))
        return ts

    bf.capture(best)
    tb = timed(bf.graph.replay, 3)
    err, flags = bf.status()
    fe_plain = torch.empty_like(bf.FEs)
    df_plain = torch.empty_like(bf.df)
    tp = timed(lambda: bf.plain_loop(fe_plain, df_plain), 2)
    fallback = np.flatnonzero(flags)
    if fallback.size:
        bf.recompute_plain(fallback)
        torch.cuda.synchronize(dev)
    same = bool(torch.equal(bf.df, df_plain) and torch.equal(bf.FEs, fe_plain))
    # the plain per-image loop on the same z windows (equal pair evaluations): separates what the
    # persistent launch, fusion and graph replay buy from what the windows buy
    fe_w = torch.empty_like(bf.FEs)
    df_w = torch.empty_like(bf.df)
    tw = timed(lambda: bf.plain_loop(fe_w, df_w, windowed=True), 2)
    same_w = bool(torch.equal(bf.df, df_w) and torch.equal(bf.FEs, fe_w))
    win_fallback = int(bf.last_plain_fallback.size)
    # both plain loops with a molecule's relaxation beside the next molecule's projection (two streams)
    fe_2 = torch.empty_like(bf.FEs)
    df_2 = torch.empty_like(bf.df)
    tp2 = timed(lambda: bf.plain_loop(fe_2, df_2, streams=2), 2)
    same_2 = bool(torch.equal(bf.df, df_2) and torch.equal(bf.FEs, fe_2))
    tw2 = timed(lambda: bf.plain_loop(fe_2, df_2, windowed=True, streams=2), 2)
    same_2 = same_2 and bool(torch.equal(bf.df, df_2) and torch.equal(bf.FEs, fe_2))
    states = bf.states
    t_b, t_p, t_w, t_p2, t_w2 = min(tb), min(tp), min(tw), min(tp2), min(tw2)
    res = {
        "workload": f"{n_mol} molecules, U(10..150) atoms, cells U(16,32) A x 20 A, 0.1 A grids, 128x128 px x 40 z",
        "batched_ms": t_b, "plain_loop_ms": t_p, "plain_windowed_loop_ms": t_w,
        "plain_loop_two_streams_ms": t_p2, "plain_windowed_loop_two_streams_ms": t_w2,
        "speedup_vs_plain_two_streams": t_p2 / t_b, "speedup_vs_windowed_plain_two_streams": t_w2 / t_b,
        "bitexact_two_stream_plain": same_2,
        "batched_pixel_z_steps_per_s": states / (t_b * 1e-3),
        "plain_pixel_z_steps_per_s": states / (t_p * 1e-3),
        "plain_windowed_pixel_z_steps_per_s": states / (t_w * 1e-3),
        "speedup": t_p / t_b, "speedup_vs_windowed_plain": t_w / t_b,
        "bitexact_windowed_plain": same_w, "windowed_plain_fallback_molecules": win_fallback,
        "launch": best.key(), "autotune_ms": timings, "autotune_whole_batch_ms": whole,
        "autotune_fused_vs_two_phase_ms": [min((v for k, v in timings.items() if "_fused" in k), default=None),
                                           min((v for k, v in timings.items() if k.endswith("_two_phase")),
                                               default=None)],
        "projected_pair_evals": bf.pair_evals, "z_margin_A": list(bf.z_margin),
        "full_grid_pair_evals": int(sum(w.pair_evals for w in mols)),
        "window_fallback_molecules": int(fallback.size), "wait_timeout": int(err),
        "bitexact_vs_plain": same,
    }
    log(f"[bench] batched: {res}")
    del bf, fe_plain, df_plain, fe_w, df_w, fe_2, df_2
    torch.cuda.empty_cache()
    return res


def substrate_leg(dev, n_poses=64, fd=True):
    """Placements of a 60-atom molecule 3 A above a 4-layer rock-salt slab (1024 ions, periodic
    x,y, 902x902x300 grid at 0.05 A, 192x192 pixels x 60 z-steps): the slab's raw grid once, then
    per placement the forward (grid over the slab base, relaxation, df) and the adjoint of an L2
    loss w.r.t. the placement (shift, rotation) and the molecule's charges, recomputing the grid
    and the probe positions in the backward (one shared grid buffer).  Every grid lives on the
    scan window (the sub-box of the 902x902x300 grid the probes can reach); a placement whose
    probes leave it is recomputed on the full grid (counted)."""
    import torch
    from spammm_amd import poses as PS, workloads as WL
    sc = WL.substrate_scene(n_poses=n_poses)
    spec = D_axis(sc)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)  # noqa: E731
    from spammm_amd import device as D, params as PP
    kl, kr = PP.stiffness_Nm_to_eVA2(sc["k_lat_Nm"]), PP.stiffness_Nm_to_eVA2(sc["k_rad_Nm"])
    tip = D.Tip(stiffness=np.array([-kl, -kl, -kl, -kr], np.float32),
                dpos0=np.array([0, 0, -sc["bond"], sc["bond"]], np.float32),
                tipC=np.array([0., 0., 1., sc["dtip"]], np.float32))
    pts = D.scan_points(sc["scan_p0"], sc["scan_da"], sc["scan_db"], *sc["nxy"])
    scene = PS.Scene(spec=spec, sub_atoms=t(sc["sub_atoms"]), sub_cpar=t(sc["sub_cLJ"]),
                     mol_xyz0=sc["mol_atoms"][:, :3].astype(np.float64), mol_cpar=t(sc["mol_cLJ"]), pts=t(pts),
                     nxy=sc["nxy"], nz=sc["nz"], tip=tip, Qs=sc["Qs"], QZs=sc["QZs"], npbc=sc["npbc"],
                     lvA=sc["lvA"], lvB=sc["lvB"], max_iter=1000, f2conv=1e-12, strict=False,
                     gate_eig_frac=0.5, gate_max_jump=0.3)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    scene.prepare()
    e1.record(stream)
    torch.cuda.synchronize(dev)
    t_base = e0.elapsed_time(e1)
    log(f"[bench] substrate base grid {t_base:.1f} ms")
    q0 = torch.tensor(sc["mol_atoms"][:, 3], dtype=torch.float64, device=dev)
    dpose = torch.tensor([0.1, -0.1, 0.0, 0.02], dtype=torch.float64, device=dev)
    t_f, t_b, gnorm, gated, cols_out = [], [], [], 0, 0
    checks, dfs, pgrads = [], [], []
    for k in range(n_poses):
        pose_k = torch.tensor(sc["poses"][k], dtype=torch.float64, device=dev)
        with torch.no_grad():
            target = PS.pose_df(pose_k + dpose, q0, scene).double()
        tcols = scene.last_columns.clone()
        pose = pose_k.clone().requires_grad_(True)
        q = q0.clone().requires_grad_(True)
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record(stream)
        dfi = PS.pose_df(pose, q, scene)
        cols = scene.last_columns & tcols
        loss = 0.5 * (((dfi.double() - target) * cols[..., None]) ** 2).sum()
        b.record(stream)
        loss.backward()
        c.record(stream)
        torch.cuda.synchronize(dev)
        t_f.append(a.elapsed_time(b))
        t_b.append(b.elapsed_time(c))
        gated += scene.last_status.get("gated", 0)
        cols_out += int(cols.numel() - cols.sum())
        gnorm.append(float(pose.grad.norm()))
        if k == 0:
            cols0 = cols.clone()
        dfs.append(dfi.detach())
        pgrads.append(pose.grad.detach().clone())
        if fd and k < 2:    # directional finite difference along -grad(pose) on two placements
            gp = pose.grad.detach()
            d = -gp / gp.norm()
            row = {"pose": k, "adjoint": -float(gp.norm())}
            for eps in (1e-3, 1e-4):
                with torch.no_grad():
                    lp = float(0.5 * (((PS.pose_df(pose_k + eps * d, q0, scene).double() - target)
                                       * cols[..., None]) ** 2).sum())
                    lm = float(0.5 * (((PS.pose_df(pose_k - eps * d, q0, scene).double() - target)
                                       * cols[..., None]) ** 2).sum())
                row[f"fd_eps_{eps:g}"] = (lp - lm) / (2 * eps)
            row["rel_err_best"] = min(abs(row["adjoint"] - row[f"fd_eps_{e:g}"]) / abs(row[f"fd_eps_{e:g}"])
                                      for e in (1e-3, 1e-4))
            checks.append(row)
        if (k + 1) % 16 == 0:
            log(f"[bench] substrate placements {k + 1}/{n_poses}")
    batched = _substrate_batched(scene, sc, dev, n_poses, q0, dpose, dfs, pgrads)
    del dfs, pgrads
    states = scene.n_scan * scene.nz
    with torch.no_grad():       # work of the grid->atom adjoint of placement 0 (its active workgroups)
        x0 = PS.place(torch.as_tensor(scene.mol_xyz0, device=dev), torch.as_tensor(scene.center, device=dev),
                      torch.tensor(sc["poses"][0], dtype=torch.float64, device=dev))
        _, PPs0 = PS._forward(scene, torch.cat([x0.float(), q0.float()[:, None]], 1).contiguous())
    nimg = (2 * scene.npbc[0] + 1) * (2 * scene.npbc[1] + 1)
    adj_pairs, adj_octets = grid_adjoint_active_pairs(scene.spec_run, PPs0, sc["mol_atoms"].shape[0] * nimg, cols0)
    res = {
        "grid_adjoint_work_placement0": {"octets": adj_octets, "pair_evals": adj_pairs,
                                         "flop": adj_pairs * ADJ_FLOP_PER_PAIR},
        "workload": f"60-atom molecule x {n_poses} placements over a 1024-ion rock-salt slab, "
                    f"{sc['n'][0]}x{sc['n'][1]}x{sc['n'][2]} grid (0.05 A), {sc['nxy'][0]}x{sc['nxy'][1]} px x {sc['nz']} z",
        "base_grid_ms": t_base, "forward_ms_mean": float(np.mean(t_f)), "adjoint_ms_mean": float(np.mean(t_b)),
        "forward_ms_median": float(np.median(t_f)), "adjoint_ms_median": float(np.median(t_b)),
        "first_placement_forward_ms": float(t_f[0]),
        "adjoint_over_forward": float(np.mean(t_b) / np.mean(t_f)),
        "pixel_z_steps_per_s_forward": states / (np.mean(t_f) * 1e-3),
        "placements_per_s_forward_plus_adjoint": 1e3 / (np.mean(t_f) + np.mean(t_b)),
        "gated_states_total": int(gated), "pose_grad_norm_median": float(np.median(gnorm)),
        "gated_columns_total": [cols_out, n_poses * scene.n_scan],
        "directional_fd": checks, "directional_fd_rel_err": [r["rel_err_best"] for r in checks],
        "scan_window": list(scene.spec_run.n), "window_fallbacks": int(scene.window_fallbacks),
        "grid_bytes": int(scene.base.numel() * 4), "peak_mem_GB": torch.cuda.max_memory_allocated(dev) / 1e9,
        "batched": batched,
        "batched_placements_per_s": batched["placements_per_s_forward_plus_adjoint"],
    }
    log(f"[bench] substrate: {res}")
    del scene
    torch.cuda.empty_cache()
    return res


def _substrate_batched(scene, sc, dev, n_poses, q0, dpose, dfs, pgrads):
    """The substrate placements again through poses.pose_df_batch (scene.batch_depth placements in
    flight, each on its own stream; the grids kept for the backward): the same L2 loss per placement,
    summed; its df images against the one-at-a-time images (bit for bit) and its pose gradients
    against theirs."""
    import torch
    from spammm_amd import poses as PS
    P = torch.tensor(np.asarray(sc["poses"][:n_poses]), dtype=torch.float64, device=dev)
    with torch.no_grad():
        target = PS.pose_df_batch(P + dpose, q0, scene).double()
    tcols = scene.last_columns.clone()
    stream = torch.cuda.current_stream(dev)
    for rep in range(2):        # the first pass allocates the kept grids and the batch buffers
        poses = P.clone().requires_grad_(True)
        q = q0.clone().requires_grad_(True)
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        torch.cuda.synchronize(dev)
        a.record(stream)
        df = PS.pose_df_batch(poses, q, scene)
        cols = scene.last_columns & tcols
        loss = 0.5 * (((df.double() - target) * cols[..., None]) ** 2).sum()
        b.record(stream)
        loss.backward()
        c.record(stream)
        torch.cuda.synchronize(dev)
    t_f, t_b = a.elapsed_time(b), b.elapsed_time(c)
    same = sum(bool(torch.equal(df[k].detach(), dfs[k])) for k in range(n_poses))
    rel = max(float((poses.grad[k] - pgrads[k]).norm() / pgrads[k].norm()) for k in range(n_poses))
    kept = min(n_poses, int(scene.batch_grid_budget // (scene.base.numel() * 4)))
    res = {"placements": n_poses, "in_flight": scene.batch_depth, "grids_kept": kept,
           "forward_ms": t_f, "adjoint_ms": t_b, "forward_ms_per_placement": t_f / n_poses,
           "adjoint_ms_per_placement": t_b / n_poses,
           "placements_per_s_forward_plus_adjoint": 1e3 * n_poses / (t_f + t_b),
           "df_bitequal_one_at_a_time": [same, n_poses], "pose_grad_max_rel_diff_one_at_a_time": rel,
           "gated_states": scene.last_status.get("gated", 0)}
    log(f"[bench] substrate batched: {res}")
    del target, df, loss
    torch.cuda.empty_cache()
    return res


def ewald_leg(dev, n_harm=40, reps=3):
    """Periodic electrostatics of the substrate configuration's 1024-ion rock-salt slab (8x8 cells,
    4 layers, 45.12 A cell) on its 902x902x300 grid at 0.05 A: the 2D Ewald probe grid (prepare =
    layer structure factors + G compaction; tip grid = plane tables + grid kernel) against the
    direct Coulomb sum over 3x3 cells that the substrate base otherwise uses (the grid-projection
    kernel with zero LJ parameters), and their difference in the probe region."""
    import torch
    from spammm_amd import _lib, device as D, workloads as WL
    from spammm_amd.ewald import SurfaceEwald
    sc = WL.substrate_scene(n_poses=1)
    spec = D_axis(sc)
    nx, ny, nz = spec.n
    stream = torch.cuda.current_stream(dev)

    def timed(fn, n):
        ts = []
        for _ in range(n):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize(dev)
            ts.append(e0.elapsed_time(e1))
        return float(np.min(ts)), float(np.mean(ts))

    # drop_tol 1e-6: the slab's FP32 coordinates break its 5.64 A periodicity at the 1e-7 level, which
    # leaves ~900 G of the 45 A supercell with structure factors of 2.7e-7 of the largest; those are
    # dropped (36 G kept, evaluated as 18 +-G pairs).  drop_tol 1e-12 keeps them; both are timed and
    # compared.
    ew_all = SurfaceEwald(dev).prepare_system(sc["sub_atoms"], sc["lvA"][:2], sc["lvB"][:2], n_harm, 1e-12)
    ew = SurfaceEwald(dev)
    t_prep = timed(lambda: ew.prepare_system(sc["sub_atoms"], sc["lvA"][:2], sc["lvB"][:2], n_harm, 1e-6), 2)[0]
    kept = ew.n_kept()
    g = torch.empty((nz, ny, nx, 4), dtype=torch.float32, device=dev)
    ew_all.tip_grid(spec, sc["Qs"], sc["QZs"], out=g)
    t_all = timed(lambda: ew_all.tip_grid(spec, sc["Qs"], sc["QZs"], out=g), 1)[0]
    g_all = g.clone()
    ew.tip_grid(spec, sc["Qs"], sc["QZs"], out=g)
    t_grid, t_grid_mean = timed(lambda: ew.tip_grid(spec, sc["Qs"], sc["QZs"], out=g), reps)
    sub = torch.as_tensor(sc["sub_atoms"], device=dev)
    zc = torch.zeros((sub.shape[0], 2), dtype=torch.float32, device=dev)
    gd = torch.empty_like(g)
    run_direct = lambda: D.grid_project(sub, zc, _lib.PPM_MODEL_LJ, spec, sc["Qs"], sc["QZs"], sc["npbc"],  # noqa: E731
                                        sc["lvA"], sc["lvB"], cap=False, out=gd)
    run_direct()
    t_direct = timed(run_direct, 2)[0]
    dz = float(spec.dC[2])
    res = {"workload": f"1024-ion rock-salt slab (8x8 cells, 4 layers), {nx}x{ny}x{nz} grid (0.05 A), "
                       f"n_harm {n_harm}",
           "n_g": _lib.load().ppm_ewald_num_g(n_harm), "kept_g_pairs": kept, "g_vectors_summed": 2 * kept,
           "prepare_ms": t_prep,
           "tip_grid_ms": t_grid, "tip_grid_ms_mean": t_grid_mean, "direct_3x3_coulomb_ms": t_direct,
           "kept_g_pairs_drop_1e-12": ew_all.n_kept(), "tip_grid_ms_drop_1e-12": t_all}
    nvox = nx * ny * nz
    res["tip_grid_GBps_written"] = nvox * 16 / (t_grid * 1e-3) / 1e9
    # executed work: 15 FLOP per voxel and kept pair (each pair's two equal terms evaluated once)
    res["tip_grid_fp32_TFLOPs"] = nvox * kept * 15 / (t_grid * 1e-3) / 1e12
    for h in (3.0, 5.0):
        iz0 = int(np.ceil((h - float(spec.p0[2])) / dz))
        dF = (g[iz0:, :, :, :3] - gd[iz0:, :, :, :3]).abs().amax().item()
        Fm = g[iz0:, :, :, :3].abs().amax().item()
        res[f"direct_vs_ewald_max_dF_above_{h:g}A"] = dF
        res[f"drop_1e-6_vs_1e-12_max_dF_above_{h:g}A"] = (g[iz0:, :, :, :3] - g_all[iz0:, :, :, :3]).abs().amax().item()
        res[f"ewald_max_F_above_{h:g}A"] = Fm
    # series truncation: the potential's change from n_harm to n_harm + 16 at probe heights 2..5.5 A
    # above the top layer, and at 1 A (near the surface the series needs more harmonics)
    hi = SurfaceEwald(dev).prepare_system(sc["sub_atoms"], sc["lvA"][:2], sc["lvB"][:2], n_harm + 16, 1e-6)
    for zlo in (1.0, 2.0):
        P = np.array([[22.56 + 0.3 * i, 22.56 - 0.2 * i, zlo + 0.5 * i] for i in range(8)], np.float32)
        res[f"phi_change_n_harm_plus_16_z_{zlo:g}_to_{zlo + 3.5:g}A"] = float(np.abs(
            hi.eval_full(P[:, 0], P[:, 1], P[:, 2]) - ew.eval_full(P[:, 0], P[:, 1], P[:, 2])).max())
    log(f"[bench] ewald: {res}")
    del g, gd, g_all
    torch.cuda.empty_cache()
    return res


def fdbm_leg(dev, reps=5):
    """FDBM product path on a PTCDA-sized grid (200x160x128 at 0.1 A): synthetic Gaussian densities of
    a 38-atom planar molecule (sample density, neutral charge density) and CO-like tip densities
    (total, dipolar charge) on a 24^3 box.  Times stage 3 with the inputs resident on the device
    (Pauli correlation + scaling, fused Poisson/electrostatic convolution, dispersion, composed
    force grid), the host-API call (uploads, F_total download included), and the forward and
    backward relaxed scans (scan_fdbm) of 100x80 pixels x 21 heights with shared_postprocess."""
    import torch
    from spammm_amd import _lib, afm, fdbm, postprocess as PP, workloads as WL
    rng = np.random.default_rng(8)
    en, pos = WL.flake_geometry(38, rng)
    Z = np.array([{"H": 1, "C": 6, "N": 7, "O": 8}[e] for e in en])
    q = rng.uniform(-0.2, 0.2, len(Z))
    q -= q.mean()
    shape, st = (200, 160, 128), 0.1
    origin = np.array([pos[:, 0].mean() - 10.0, pos[:, 1].mean() - 8.0, -6.4])

    def gauss(centers, sigma, weights, shp, org):
        ax = [torch.arange(shp[a], dtype=torch.float64, device=dev) * st + float(org[a]) for a in range(3)]
        X, Y, Zg = torch.meshgrid(*ax, indexing="ij")
        out = torch.zeros(shp, dtype=torch.float64, device=dev)
        for c, w in zip(centers, weights):
            r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Zg - c[2]) ** 2
            out += float(w) * torch.exp(-r2 / (2 * sigma * sigma)) / (2 * np.pi * sigma * sigma) ** 1.5
        return out.float().contiguous()

    rho = gauss(pos, 0.55, Z, shape, origin)
    rdiff = gauss(pos, 0.45, q, shape, origin) - gauss(pos, 0.65, q, shape, origin)
    tshape = (24, 24, 24)
    t0 = -0.5 * st * (np.array(tshape) - 1)
    tip_tot = gauss([[0, 0, 0]], 0.6, [6.0], tshape, t0).cpu().numpy()
    tip_del = (gauss([[0, 0, 0.2]], 0.5, [0.4], tshape, t0) - gauss([[0, 0, -0.2]], 0.5, [0.4], tshape, t0)).cpu().numpy()
    A, beta = 18.0, 0.9
    stream = torch.cuda.current_stream(dev)
    # device-resident stage 3
    tt = fdbm._place_tip(tip_tot, shape, False)
    td = fdbm._place_tip(tip_del, shape, False)
    atoms_w = torch.as_tensor(fdbm._c6_atoms(pos, Z, None, 30.0, 1.0), device=dev)
    bufs = {}

    def stage3_device():
        bufs["p"] = fdbm.pauli_scale_(fdbm.fft_combine(_lib.PPM_FFT_CORRELATE, rho, tt, st, st ** 3,
                                                       out=bufs.get("p")), A, beta)
        bufs["e"] = fdbm.fft_combine(_lib.PPM_FFT_POISSON_CONVOLVE, rdiff, fdbm.flip3(td), st,
                                     fdbm.POISSON_SCALE * st ** 3, out=bufs.get("e"))
        bufs["v"] = fdbm.dispersion(atoms_w, origin, st, shape, out=bufs.get("v"))
        bufs["g"] = fdbm.compose_gradient((bufs["p"], bufs["e"], bufs["v"]), st, grid=bufs.get("g"))

    stage3_device()
    torch.cuda.synchronize(dev)
    ts = []
    for _ in range(reps):
        a_, b_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a_.record(stream)
        stage3_device()
        b_.record(stream)
        torch.cuda.synchronize(dev)
        ts.append(a_.elapsed_time(b_))
    # host-API call (the reference's signature: host arrays in, F_total out)
    am = afm.AFMulator(use_morse=False)
    rho_h, rdiff_h = rho.cpu().numpy(), rdiff.cpu().numpy()
    t0w = time.perf_counter()
    fdbm.stage3_fdbm_fields_fast(am, rho_h, rdiff_h, tip_tot, tip_del, origin, st, shape, pos, Z, A, beta,
                                 download_fields=False)
    torch.cuda.synchronize(dev)
    t_api = time.perf_counter() - t0w
    h_df, h_Fz, h_scan = PP.afm_df_height_stacks(3.7, 4.7, 0.1, amp=1.0)
    xs = (pos[:, 0].mean() - 5.0 + 0.1 * np.arange(100)).astype(np.float32)
    ys = (pos[:, 1].mean() - 4.0 + 0.1 * np.arange(80)).astype(np.float32)
    kw = dict(mol_z=float(pos[:, 2].max()), K_LAT=afm.stiffness_Nm_to_eVA2(0.5), K_RAD=20.0, bond_length=3.0)
    t0s = time.perf_counter()
    FEf, disp = am.scan_fdbm(xs, ys, h_scan, **kw)
    FEb, _ = am.scan_fdbm(xs, ys, h_scan, reverse=True, **kw)
    t_scan = time.perf_counter() - t0s
    spec = PP.ScanSpec(scan_xs=xs, scan_ys=ys, h_df=h_df, h_Fz=h_Fz, h_scan=h_scan, amplitude=1.0,
                       osc_dir=(0., 0., 1.), K_LAT=kw["K_LAT"], K_RAD=20.0, bond_length=3.0)
    t0p = time.perf_counter()
    res = PP.shared_postprocess(FEf, spec, FEs_bwd=FEb, tip_disp=disp, backend_name="fdbm", fft_path="GPU")
    t_post = time.perf_counter() - t0p
    nvox = int(np.prod(shape))
    out = {
        "workload": f"FDBM: 38-atom planar molecule, synthetic Gaussian densities, {shape[0]}x{shape[1]}x{shape[2]} "
                    f"grid at 0.1 A, tip 24^3, scan {len(xs)}x{len(ys)} px x {len(h_scan)} heights fwd + bwd",
        "stage3_device_ms": float(np.median(ts)), "stage3_host_api_s": t_api,
        "grid_points_per_s_stage3": nvox / (np.median(ts) * 1e-3),
        "scan_fwd_bwd_s": t_scan, "postprocess_s": t_post,
        "df_range": [float(res.df.min()), float(res.df.max())], "E_diss_max": float(res.E_diss.max()),
        "fft_count_per_stage3": 5,
    }
    log(f"[bench] fdbm: {out}")
    return out


def D_axis(sc):
    from spammm_amd import device as D
    return D.axis_aligned_spec(sc["p0"], sc["L"], sc["n"])


def _pmc_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 --pmc summary
    (profiles/pmc_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same
    counters' reading for a 1 GiB copy), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        return rec[workload][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _direct_group(n_scan, na):
    """Lanes per scan point the direct relaxation launch uses (same rule as the library)."""
    g = 1
    while g < 64 and n_scan * g * 2 <= 4 * 1024 * 64 and g < na:
        g *= 2
    return g


def direct_leg(dev, reps=5):
    """Gridless direct Morse path: 128-atom flake (the path's atom limit), 64x64 pixels x 60 z-steps
    (the largest case of the reference's performance matrix for this path), CO-like tip; forward
    relaxation and the FP64 implicit VJP timed with HIP events."""
    import torch
    from spammm_amd import device as D, params as P, workloads as WL
    rng = np.random.default_rng(5)
    na, nxy, nz, dtip = 128, (64, 64), 60, -0.1
    en, pos = WL.flake_geometry(na, rng)
    atoms = np.zeros((na, 4), np.float32)
    atoms[:, :3] = pos
    atoms[:, 3] = rng.uniform(-0.3, 0.3, na)
    atoms[:, 3] -= atoms[:, 3].mean()
    cMs = P.mix_morse(en, P.ELEMENTS_REFERENCE, 1.452, 0.0006808, -1.8).astype(np.float32)
    lo, hi = pos[:, :2].min(0) - 1.0, pos[:, :2].max(0) + 1.0
    step = (hi - lo) / np.array(nxy)
    p0 = np.array([lo[0] + step[0] / 2, lo[1] + step[1] / 2, 13.0], np.float32)
    pts = D.scan_points(p0, np.array([step[0], 0, 0], np.float32), np.array([0, step[1], 0], np.float32), *nxy)
    kl, kr = P.stiffness_Nm_to_eVA2(0.25), P.stiffness_Nm_to_eVA2(20.0)
    tip = D.Tip(stiffness=np.array([-kl, -kl, -kl, -kr], np.float32), tipC=np.array([0., 0., 1., dtip], np.float32))
    Qs, QZs = np.array([-0.05, 0, 0, 0], np.float32), np.zeros(4, np.float32)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)  # noqa: E731
    A, M, Pt = t(atoms), t(cMs), t(pts)
    stream = torch.cuda.current_stream(dev)
    ns = nxy[0] * nxy[1]
    FEs = torch.empty((ns, nz, 4), dtype=torch.float32, device=dev)
    PPs = torch.empty_like(FEs)
    U = torch.zeros_like(FEs)
    U[..., 2] = torch.randn((ns, nz), device=dev, generator=torch.Generator(device=dev).manual_seed(1))

    def timed(fn):
        ts = []
        for r in range(reps + 1):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize(dev)
            if r:
                ts.append(a.elapsed_time(b))
        return float(np.mean(ts))

    t_f = timed(lambda: D.relax_direct_morse(A, M, Pt, nz, tip, Qs, QZs, 100, 1e-8, scan_ny=nxy[1], FEs=FEs,
                                             PPs=PPs))
    res = {}
    t_b = timed(lambda: res.setdefault("v", D.direct_morse_vjp(A, M, Pt, nz, tip, Qs, QZs, PPs, U)))
    pp = PPs.cpu().numpy()
    evals = float(np.abs(pp[..., 3]).sum() + (pp[..., 3] < 0).sum())     # FE evaluations (+final for unconverged)
    flop = evals * na * 54
    tflops = flop / (t_f * 1e-3) / 1e12
    diag = res["v"][2].cpu().numpy()
    out = {
        "workload": f"{na}-atom flake, Morse + 1 charge site, {nxy[0]}x{nxy[1]} pixels x {nz} z-steps, "
                    "FIRE |F|<1e-4 <= 100 it",
        "forward_ms": t_f, "vjp_ms": t_b, "vjp_over_forward": t_b / t_f,
        "pixel_z_steps_per_s": ns * nz / (t_f * 1e-3),
        "mean_fire_evals": evals / (ns * nz), "converged_fraction": float((pp[..., 3] > 0).mean()),
        "forward_TFLOPs_algorithmic": tflops, "forward_frac_of_fp32_peak": tflops / PEAK_FP32_TFLOPS,
        "vjp_states_used_fraction": float((diag[..., 3] == 0).mean()),
        "lanes_per_point": _direct_group(ns, na),
    }
    log(f"[bench] direct: {out}")
    return out


def _diff_case(dev, z0, adam_steps, lr, fd_steps, order4=False, fd=True):
    """One height range of the differentiable flake case (see diff_leg)."""
    import torch
    from spammm_amd import diff, gradcheck as GC, params as P, workloads as WL
    w = WL.flake_differentiable(z0)
    prob = diff.Problem.from_workload(w, dev, max_iter=1000, f2conv=1e-12, check_before=True, polish=3,
                                      strict=False, gate_eig_frac=0.5, gate_max_jump=0.3, window_margin=(2.0, 1.0))
    R0, E0 = P.mix_R0E0(w.enames, P.ELEMENTS_BENCH, 1.66, 0.0091)
    R0, E0 = torch.tensor(R0, device=dev), torch.tensor(E0, device=dev)
    xyz0 = torch.tensor(w.atoms[:, :3], dtype=torch.float64, device=dev)
    q0 = torch.tensor(w.atoms[:, 3], dtype=torch.float64, device=dev)
    xyz_t = torch.tensor(WL.perturbed_positions(w), device=dev)
    with torch.no_grad():
        target = diff.df_image(xyz_t, q0, R0, E0, prob).double()
    tcols, tgate = prob.last_columns.clone(), dict(prob.last_gate)

    stream = torch.cuda.current_stream(dev)
    t_f, t_b = [], []
    for rep in range(4):
        x = xyz0.clone().requires_grad_(True)
        q = q0.clone().requires_grad_(True)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(stream)
        dfi = diff.df_image(x, q, R0, E0, prob)
        cols = prob.last_columns & tcols
        loss = diff.masked_l2(dfi, target, cols)
        e1.record(stream)
        loss.backward()
        e2.record(stream)
        torch.cuda.synchronize()
        if rep > 0:
            t_f.append(e0.elapsed_time(e1))
            t_b.append(e1.elapsed_time(e2))
    tf, tb = float(np.mean(t_f)), float(np.mean(t_b))
    gx, gq = x.grad.detach(), q.grad.detach()
    gating = {"forward": dict(prob.last_gate), "target": tgate, "columns_kept_both": int(cols.sum()),
              "columns": int(cols.numel())}
    with torch.no_grad():
        atoms = torch.cat([xyz0.float(), q0.float()[:, None]], 1).contiguous()
        _, _, PPs, _, run_spec = diff._forward_parts(atoms, diff.lj_cpar(R0, E0), prob)
    nimg = (2 * prob.npbc[0] + 1) * (2 * prob.npbc[1] + 1)
    adj_pairs, adj_octets = grid_adjoint_active_pairs(run_spec, PPs, xyz0.shape[0] * nimg, cols)
    status = dict(prob.last_status)
    # central differences of the FP64 device forward on 32 random parameters (seed 3), same gated loss
    na = xyz0.shape[0]
    sel = np.random.default_rng(3).choice(4 * na, 32, replace=False)
    t0 = time.perf_counter()
    fdres = None
    if fd:
        fdres = GC.fd_check(GC.Fp64Forward(prob, R0, E0), target.reshape(prob.n_scan, prob.nz), xyz0, q0, gx, gq,
                            sel, steps=fd_steps, columns=cols, order4_step=fd_steps[0] if order4 else None)
        fdres["wall_s"] = time.perf_counter() - t0
    # Adam on the positions, the loss gated on the columns stable in the forward and in the target
    x = xyz0.clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    hist = []
    t0 = time.perf_counter()
    for it in range(adam_steps):
        opt.zero_grad()
        loss = diff.masked_l2(diff.df_image(x, q0, R0, E0, prob), target, prob.last_columns & tcols)
        loss.backward()
        opt.step()
        hist.append(float(loss.detach()))
    torch.cuda.synchronize()
    t_adam = time.perf_counter() - t0
    d0, d1 = xyz0 - xyz_t, x.detach() - xyz_t
    rxy = [float(d[:, :2].pow(2).sum(1).mean().sqrt()) for d in (d0, d1)]
    rz = [float(d[:, 2].pow(2).mean().sqrt()) for d in (d0, d1)]
    return {
        "tip_z": [z0, z0 + (w.nz - 1) * float(w.dtip)], "probe_z_rest": [z0 - 4.0, z0 - 4.0 + (w.nz - 1) * float(w.dtip)],
        "forward_ms": tf, "adjoint_ms": tb, "adjoint_over_forward": tb / tf,
        "pixel_z_steps_per_s_forward": w.n_scan * w.nz / (tf * 1e-3),
        "grid_run": list(run_spec.n), "grid_full": list(prob.spec.n), "window_fallbacks": prob.window_fallbacks,
        "adjoint_status": status, "gating": gating, "fd_fp64": fdres,
        "grid_adjoint_work": {"octets": adj_octets, "pair_evals": adj_pairs,
                              "flop": adj_pairs * ADJ_FLOP_PER_PAIR, "flop_per_pair": ADJ_FLOP_PER_PAIR},
        "adam": {"steps": adam_steps, "lr": lr, "wall_s": t_adam, "loss_first": hist[0] if hist else None,
                 "loss_last": hist[-1] if hist else None, "loss_every5": hist[::5],
                 "rmsd_xy_start_end_A": rxy, "rmsd_z_start_end_A": rz},
    }


# grid->atom adjoint per pair (pair_adjoint, LJ + tip charge at z offset 0), counted on its formulas
# like the forward's 52 (add/mul 1, FMA 2, rsqrt 1): d = v - a 3, |d|^2 + eps 6, 1/r and 1/r^2 2,
# g.d 5, g.d/r^2 1, r^-6, r^-8 3, c12 r^-6 1, fr 3, dfr/ds 4, charge term 6, dq 4, the common factor
# 4, the three position components 12 = 54; the (c6, c12) columns add 9 (63).  Both legs that report
# it differentiate positions and charges only.
ADJ_FLOP_PER_PAIR = 54


def grid_adjoint_active_pairs(spec, PPs, n_sites, columns=None, V=8):
    """Pair evaluations of k_grid_adjoint: it evaluates the octets (a grid column x V planes, V = 8)
    that hold a corner of the trilinear cell of some relaxed probe whose pixel column is in the loss
    (the only voxels with a nonzero cotangent), x V voxels x atom sites.  Returns (pairs, octets)."""
    nx, ny, nz = (int(v) for v in spec.n)
    step = np.array([spec.dA[0], spec.dB[1], spec.dC[2]], np.float64)
    p = PPs[..., :3].reshape(-1, 3).double().cpu().numpy()
    if columns is not None:
        keep = np.repeat(columns.reshape(-1).cpu().numpy().astype(bool), p.shape[0] // columns.numel())
        p = p[keep]
    i0 = np.floor((p - np.asarray(spec.p0[:3], np.float64)) / step).astype(np.int64)
    octets = set()
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                ia = (i0[:, 0] + cx) % nx
                ib = (i0[:, 1] + cy) % ny
                ic = (i0[:, 2] + cz) % nz
                octets.update(np.unique((ic // V) * (nx * ny) + ib * nx + ia).tolist())
    return len(octets) * V * n_sites, len(octets)


def diff_leg(dev, adam_steps=50, fd=True):
    """Differentiable df image of the flake molecule, 128x128 pixels x 40 z-steps, L2 loss to a target
    made from positions perturbed by N(0, 0.1 A), taken over the pixel columns whose probes stay on one
    equilibrium branch in the forward and in the target (diff.stable_columns): forward and adjoint
    timed with HIP events (adjoint/forward ratio), central differences of the FP64 device forward on 32
    random parameters (seed 3) at h = 1e-3 and 1e-4 A (1e-4, 1e-5 e), a fourth-order stencil at 1e-3,
    and 50 Adam steps (lr 1e-2) on the positions.  Run at the configured tip heights (10 -> 6.1 A:
    probe 6 -> 2.1 A) and 1 A higher (probe 7 -> 3.1 A)."""
    steps = ((1e-3, 1e-4), (1e-4, 1e-5))
    res = {"workload": "flake 200 atoms, 300x300x240 grid, 128x128 pixels x 40 z, Giessibl df, gated L2 loss",
           "configured_heights": _diff_case(dev, 10.0, adam_steps, 1e-2, steps, order4=True, fd=fd),
           "smooth_heights": _diff_case(dev, 11.0, adam_steps, 1e-2, steps, fd=fd)}
    c = res["configured_heights"]
    res["forward_ms"], res["adjoint_ms"], res["adjoint_over_forward"] = (c["forward_ms"], c["adjoint_ms"],
                                                                       c["adjoint_over_forward"])

    def fd_summary(f):
        if f is None:
            return None
        out = {p["h"]: p["within_tol"] for p in f["per_step"]}
        if "order4" in f:
            out[f["order4"]["h"] + " order4"] = f["order4"]["within_tol"]
        out["plateau"] = f["plateau_within_tol"]
        out["of"] = f["n"]
        return out

    def adam_summary(a):
        return {"loss": [a["loss_first"], a["loss_last"]], "rmsd_xy_A": a["rmsd_xy_start_end_A"],
                "rmsd_z_A": a["rmsd_z_start_end_A"]}

    res["fd_fp64_configured_heights"] = fd_summary(c["fd_fp64"])
    res["fd_fp64_smooth_heights"] = fd_summary(res["smooth_heights"]["fd_fp64"])
    res["gated_columns_configured_heights"] = [c["gating"]["columns"] - c["gating"]["columns_kept_both"],
                                               c["gating"]["columns"]]
    res["adam_configured_heights"] = adam_summary(c["adam"])
    res["adam_smooth_heights"] = adam_summary(res["smooth_heights"]["adam"])
    log(f"[bench] differentiable: {res}")
    return res


def accuracy_leg(w, sp, grid, pts_h, FEs, PPs, df, tip):
    """Error of every output against the oracle on a sample (FP64 and FP32-faithful)."""
    from oracle import oracle as O, giessibl_ref as GQ
    t0 = time.perf_counter()
    rng = np.random.default_rng(11)
    G = grid.cpu().numpy()
    idx = rng.choice(w.nvox, 1000, replace=False)
    Go = O.grid_voxels(0, w.atoms, w.cLJ, w.n, sp.p0, sp.dA, sp.dB, sp.dC, w.Qs, w.QZs, idx, w.npbc, w.lvA, w.lvB)
    Gg = G.reshape(-1, 4)[idx].astype(np.float64)
    mag = np.linalg.norm(Go, axis=1)
    gerr = np.abs(Gg - Go).max(1)
    grid_rel = gerr / np.maximum(mag, 1e-30)
    # the FP32 CPU oracle at the same voxels: the error any FP32 sum of these terms makes
    nx, ny = w.n[0], w.n[1]
    G32 = np.empty((idx.size, 4))
    for k, v in enumerate(idx):
        pv = (np.asarray(sp.p0, np.float64) + (v % nx) * np.asarray(sp.dA, np.float64)
              + ((v // nx) % ny) * np.asarray(sp.dB, np.float64) + (v // (nx * ny)) * np.asarray(sp.dC, np.float64))
        G32[k] = O.grid_project(0, w.atoms, w.cLJ, (1, 1, 1), pv.astype(np.float32), sp.dA, sp.dB, sp.dC, w.Qs,
                                w.QZs, w.npbc, w.lvA, w.lvB, cap=True, precision="f32").reshape(4)
    gerr32 = np.abs(G32 - Go).max(1)
    # the full-size parity bar: 1e-5 |(F,E)| + 4e-6 sum|terms| (rescaled on capped voxels) + 1e-6
    ic, rem = np.divmod(idx, nx * ny)
    ib, ia = np.divmod(rem, nx)
    Pv = (np.asarray(sp.p0, np.float64)[None, :] + np.outer(ia, sp.dA) + np.outer(ib, sp.dB) + np.outer(ic, sp.dC))
    S = O.abs_term_sums(w.atoms, w.cLJ, w.Qs[0], w.npbc, w.lvA, w.lvB, Pv)
    cap = np.minimum(1.0, 100.0 / np.maximum(np.abs(Go[:, 3]), 1e-300))
    tbar = 1e-5 * mag + 4e-6 * S * cap + 1e-6
    sel = rng.choice(w.n_scan, 64, replace=False)
    F = FEs.cpu().numpy()[sel]
    P = PPs.cpu().numpy()[sel]
    args = (G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts_h[sel], w.nz, tip.tipA, tip.tipB, tip.tipC, tip.stiffness,
            tip.dpos0, tip.relax, tip.surfFF, w.max_iter, w.f2conv, False)
    F32, P32 = O.relax_tilted(*args, precision="f32")
    F64, P64 = O.relax_tilted(*args, precision="f64")
    dpp = np.abs(P[..., :3] - P64[..., :3]).max(-1)
    dfz = np.abs(F[..., 2] - F64[..., 2])
    z = w.scan_p0[2] + w.dtip * np.arange(w.nz)
    dfo = GQ.giessibl_df_piecewise(F[..., 2].astype(np.float64), z, w.amp, w.f0, w.k_cant)
    dfg = df.cpu().numpy().reshape(w.n_scan, w.nz)[sel]
    ddf = np.abs(dfg - dfo)
    rms = lambda x: float(np.sqrt(np.mean(np.square(x))))  # noqa: E731
    res = {
        # per voxel: max channel error relative to |(F,E)|.  The small-case parity bar is
        # |d| <= 1e-5 |(F,E)| + 1e-6; high above the flake the net field (~1e-2) is a sum of
        # ~1800 Coulomb terms of ~1e-2 each, so any FP32 sum errs by a few 1e-6 there (the FP32
        # oracle's own error at the same voxels is reported beside the GPU's) and the full-size
        # bar adds 4e-6 sum|terms|: `*_over_term_bar_max` <= 1 is the full-size parity test's pass
        "grid_rel_err_vs_fp64_max": float(grid_rel.max()), "grid_rel_err_vs_fp64_rms": rms(grid_rel),
        "grid_abs_err_vs_fp64_max": float(gerr.max()),
        "grid_err_over_small_case_bar_max": float((gerr / (1e-5 * mag + 1e-6)).max()),
        "grid_fp32_oracle_err_over_small_case_bar_max": float((gerr32 / (1e-5 * mag + 1e-6)).max()),
        "grid_err_over_term_bar_max": float((gerr / tbar).max()),
        "grid_fp32_oracle_err_over_term_bar_max": float((gerr32 / tbar).max()),
        "grid_fp32_oracle_abs_err_vs_fp64_max": float(gerr32.max()),
        "relax_bitexact_vs_fp32_oracle": bool(np.array_equal(F, F32) and np.array_equal(P, P32)),
        "pp_err_vs_fp64_max_A": float(dpp.max()), "pp_err_vs_fp64_rms_A": rms(dpp),
        "pp_err_vs_fp64_median_A": float(np.median(dpp)),
        "fz_err_vs_fp64_max": float(dfz.max()), "fz_err_vs_fp64_rms": rms(dfz),
        "fz_err_vs_fp64_median": float(np.median(dfz)),
        "df_err_vs_fp64_max": float(ddf.max()), "df_err_vs_fp64_rms": rms(ddf),
        "df_max_rel_err_vs_fp64": float(ddf.max() / max(np.abs(dfo).max(), 1e-30)),
        "sample": "1000 grid voxels, 64 pixels x all z; relaxation errors at the reference's stopping "
                  "rule (|F| < 1e-4 eV/A, <= 100 iterations), FP64 oracle relaxed on the same grid",
    }
    log(f"[bench] accuracy leg {time.perf_counter() - t0:.1f}s: {res}")
    return res


def _cpu_model():
    """The host CPU's model name (from /proc/cpuinfo, as lscpu reports it) and its logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs on the host)"


def cpu_leg(w, sp, grid, pts_h, tip, states):
    """C++ oracle (FP32 mode, OpenMP) on a bounded sample, scaled to the whole step."""
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    G = grid.cpu().numpy()
    # grid: a slab of z-planes (every voxel costs the same: a direct sum over all sites)
    nslab = max(1, w.n[2] // 24)
    p0 = sp.p0.copy()
    t0 = time.perf_counter()
    O.grid_project(0, w.atoms, w.cLJ, (w.n[0], w.n[1], nslab), p0, sp.dA, sp.dB, sp.dC, w.Qs, w.QZs, w.npbc, w.lvA,
                   w.lvB, cap=True, precision="f32")
    t_grid = (time.perf_counter() - t0) * w.n[2] / nslab
    # relax: a random subset of pixels on the same grid
    rng = np.random.default_rng(17)
    nsel = min(w.n_scan, 4096)
    sel = rng.choice(w.n_scan, nsel, replace=False)
    t0 = time.perf_counter()
    O.relax_tilted(G, sp.n, sp.dinvA, sp.dinvB, sp.dinvC, pts_h[sel], w.nz, tip.tipA, tip.tipB, tip.tipC,
                   tip.stiffness, tip.dpos0, tip.relax, tip.surfFF, w.max_iter, w.f2conv, False, precision="f32")
    t_relax = (time.perf_counter() - t0) * w.n_scan / nsel
    t_step = t_grid + t_relax
    res = {
        "value": states / t_step,
        "unit": "pixel*z-steps/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "sample": (f"grid: {nslab} of {w.n[2]} z-planes; relax: {nsel} of {w.n_scan} pixels x {w.nz} z; "
                   f"scaled to the full step (grid {t_grid:.2f} s + relax {t_relax:.2f} s)"),
        "grid_points_per_s": w.nvox / t_grid,
        "relax_pixel_z_steps_per_s": states / t_relax,
    }
    log(f"[bench] cpu leg: {res}")
    return res


if __name__ == "__main__":
    main()
