"""The bench result line stays short and keeps the headline fields: the line is built from stub
legs shaped like the real ones (including the large per-parameter records that go to the detail
file only), on the CPU."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

HEAD_KEYS = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
             "vs_baseline", "dtype", "data", "config", "roofline", "stages", "cpu_baseline")


def _head():
    return {
        "metric": "relaxed pixel*z-steps/s on 1 MI355X (batched vs plain); adjoint/forward ratio",
        "value_is": "flake grid + relaxed scan + Giessibl df per image, consecutive images on two streams",
        "value": 240250000.123456, "unit": "pixel*z-steps/s", "n_gpus": 1, "steps": 20, "warmup": 3,
        "ms_per_step": 16.367123, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "flake: 200 atoms LJ+charge, 300x300x240 float4 grid (3x3 periodic images), "
                               "256x256 pixels x 60 z-steps, FIRE |F|<1e-4 <= 100 it, Giessibl df A=1.0 A",
                   "grid": [300, 300, 240], "atoms": 200, "pixels": [256, 256], "z_steps": 60},
        "roofline": {"kernel": "k_grid_project (grid projection)", "bound": "valu",
                     "compute_unit": "FP32 VALU, no matrix instructions; peak = published MI355X FP32",
                     "achieved": 151.66123, "peak": 157.3, "unit": "TFLOP/s", "frac": 0.9641234,
                     "traffic": 345794965.3, "work_unit": "38880000000 pair evaluations x 52 FLOP per launch",
                     "avg_launch_ms": 14.401, "avg_launch_ms_one_stream": 13.331, "frac_one_stream": 0.9641234},
        "stages": {"one_stream": {"ms_per_step": 15.546123, "value": 252940000.123},
                   "two_streams": {"grid_ms": 14.40123, "relax_ms": 3.512345, "outputs_equal_one_stream": True},
                   "grid_points_per_s": 1620300000.0, "grid_ms": 13.331, "relax_ms": 3.0272,
                   "relax_pixel_z_steps_per_s": 1298900000.0, "relax_mean_fire_iters": 25.96,
                   "relax_converged_fraction": 0.99876, "relax_algorithmic_GBps": 4524.0,
                   "relax_algorithmic_frac_of_hbm_peak": 0.56551, "relax_hbm_bytes_per_launch": 286130747.1,
                   "df": "fused into the relaxation launch (ppm_relax_strokes_tilted_df)"},
        "cpu_baseline": {"value": 222790.0, "unit": "pixel*z-steps/s", "cores": 16, "kind": "port",
                         "cpu_model": "AMD EPYC (16 logical CPUs on the host)",
                         "sample": "grid: 10 of 240 z-planes; relax: 4096 of 65536 pixels x 60 z; scaled to the "
                                   "full step (grid 17.25 s + relax 0.40 s)",
                         "grid_points_per_s": 1252000.0, "relax_pixel_z_steps_per_s": 9905300.0},
        "detail_file": "bench_detail.json",
    }


def _fd(n_steps=2):
    rows = [(f"xyz[{i},{i % 3}]", 0.123456 * i, [0.1234567 * i] * n_steps) for i in range(32)]
    return {"tol": 1e-3, "n": 32, "rows": rows,
            "per_step": [{"h": "0.001 A / 0.0001 e", "within_tol": 30, "median_rel_err": 9.2e-05,
                          "max_rel_err": 1.7e-03}] * n_steps}


def _legs():
    """Legs with every summary key the real legs carry, plus bulky detail-only entries."""
    adam = {"loss": [17960.0, 160.62], "rmsd_xy_A": [0.14212, 0.12034], "rmsd_z_A": [0.097573, 0.043297]}
    fdsum = {"0.001 A / 0.0001 e": 30, "0.0001 A / 1e-05 e": 32, "0.001 A / 0.0001 e order4": 30, "plateau": 32,
             "of": 32}
    return {
        "differentiable": {"forward_ms": 19.274, "adjoint_ms": 6.0912, "adjoint_over_forward": 0.3161,
                           "fd_fp64_configured_heights": fdsum, "fd_fp64_smooth_heights": fdsum,
                           "gated_columns_configured_heights": [4354, 16384], "adam_configured_heights": adam,
                           "adam_smooth_heights": adam, "configured_heights": {"fd_fp64": _fd()},
                           "smooth_heights": {"fd_fp64": _fd()}},
        "batched": {"batched_ms": 499.52, "plain_loop_ms": 2614.6, "plain_windowed_loop_ms": 1504.0,
                    "speedup": 5.2341, "speedup_vs_windowed_plain": 3.0108, "batched_pixel_z_steps_per_s": 6.7173e8,
                    "speedup_vs_plain_two_streams": 4.1234567, "speedup_vs_windowed_plain_two_streams": 2.5123456,
                    "launch": "vz8_tw4_w3_fused_lag8", "autotune_fused_vs_two_phase_ms": [29.041, 29.149],
                    "bitexact_vs_plain": True, "bitexact_windowed_plain": True, "window_fallback_molecules": 0,
                    "wait_timeout": 0, "autotune_ms": {f"vz{v}_tw{t}_w{w}_fused_lag{l}": 30.0 + v + t + w + l
                                                       for v in (4, 8) for t in (4, 8, 16) for w in (2, 3, 4)
                                                       for l in (0, 2, 8)}},
        "direct_morse": {"forward_ms": 3.6843, "vjp_ms": 0.56994, "vjp_over_forward": 0.15469,
                         "pixel_z_steps_per_s": 66704000.0},
        "substrate_poses": {"forward_ms_median": 15.285, "adjoint_ms_median": 13.981, "adjoint_over_forward": 0.92071,
                            "placements_per_s_forward_plus_adjoint": 33.98, "batched_placements_per_s": 52.123,
                            "window_fallbacks": 0,
                            "gated_states_total": 39059, "gated_columns_total": [140093, 2359296],
                            "directional_fd_rel_err": [4.9427e-05, 3.7668e-05],
                            "directional_fd": [{"pose": k, "adjoint": -68241.49, "fd_eps_0.001": -68244.86}
                                               for k in range(2)]},
        "fdbm": {"stage3_device_ms": 0.50279, "grid_points_per_s_stage3": 8146600000.0},
        "ewald": {"tip_grid_ms": 1.221, "direct_3x3_coulomb_ms": 740.73, "kept_g_pairs": 18},
        "accuracy": {"relax_bitexact_vs_fp32_oracle": True, "pp_err_vs_fp64_max_A": 0.0095036,
                     "pp_err_vs_fp64_rms_A": 0.00056819, "fz_err_vs_fp64_max": 0.00025996,
                     "fz_err_vs_fp64_rms": 5.8436e-06, "df_max_rel_err_vs_fp64": 1.5094e-07,
                     "grid_err_over_term_bar_max": 0.43531},
    }


def test_result_line_is_short_and_complete():
    out = _head()
    out.update(bench.summarize(_legs()))
    line = bench.compact_line(out)
    print(f"result line: {len(line)} bytes")
    assert len(line) <= bench.LINE_LIMIT
    parsed = json.loads(line)
    for k in HEAD_KEYS:
        assert k in parsed, k
    assert parsed["roofline"]["bound"] == "valu"
    assert 0 < parsed["roofline"]["frac"] <= 1.5
    for leg in ("differentiable", "batched", "substrate_poses", "accuracy"):
        assert leg in parsed, leg
    assert "rows" not in line and "autotune_ms" not in line     # detail-only records stay out
    assert "dropped_from_line" not in parsed


def test_overlong_line_drops_leg_summaries_not_the_headline():
    out = _head()
    legs = bench.summarize(_legs())
    legs["accuracy"]["pp_err_vs_fp64_max_A"] = [0.123456789] * 600        # a leg summary grown too long
    out.update(legs)
    line = bench.compact_line(out)
    assert len(line) <= bench.LINE_LIMIT
    parsed = json.loads(line)
    for k in HEAD_KEYS:
        assert k in parsed, k
    assert parsed["dropped_from_line"] == ["accuracy"]
    assert "batched" in parsed
