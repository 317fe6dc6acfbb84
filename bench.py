#!/usr/bin/env python3
"""Benchmark of the Debye-SAXS hot path on one MI355X.

Workload (hydrated forward, sasa-hydrated-20000x256): 20,000 atoms at 0.045 A^-3 in a
1:1.3:1.8 ellipsoid, SASA hydration (256 Fibonacci dots, probe 1.4 A,
contrast 0.03 e/A^3), 256 q over 0-0.6 A^-1.  A step is one full device-
resident forward pass: SASA dot counts -> hydration classes -> Debye pair
kernel -> FP64 reduction, over positions already in HBM.

Prints ONE JSON line.  value = forward pair.q terms/s; roofline = the pair
kernel's algorithmic FP32 FLOP rate (P (5 n_q + 9) + 2 N n_q per launch)
over its HIP-event duration on the engine stream, against the FP32 peak.
"""
import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "debye_pair_kernel_sasa-hydrated-20000x256_counters.txt")
# autotuner cache (one line per arch / path / n_q / size bucket / hydration key); a missing key is
# tuned on the spot and appended
TUNING_CACHE = os.path.join(ROOT, "resaxs_amd", "tuning_gfx950.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_engine(w, hydrate=True):
    from resaxs_amd import RSX_DOTS_FIBONACCI, Engine
    e = Engine(0)
    e.set_q(w.q)
    e.set_factors(w.table, w.water_index)
    if hydrate:
        e.set_sasa_params(w.radii, w.probe, RSX_DOTS_FIBONACCI, w.n_half)
        e.set_water_weight(w.w)
    e.initialize(len(w.bodies))
    return e


def read_traffic():
    """HBM bytes per pair-kernel launch from the committed rocprofv3 counter summary."""
    try:
        text = open(TRAFFIC_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def forward_bench(steps, warmup):
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000()
    n, nq = len(w.bodies), len(w.q)
    e = make_engine(w)
    log(f"bench: {w.name}: host calc_curve (layout + first result)")
    I_host = e.calc_curve(w.bodies)
    tuning = e.autotune(TUNING_CACHE)
    log(f"bench: tuning {tuning}")
    I_host = e.calc_curve(w.bodies)
    d_b = e.dev_alloc(n * 16)
    d_out = e.dev_alloc(nq * 8)
    e.h2d(d_b, w.bodies)
    for _ in range(warmup):
        e.calc_curve_dev_async(d_b, n, d_out)
    e.synchronize()
    log(f"bench: timing {steps} device-resident steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        e.calc_curve_dev_async(d_b, n, d_out)
    e.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / steps
    I_dev = e.d2h(np.zeros(nq, dtype=np.float64), d_out)
    # same kernels and layout; only the q = 0 value is summed in another order
    if not np.all(np.abs(I_dev - I_host) <= 1e-12 * np.abs(I_host)):
        raise RuntimeError("device-resident step differs from the host-path result")
    # pair-kernel duration from HIP events on the engine stream
    _, pair_ms = e.time_forward(10)
    sasa_ms = e.time_sasa(10)
    e.dev_free(d_b)
    e.dev_free(d_out)
    e.close()
    return w, I_host, ms_per_step, pair_ms, tuning, sasa_ms


def sasa_summary(w, ms):
    """SASA at its algorithmic unit: one (dot, candidate) test over every atom j != i with
    |x_i - x_j| < r_i + r_j + 2p, 10 FP32 FLOP each (distance, subtract, compare); the
    count is a property of the structure, not of the kernel (which stops scanning an atom
    once all its dots are buried, so it executes fewer)."""
    from scipy.spatial import cKDTree
    R = (w.radii.astype(np.float64) + w.probe)[w.types]
    pairs = cKDTree(w.pos.astype(np.float64)).query_pairs(2.0 * R.max(), output_type="ndarray")
    d = np.linalg.norm(w.pos[pairs[:, 0]].astype(np.float64) - w.pos[pairs[:, 1]], axis=1)
    close = int(np.count_nonzero(d < R[pairs[:, 0]] + R[pairs[:, 1]]))
    tests = 2.0 * close * 2 * w.n_half
    return {"workload": f"sasa-{len(w.pos)}-atoms-{2 * w.n_half}-dots", "ms": round(ms, 4),
            "kernels": "sasa_bin + sasa_scan_tiles/top/add + sasa_place + sasa_dots (HIP events, engine stream)",
            "tests": tests, "tests_per_s": tests / (ms * 1e-3),
            "achieved_tflops": round(10 * tests / (ms * 1e-3) / 1e12, 3),
            "frac": round(10 * tests / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4)}


def forward_50k_bench():
    """The 50,000-atom scale: plain forward of 50,000 atoms x 256 q (the incremental
    workload's structure), pair-kernel time from HIP events on the engine stream."""
    from resaxs_amd import configs
    w = configs.incremental_50000()
    n, nq = len(w.bodies), len(w.q)
    e = make_engine(w, hydrate=False)
    log(f"bench: {n} x {nq} plain forward (autotuned layout)")
    e.calc_curve(w.bodies)
    tuning = e.autotune(TUNING_CACHE)
    e.calc_curve(w.bodies)
    _, pair_ms = e.time_forward(5)
    e.close()
    flops = configs.forward_flops(n, nq)
    return {"workload": f"plain-{n}x{nq}", "kernel_ms": round(pair_ms, 3),
            "terms_per_s": configs.pair_terms(n, nq) / (pair_ms * 1e-3),
            "achieved_tflops": round(flops / (pair_ms * 1e-3) / 1e12, 2),
            "frac": round(flops / (pair_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
            "kernel": f"variant {tuning['variant']} K={tuning['k']} L={tuning['col_range']}"}


def adjoint_bench(reps=3):
    """Forward + adjoint (chi^2 gradient) on the fit configuration (10,000 x 200).
    Device times (HIP events): the default forward pair kernel, the fit path's FP64
    forward pair kernel and the adjoint pass; plus host wall clock of the C-ABI calls."""
    from resaxs_amd import configs
    w = configs.fit_10000()
    n, nq = len(w.bodies), len(w.q)
    e = make_engine(w, hydrate=False)
    I_ref = e.calc_curve(w.bodies)
    e.autotune(TUNING_CACHE)
    e.calc_curve(w.start_bodies)
    _, ms_fwd_default = e.time_forward(5)
    default = e.get_tuning()
    rng = w.noise_rng
    i_exp = (I_ref * (1.0 + 0.01 * rng.normal(size=nq))).astype(np.float32)
    sigma = (0.01 * I_ref).astype(np.float32)
    e.chi2_grad(w.start_bodies, i_exp, sigma)   # moves the engine to the gradient forward
    t0 = time.perf_counter()
    for _ in range(reps):
        e.chi2_grad(w.start_bodies, i_exp, sigma)
    wall = (time.perf_counter() - t0) * 1e3 / reps
    _, ms_fwd_fit = e.time_forward(5)
    fit = e.get_tuning()
    fit_kernel = f"variant {fit['variant']} K={fit['k']}"
    ms_adj = e.time_adjoint(5)
    e.close()
    terms = n * (n - 1) / 2.0 * nq
    return {"workload": w.name, "ms_fwd_fit": round(ms_fwd_fit, 4),
            "fwd_fit_kernel": fit_kernel, "ms_adjoint": round(ms_adj, 4),
            "ratio_fwd_adjoint_to_fwd": round((ms_fwd_fit + ms_adj) / ms_fwd_fit, 3),
            "ms_fwd_fastest_fp32": round(ms_fwd_default, 4),
            "fwd_fastest_kernel": f"variant {default['variant']} K={default['k']}",
            "ratio_to_fastest_fp32_fwd": round((ms_fwd_fit + ms_adj) / ms_fwd_default, 3),
            "terms_per_s": terms / ((ms_fwd_fit + ms_adj) * 1e-3),
            "ms_wall_chi2_grad": round(wall, 3),
            "note": "the gradient path computes I(q) with the per-lane-sum FP32 recurrence in runs of at most "
                    "51 q (chi^2 weights amplify forward error by I/sigma, and the recurrence error grows "
                    "with the run length); ratio_fwd_adjoint_to_fwd compares with that forward, "
                    "ratio_to_fastest_fp32_fwd with the FP32 forward of the same layout",
            "timing": "device: HIP events on the engine stream; wall: host clock around rsx_chi2_grad"}


def incremental_bench(hydrate, n_moves=2000):
    """The incremental workload: 50,000 atoms x 256 q, n_moves committed rigid moves of
    500-atom segments, running FP64 I(q) re-baselined every 200 moves.  Host wall clock
    around recalc_curve + commit (the synchronous calls a Monte Carlo caller makes),
    work per move 2 P_move (5 n_q + 9) with P_move = m (N - m) + m (m - 1) / 2.  The
    running I(q) after the next-to-last move (with 2000 moves: move 1999, 199 moves
    since the last re-baseline) is compared with a fresh full recompute by the FP64
    pair kernel."""
    from resaxs_amd import configs
    w = configs.incremental_50000()
    n, nq, m = len(w.pos), len(w.q), w.seg
    e = make_engine(w, hydrate=hydrate)
    e.set_rebaseline(200)
    pos = w.pos.copy()
    b = configs.bodies_of(pos, w.types)
    e.calc_curve(b)
    rng = w.move_rng
    log(f"bench: incremental, {n_moves} moves of {m} atoms on {n} atoms, hydration={hydrate}")
    ts, err_last = [], None
    for k in range(n_moves):
        start, newp = configs.random_move(rng, pos, m)
        pos[start:start + m] = newp
        b[start:start + m, :3] = newp
        t0 = time.perf_counter()
        I = e.recalc_curve(b, start, m)
        e.commit()
        ts.append(time.perf_counter() - t0)
        if k == n_moves - 2 and n_moves >= 2:
            f = make_engine(w, hydrate=hydrate)
            f.set_variant(0)
            I_full = f.calc_curve(b)
            err_last = float(np.max(np.abs(I - I_full) / np.abs(I_full)))
            f.close()
    e.close()
    p_move = m * (n - m) + m * (m - 1) / 2.0
    ms = float(np.mean(ts)) * 1e3
    return {"workload": f"{w.name}{'-sasa' if hydrate else ''}", "moves": n_moves, "segment": m,
            "rebaseline_every": 200, "ms_per_move_mean": round(ms, 4),
            "ms_per_move_median": round(float(np.median(ts)) * 1e3, 4),
            "terms_per_s": 2 * p_move * nq / (ms * 1e-3),
            "achieved_tflops": round(2 * p_move * (5 * nq + 9) / (ms * 1e-3) / 1e12, 2),
            "rel_err_vs_fp64_full": err_last,
            "moves_since_rebaseline_at_check": (n_moves - 2) % 200 + 1,
            "timing": "host wall clock around rsx_recalc_curve_f64 + rsx_commit (host bodies in, I(q) out)"
                      + ("; SASA recounted for the atoms within reach of the moved segment" if hydrate else "")}


def batched_bench(reps=5):
    """The batched ensemble: 512 frames of 1,000-15,000 atoms x 128 q, one graph-captured
    launch per replay; per-frame I(q) and the uniform mean to device buffers."""
    from resaxs_amd import Engine, configs
    w = configs.ensemble_512()
    nq = len(w.q)
    e = Engine(0)
    e.set_q(w.q)
    e.set_factors(w.table, w.water_index)
    e.initialize(1)
    t0 = time.perf_counter()
    e.batch_set_frames(w.bodies, w.offsets)
    layout_s = time.perf_counter() - t0
    fr = e.dev_alloc(len(w.sizes) * nq * 8)
    mn = e.dev_alloc(nq * 8)
    e.batch_calc_dev_async(fr, mn)
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        e.batch_calc_dev_async(fr, mn)
    e.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    tuning = e.get_tuning()
    launch = e.batch_launch()
    e.dev_free(fr)
    e.dev_free(mn)
    e.close()
    terms = float(sum(s * (s - 1) / 2.0 for s in w.sizes)) * nq
    flops = float(sum(configs.forward_flops(int(s), nq) for s in w.sizes))
    return {"workload": f"ensemble-{len(w.sizes)}-frames-1000..15000x{nq}", "ms_per_batch": round(ms, 3),
            "terms_per_s": terms / (ms * 1e-3), "achieved_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
            "frac": round(flops / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4), "layout_s": round(layout_s, 3),
            "kernel": f"variant {tuning['variant']} K={tuning['k']} " + (
                f"(work queue, {launch['persistent_waves']} resident waves)" if launch["persistent_waves"] > 0
                else "(one wave per (job, q chunk) item)") + f", {launch['n_jobs']} jobs",
            "timing": "host wall clock over graph replays on device-resident frames, outputs to device"}


def fit_driver_bench(n_models=8, n=4000, nq=100):
    """The calc_saxs ensemble fit (grid search over expansion factor and water weight with
    the least-squares scale) through rsx_fit_ensemble: n_models conformations of n atoms,
    nq q in [0.01, 0.5], target = 2 x the ensemble mean at c = 1.012, w = 0.8 (the
    engine's own, hydrated), 1 % sigmas.  Host wall clock of the whole search."""
    from resaxs_amd import configs, fit_ensemble
    rng = np.random.default_rng(6006)
    radius = (n / 0.045 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
    base = configs.in_ball(rng, n, radius)
    types = rng.choice(np.array([1, 2, 3, 5, 7, 8, 9, 10, 13], dtype=np.uint32), size=n).astype(np.uint32)
    bodies = np.concatenate([configs.bodies_of((base + rng.normal(scale=0.5, size=base.shape)), types)
                             for _ in range(n_models)])
    offsets = np.arange(n_models + 1, dtype=np.int32) * n
    q = np.linspace(0.01, 0.5, nq).astype(np.float32)
    ones = np.ones(nq, dtype=np.float32)
    c0, w0 = 1.012, 0.8
    *_, iq = fit_ensemble(bodies, offsets, q, ones, ones, scale=1.0, water=(w0, w0), expansion=(c0, c0))
    target = (2.0 * iq).astype(np.float32)
    t0 = time.perf_counter()
    c, w, s, chi, ev, _ = fit_ensemble(bodies, offsets, q, target, (0.01 * target).astype(np.float32))
    sec = time.perf_counter() - t0
    return {"workload": f"fit-{n_models}-models-{n}x{nq}", "seconds": round(sec, 3), "evaluations": ev,
            "ms_per_model_evaluation": round(sec * 1e3 / ev, 4), "found": [round(c, 5), round(w, 4), round(s, 5)],
            "target": [c0, w0, 2.0], "chi": chi,
            "timing": "host wall clock around rsx_fit_ensemble (engines created, laid out and SASA counted inside)"}


def histogram_bench(bin_width=0.01, reps=3):
    """The distance-histogram Debye (approximate) on the 50,000-atom plain workload x 256 q:
    host wall clock of rsx_calc_curve_hist (bodies from the host, I(q) back) against the
    exact engine's forward on the same engine, and its error relative to that forward."""
    from resaxs_amd import configs
    w = configs.incremental_50000()
    n, nq = len(w.bodies), len(w.q)
    e = make_engine(w, hydrate=False)
    I_exact = e.calc_curve(w.bodies)
    _, pair_ms = e.time_forward(3)
    e.calc_curve_hist(w.bodies, bin_width)
    t0 = time.perf_counter()
    for _ in range(reps):
        I_h = e.calc_curve_hist(w.bodies, bin_width)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    e.close()
    return {"workload": f"plain-{n}x{nq}", "bin_width_A": bin_width, "ms_wall": round(ms, 3),
            "pairs_per_s": n * (n - 1) / 2.0 / (ms * 1e-3), "exact_pair_kernel_ms": round(pair_ms, 3),
            "max_rel_diff_vs_exact": float(np.max(np.abs(I_h - I_exact) / np.abs(I_exact))),
            "timing": "host wall clock around rsx_calc_curve_hist (histogram, transform, host combination)"}


def cpu_baseline(w, I_gpu, rows=2560):
    """The reference's CPU path timed on this host, 1 thread, on a bounded sample: the
    pairs among the first `rows` atoms (the first rows of the lower triangle) at all
    n_q.  kind "reference": host_debye<double, float, cl_float4>::calc_curve compiled
    from the reference's sources (oracle/_ref); kind "port": its restatement in the
    oracle where that library was not built.  Plus the FP64 oracle's full I(q) as the
    accuracy check of the GPU result."""
    from oracle import pyoracle as po
    nq = len(w.q)
    terms = rows * (rows - 1) / 2.0 * nq
    pos, types = w.pos[:rows], w.types[:rows]
    if po.ref_lib() is not None:
        log(f"bench: CPU baseline, reference host_debye over the first {rows} atoms, 1 thread")
        t0 = time.perf_counter()
        po.ref_host_debye(w.q, pos, types, w.table)
        t = time.perf_counter() - t0
        kind, what = "reference", "host_debye<double,float,cl_float4>::calc_curve (host_debye.cpp:54-109, " \
                                  "compiled from the reference's sources)"
    else:
        log(f"bench: CPU baseline, restated reference loop over the first {rows} atoms, 1 thread")
        t0 = time.perf_counter()
        po.debye_ref_rows(w.q, pos, types, w.table, 0, rows, nthreads=1)
        t = time.perf_counter() - t0
        kind, what = "port", "the oracle's restatement of host_debye.cpp:54-109"
    log("bench: FP64 oracle of the full workload (accuracy check)")
    dots = np.stack([po.fibonacci_dots(w.n_half, r, w.probe) for r in w.radii])
    rr = (w.radii.astype(np.float32) + np.float32(w.probe)).astype(np.float32)
    counts = po.sasa_counts(w.pos, w.types, (rr * rr).astype(np.float32), dots)
    s = (counts.astype(np.float32) / np.float32(dots.shape[1])).astype(np.float32)
    F = po.body_factors(w.pos, w.types, w.q, w.table, w.water_index, w.w, s)
    I_ref = po.debye(w.q, w.pos, F)
    err = np.abs(I_gpu - I_ref) / np.abs(I_ref)
    base = {"value": terms / t, "unit": "pair.q terms/s", "cores": 1, "kind": kind,
            "sample": f"{what} on the first {rows} atoms of the {len(w.pos)}-atom workload (plain factors) x "
                      f"{nq} q = {terms:.3e} pair.q terms in {t:.1f} s",
            "cpu": platform.processor() or platform.machine()}
    acc = {"max": float(err.max()), "median": float(np.median(err)), "reference": "FP64 oracle, full I(q)"}
    return base, acc


def cpu_baseline_all_cores(w, rows=6400):
    """The same loop on all of this host's cores (at most 16, the host share of one GPU):
    the oracle's OpenMP restatement of host_debye.cpp:54-109 over the pairs among the
    first `rows` atoms at all n_q (the reference's own loop is single-threaded)."""
    from oracle import pyoracle as po
    nq = len(w.q)
    threads = max(1, min(16, os.cpu_count() or 1))
    terms = rows * (rows - 1) / 2.0 * nq
    log(f"bench: CPU baseline on {threads} threads, restated reference loop over the first {rows} atoms")
    t0 = time.perf_counter()
    po.debye_ref_rows(w.q, w.pos[:rows], w.types[:rows], w.table, 0, rows, nthreads=threads)
    t = time.perf_counter() - t0
    return {"value": terms / t, "unit": "pair.q terms/s", "cores": threads, "kind": "port",
            "sample": f"the oracle's restatement of host_debye.cpp:54-109 (OpenMP over rows) on the first {rows} "
                      f"atoms of the {len(w.pos)}-atom workload (plain factors) x {nq} q = {terms:.3e} pair.q "
                      f"terms in {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the batched-ensemble and incremental-move measurements")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--section", choices=("fwd_adjoint", "incremental", "incremental_sasa", "batched"),
                    help="run only this workload section, shortened for profiling (incremental: 100 moves; "
                         "batched: 2 replays), and print its JSON")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.section:
        run = {"fwd_adjoint": lambda: adjoint_bench(),
               "incremental": lambda: incremental_bench(False, n_moves=100),
               "incremental_sasa": lambda: incremental_bench(True, n_moves=100),
               "batched": lambda: batched_bench(reps=2)}[args.section]
        print(json.dumps({"section": args.section, "result": run()}), flush=True)
        return
    from resaxs_amd import configs
    if args.copy_probe:
        from resaxs_amd import Engine
        e = Engine(0)
        ms = e.time_copy(args.copy_probe << 20, 3)
        log(f"bench: copy probe {args.copy_probe} MiB: {ms:.4f} ms/launch, "
            f"{2 * (args.copy_probe << 20) / ms / 1e6:.1f} GB/s (read + write)")
        e.close()
    w, I_gpu, ms_per_step, pair_ms, tuning, sasa_ms = forward_bench(args.steps, args.warmup)
    kernel = ("debye_uniform_kernel", "debye_shared_kernel", "debye_rec32_kernel",
              "debye_rec32r_kernel")[tuning["variant"]]
    n, nq = len(w.pos), len(w.q)
    terms = configs.pair_terms(n, nq)
    flops = configs.forward_flops(n, nq)
    achieved = flops / (pair_ms * 1e-3) / 1e12
    line = {
        "metric": "Debye pair.q terms/s (forward, device-resident step incl. SASA hydration)",
        "value": terms / (ms_per_step * 1e-3),
        "unit": "pair.q terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded)",
        "config": {"workload": w.name, "n_atoms": n, "n_q": nq, "q_max": float(w.q[-1]),
                   "hydration": "SASA, 256 Fibonacci dots, probe 1.4 A, contrast 0.03 e/A^3"},
        "roofline": {"bound": "mfma", "pipe": "VALU (FP32 + FP64 vector; no MFMA)",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": read_traffic(),
                     "kernel": f"{kernel}<{tuning['k']}>", "col_range": tuning["col_range"],
                     "kernel_ms": round(pair_ms, 4), "flop_per_launch": flops},
        "sasa": sasa_summary(w, sasa_ms),
    }
    if not args.no_adjoint:
        line["forward_50k"] = forward_50k_bench()
        log("bench: forward + adjoint on the fit configuration")
        line["fwd_adjoint"] = adjoint_bench()
    if not args.no_workloads:
        log("bench: batched ensemble")
        line["batched"] = batched_bench()
        line["incremental"] = [incremental_bench(False), incremental_bench(True)]
        log("bench: ensemble fit driver")
        line["fit_driver"] = fit_driver_bench()
        log("bench: distance-histogram Debye")
        line["histogram"] = histogram_bench()
    if not args.no_cpu_baseline:
        base, acc = cpu_baseline(w, I_gpu)
        line["cpu_baseline"] = base
        line["cpu_baseline_all_cores"] = cpu_baseline_all_cores(w)
        line["rel_err_vs_fp64"] = acc
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
