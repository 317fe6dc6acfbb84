#!/usr/bin/env python3
"""Benchmark of the Debye-SAXS hot path on one MI355X.

Headline workload (plain-50000x256): 50,000 atoms at 0.045 A^-3 in a 1:1.3:1.8
ellipsoid (the incremental workload's structure, plain factors), 256 q over 0-0.6 A^-1.
A step is one full device-resident forward pass on positions already in HBM: columns
gathered from the bodies -> Debye pair kernel -> fixed-order FP64 reduction.  Every
step runs on its own set of positions (each atom jittered by up to 0.2 A from the
laid-out structure, all sets uploaded before the timed region), so the timed path
reuses the layout's slot order and job table but not its positions.

Prints ONE JSON line.  value = forward pair.q terms/s of the step; roofline = the pair
kernel's algorithmic FP32 FLOP rate (P (5 n_q + 9) + 2 N n_q per launch) over its
HIP-event duration on the engine stream, against the FP32 vector peak.  Further keys
carry the other workloads (hydrated 20k step with SASA, forward + adjoint, batched
ensemble, incremental moves, fit driver, histogram, non-uniform q) and the CPU
baselines.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBPS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")
# HBM bytes per launch from the committed rocprofv3 counter summaries
TRAFFIC_SUMMARY = os.path.join(PROFILES, "debye_pair_kernel_plain-50000x256_counters.txt")
SASA_SUMMARY = os.path.join(PROFILES, "sasa_kernels_sasa-hydrated-20000x256_counters.txt")
# full-recompute period of the incremental workload: every 200 moves, as the workload defines
# it; the same moves are also timed with the engine's default period (10000: the running FP64
# I(q) does not drift measurably over the 2000 moves, so no full recompute falls inside them)
REBASELINE = 200
REBASELINE_DEFAULT = 10000
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


def read_summary(path, key):
    """A number `key: value` from a committed rocprofv3 summary (None if absent)."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(re.escape(key) + r":\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def jittered(w, n_sets, seed=2024, shrink=0.99, amp=0.2):
    """n_sets position sets: the structure pulled 1 % towards its centroid, every atom then
    moved by up to amp A per axis (the sets stay inside the laid-out bounding radius)."""
    rng = np.random.default_rng(seed)
    c = w.pos.astype(np.float64).mean(axis=0)
    from resaxs_amd import configs
    out = []
    for _ in range(n_sets):
        p = c + shrink * (w.pos - c) + rng.uniform(-amp, amp, size=w.pos.shape)
        out.append(configs.bodies_of(p.astype(np.float32), w.types))
    return out


def device_steps(e, sets, steps, warmup):
    """Upload the position sets, run warmup steps, then time `steps` device-resident forward
    passes, each on the next set (host wall clock between two stream synchronisations).
    Returns (ms per step, the device buffers' last I(q), the bodies of the last step)."""
    nq = e.nq
    n = len(sets[0])
    bufs = []
    for b in sets:
        p = e.dev_alloc(b.nbytes)
        e.h2d(p, b)
        bufs.append(p)
    d_out = e.dev_alloc(nq * 8)
    for k in range(warmup):
        e.calc_curve_dev_async(bufs[k % len(bufs)], n, d_out)
    e.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        e.calc_curve_dev_async(bufs[(warmup + k) % len(bufs)], n, d_out)
    e.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    I = e.d2h(np.zeros(nq, dtype=np.float64), d_out)
    last = sets[(warmup + steps - 1) % len(sets)]
    for p in bufs:
        e.dev_free(p)
    e.dev_free(d_out)
    return ms, I, last


def forward_bench(steps, warmup):
    """The headline: plain 50,000 x 256 forward, device-resident steps on fresh positions."""
    from resaxs_amd import configs
    w = configs.incremental_50000()
    n, nq = len(w.bodies), len(w.q)
    e = make_engine(w, hydrate=False)
    log(f"bench: plain-{n}x{nq}: host calc_curve (layout + first result)")
    e.calc_curve(w.bodies)
    tuning = e.autotune(TUNING_CACHE)
    log(f"bench: tuning {tuning}")
    e.calc_curve(w.bodies)
    sets = jittered(w, min(steps + warmup, 32))
    log(f"bench: timing {steps} device-resident steps on fresh positions")
    ms_per_step, I_dev, last = device_steps(e, sets, steps, warmup)
    # pair-kernel duration from HIP events on the engine stream (the last step's columns)
    _, pair_ms = e.time_forward(10)
    t = e.get_tuning()
    # accuracy of the last step: the same positions laid out afresh and summed by the FP64
    # recurrence kernel (the oracle's 50,000-atom sum takes minutes on the host; the FP64
    # kernel is held to it in the GPU tests)
    f = make_engine(w, hydrate=False)
    f.set_variant(0)
    I_f64 = f.calc_curve(last)
    f.close()
    e.close()
    err = np.abs(I_dev - I_f64) / np.abs(I_f64)
    return w, ms_per_step, pair_ms, t, {"max": float(err.max()), "median": float(np.median(err)),
                                         "reference": "GPU FP64 three-term recurrence kernel on the last step's "
                                                      "positions, laid out afresh"}


def hydrated_bench(steps, warmup):
    """The 20,000-atom SASA-hydrated forward (256 Fibonacci dots, probe 1.4 A, contrast
    0.03 e/A^3) x 256 q: device-resident steps on fresh positions (SASA counts, classes,
    pair kernel, reduction on the device each step); the SASA count timed on its own."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000()
    n, nq = len(w.bodies), len(w.q)
    e = make_engine(w)
    log(f"bench: {w.name}: host calc_curve")
    I_host = e.calc_curve(w.bodies)
    tuning = e.autotune(TUNING_CACHE)
    I_host = e.calc_curve(w.bodies)
    sets = jittered(w, min(steps + warmup, 32), seed=2025)
    ms_per_step, _, _ = device_steps(e, sets, steps, warmup)
    _, pair_ms = e.time_forward(10)
    # the laid-out positions once more: the device path equals the host path (same kernels and
    # layout; only the q = 0 value is summed in another order)
    _, same, _ = device_steps(e, [w.bodies], 1, 0)
    if not np.all(np.abs(same - I_host) <= 1e-12 * np.abs(I_host)):
        raise RuntimeError("device-resident step differs from the host-path result")
    sasa_ms = e.time_sasa(10)
    t = e.get_tuning()
    e.close()
    flops = configs.forward_flops(n, nq)
    kernel = ("debye_uniform_kernel", "debye_shared_kernel", "debye_rec32_kernel", "debye_rec32r_kernel")[t["variant"]]
    return w, I_host, {
        "workload": w.name, "ms_per_step": round(ms_per_step, 4),
        "terms_per_s": configs.pair_terms(n, nq) / (ms_per_step * 1e-3),
        "pair_kernel": f"{kernel}<{t['k']}> L={t['col_range']}", "pair_kernel_ms": round(pair_ms, 4),
        "achieved_tflops": round(flops / (pair_ms * 1e-3) / 1e12, 3),
        "frac": round(flops / (pair_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
        "positions": "every step on its own jittered position set (up to 0.2 A per axis); the layout's class "
                     "order is that of the laid-out positions, so changed exposed-dot counts break class runs",
        "timing": "host wall clock over device-resident steps; pair kernel from HIP events on the engine stream",
    }, sasa_ms


def sasa_summary(w, ms, probe=None, n_dots=None, hbm=True):
    """SASA at its algorithmic unit: one (dot, candidate) test over every atom j != i with
    |x_i - x_j| < r_i + r_j + 2p, 10 FP32 FLOP each (distance, subtract, compare); the
    count is a property of the structure, not of the kernel (which stops scanning an atom
    once all its dots are buried, so it executes fewer).  HBM bytes per count from the
    committed counter summary of the SASA kernels (taken on the 256-dot sphere)."""
    from scipy.spatial import cKDTree
    probe = w.probe if probe is None else probe
    n_dots = 2 * w.n_half if n_dots is None else n_dots
    R = (w.radii.astype(np.float64) + probe)[w.types]
    pairs = cKDTree(w.pos.astype(np.float64)).query_pairs(2.0 * R.max(), output_type="ndarray")
    d = np.linalg.norm(w.pos[pairs[:, 0]].astype(np.float64) - w.pos[pairs[:, 1]], axis=1)
    close = int(np.count_nonzero(d < R[pairs[:, 0]] + R[pairs[:, 1]]))
    tests = 2.0 * close * n_dots
    hbm = read_summary(SASA_SUMMARY, "hbm_bytes_per_count") if hbm else None
    out = {"workload": f"sasa-{len(w.pos)}-atoms-{n_dots}-dots-probe-{probe}", "ms": round(ms, 4),
           "kernels": "fill + sasa_bin + sasa_scan_tiles/top/add + sasa_place + sasa_dots (HIP events, engine stream)",
           "tests": tests, "tests_per_s": tests / (ms * 1e-3),
           "achieved_tflops": round(10 * tests / (ms * 1e-3) / 1e12, 3),
           "frac": round(10 * tests / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
           "hbm_bytes": hbm}
    if hbm:
        out["achieved_GBps"] = round(hbm / (ms * 1e-3) / 1e9, 1)
        out["hbm_frac"] = round(hbm / (ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4)
    return out


def sasa_reference_sphere_bench(reps=20):
    """The SASA count on the 20,000-atom ellipsoid with the reference GPU hydration's dot
    sphere: 32 dodecahedron + icosahedron dots, probe 1.8 A (saxs_alg_cl_gpu_pt_wf.cpp:91,
    rsasa_cl_gpu_alg.cpp:73)."""
    from resaxs_amd import RSX_DOTS_DODECA_ICOSA, configs
    w = configs.sasa_ellipsoid_20000()
    e = make_engine(w)
    e.set_sasa_params(w.radii, 1.8, RSX_DOTS_DODECA_ICOSA, 0)
    e.calc_curve(w.bodies)
    ms = e.time_sasa(reps)
    e.close()
    return sasa_summary(w, ms, probe=1.8, n_dots=32, hbm=False)


def direct_bench(reps=5):
    """A non-uniform q grid (the direct kernel: one v_sin per term, what a measured profile's
    q values take): 20,000 plain atoms x 256 q spaced as q_k = 0.6 (k / 255)^1.5."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000()
    q = (0.6 * (np.arange(256) / 255.0) ** 1.5).astype(np.float32)
    n, nq = len(w.bodies), len(q)
    from resaxs_amd import Engine
    e = Engine(0)
    e.set_q(q)
    rng = np.random.default_rng(77)
    e.set_factors(configs.factor_table(rng, q), configs.WATER)
    e.initialize(n)
    e.calc_curve(w.bodies)
    t = e.get_tuning()
    _, pair_ms = e.time_forward(reps)
    e.close()
    flops = configs.forward_flops(n, nq)
    return {"workload": f"plain-{n}x{nq}-nonuniform-q", "pair_kernel": f"debye_direct_kernel<{t['k']}>",
            "recurrence": bool(t["recurrence"]), "pair_kernel_ms": round(pair_ms, 4),
            "terms_per_s": configs.pair_terms(n, nq) / (pair_ms * 1e-3),
            "achieved_tflops": round(flops / (pair_ms * 1e-3) / 1e12, 3),
            "frac": round(flops / (pair_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
            "timing": "HIP events on the engine stream"}


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
    lay = e.adjoint_layout()
    e.close()
    terms = n * (n - 1) / 2.0 * nq
    w_adj = n * (n - 1) / 2.0 * (8 * nq + 19)  # adjoint FLOP: 8 per pair.q term + 19 per pair
    return {"workload": w.name, "ms_fwd_fit": round(ms_fwd_fit, 4),
            "fwd_fit_kernel": fit_kernel, "ms_adjoint": round(ms_adj, 4),
            "adjoint_kernel": f"{lay['form']} form K={lay['k']} L={lay['col_range']}",
            "adjoint_tflops": round(w_adj / (ms_adj * 1e-3) / 1e12, 2),
            "adjoint_frac": round(w_adj / (ms_adj * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
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


def incremental_bench(hydrate, n_moves=2000, rebaseline=REBASELINE):
    """The incremental workload: 50,000 atoms x 256 q, n_moves committed rigid moves of
    500-atom segments, running FP64 I(q) re-baselined (a full recompute in place of the
    update) every `rebaseline` moves, those moves timed with the rest.  Host wall
    clock around recalc_curve + commit (the synchronous calls a Monte Carlo caller makes),
    work per move 2 P_move (5 n_q + 9) with P_move = m (N - m) + m (m - 1) / 2.  The
    running I(q) after the next-to-last move is compared with a fresh full recompute by the
    FP64 pair kernel."""
    from resaxs_amd import configs
    w = configs.incremental_50000()
    n, nq, m = len(w.pos), len(w.q), w.seg
    e = make_engine(w, hydrate=hydrate)
    e.set_rebaseline(rebaseline)
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
            "rebaseline_every": rebaseline, "ms_per_move_mean": round(ms, 4),
            "ms_per_move_median": round(float(np.median(ts)) * 1e3, 4),
            "rebaselines": n_moves // rebaseline,
            "ms_rebaseline_move_mean": (round(float(np.mean(ts[rebaseline - 1::rebaseline])) * 1e3, 3)
                                        if n_moves >= rebaseline else None),
            "terms_per_s": 2 * p_move * nq / (ms * 1e-3),
            "achieved_tflops": round(2 * p_move * (5 * nq + 9) / (ms * 1e-3) / 1e12, 2),
            "rel_err_vs_gpu_fp64_full": err_last,
            "rel_err_reference": "a fresh full recompute of the same positions by the GPU's FP64 three-term "
                                 "recurrence kernel (variant 0)",
            "moves_since_rebaseline_at_check": (n_moves - 2) % rebaseline + 1,
            "timing": "host wall clock around rsx_recalc_curve_f64 + rsx_commit (host bodies in, I(q) out)"
                      + ("; SASA recounted for the atoms within reach of the moved segment" if hydrate else "")}


def batched_bench(reps=5, adjoint=True):
    """The batched ensemble: 512 frames of 1,000-15,000 atoms x 128 q, one graph-captured
    launch per replay; per-frame I(q) and the uniform mean to device buffers.  With adjoint,
    also the batched VJP (gradients of every frame from per-frame weights dL/dI_f(q))."""
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
    tuning = e.batch_autotune(TUNING_CACHE)
    fr = e.dev_alloc(len(w.sizes) * nq * 8)
    mn = e.dev_alloc(nq * 8)
    e.batch_calc_dev_async(fr, mn)
    e.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        e.batch_calc_dev_async(fr, mn)
    e.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    launch = e.batch_launch()
    e.dev_free(fr)
    e.dev_free(mn)
    adj = None
    if adjoint:
        # the batch's gradients from per-frame weights dL/dI_f(q): one adjoint launch + one reduce
        wts = np.random.default_rng(7).normal(size=(len(w.sizes), nq))
        t0 = time.perf_counter()
        e.batch_vjp(wts)
        vjp_first_s = time.perf_counter() - t0
        ms_adj = e.time_batch_vjp(3)
        pairs = float(sum(s * (s - 1) / 2.0 for s in w.sizes))
        w_adj = pairs * (8 * nq + 19)  # adjoint FLOP: 8 per pair.q term + 19 per pair
        lay = e.adjoint_layout(batch=True)
        adj = {"ms_adjoint": round(ms_adj, 3), "ratio_fwd_adjoint_to_fwd": round((ms + ms_adj) / ms, 3),
               "terms_per_s_fwd_adjoint": pairs * nq / ((ms + ms_adj) * 1e-3),
               "adjoint_tflops": round(w_adj / (ms_adj * 1e-3) / 1e12, 2),
               "adjoint_frac": round(w_adj / (ms_adj * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
               "kernel": f"{lay['form']} form K={lay['k']} L={lay['col_range']}",
               "first_call_s": round(vjp_first_s, 3),
               "timing": "HIP events on the engine stream around the adjoint kernel + per-atom reduce "
                         "(weights resident); first_call_s is the host wall clock of the first rsx_batch_vjp "
                         "(job layout, uploads, gradient download)"}
    e.close()
    terms = float(sum(s * (s - 1) / 2.0 for s in w.sizes)) * nq
    flops = float(sum(configs.forward_flops(int(s), nq) for s in w.sizes))
    return {"workload": f"ensemble-{len(w.sizes)}-frames-1000..15000x{nq}", "ms_per_batch": round(ms, 3),
            "adjoint": adj,
            "terms_per_s": terms / (ms * 1e-3), "achieved_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
            "frac": round(flops / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4), "layout_s": round(layout_s, 3),
            "kernel": f"variant {tuning['variant']} K={tuning['k']} L={tuning['col_range']} " + (
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
            "timing": "host wall clock around rsx_fit_ensemble (one engine created, the models laid out as batch "
                      "frames and SASA counted inside)"}


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


def cpu_baseline(w, rows=2900):
    """The reference's CPU path timed on this host, 1 thread, on a bounded sample of the
    headline workload: the pairs among its first `rows` atoms (a compact cluster: the
    atoms are in space-filling-curve order) at all n_q.  kind "reference":
    host_debye<double, float, cl_float4>::calc_curve compiled from the reference's sources
    (oracle/_ref); kind "port": its restatement in the oracle where that library was not
    built."""
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
    return {"value": terms / t, "unit": "pair.q terms/s", "cores": 1, "kind": kind,
            "sample": f"{what} on the first {rows} atoms of the {len(w.pos)}-atom workload x {nq} q = "
                      f"{terms:.3e} pair.q terms in {t:.1f} s",
            "cpu": cpu_model()}


def cpu_config1():
    """The plain 1,888-atom ball x 101 q timed in full on the CPU (the reference's own host_debye,
    1 thread) and on the GPU (host-buffer calc_curve, wall clock), with both results against the
    FP64 oracle."""
    from oracle import pyoracle as po
    from resaxs_amd import configs
    w = configs.plain_ball_1888()
    n, nq = len(w.pos), len(w.q)
    terms = n * (n - 1) / 2.0 * nq
    out = {"workload": w.name, "terms": terms}
    F = po.body_factors(w.pos, w.types, w.q, w.table)
    I_ref = po.debye(w.q, w.pos, F)
    if po.ref_lib() is not None:
        t0 = time.perf_counter()
        I_cpu = po.ref_host_debye(w.q, w.pos, w.types, w.table)
        t = time.perf_counter() - t0
        out.update({"cpu_s": round(t, 3), "cpu_terms_per_s": terms / t, "cpu_kind": "reference (1 thread)",
                    "cpu_rel_err_vs_fp64": float(np.max(np.abs(I_cpu - I_ref) / np.abs(I_ref)))})
    e = make_engine(w, hydrate=False)
    e.calc_curve(w.bodies)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        I_gpu = e.calc_curve(w.bodies)
    t = (time.perf_counter() - t0) / reps
    e.close()
    out.update({"gpu_ms": round(t * 1e3, 4), "gpu_terms_per_s": terms / t,
                "gpu_rel_err_vs_fp64": float(np.max(np.abs(I_gpu - I_ref) / np.abs(I_ref))),
                "timing": "CPU: one full host_debye call; GPU: host wall clock of rsx_calc_curve_f64 with host "
                          "buffers (layout, upload, kernels, download)"})
    return out


def hydrated_accuracy(w, I_gpu):
    """The hydrated 20,000-atom I(q) (host path) against the FP64 oracle with the oracle's own
    exposed-dot counts."""
    from oracle import pyoracle as po
    log("bench: FP64 oracle of the hydrated 20,000-atom workload (accuracy check)")
    dots = np.stack([po.fibonacci_dots(w.n_half, r, w.probe) for r in w.radii])
    rr = (w.radii.astype(np.float32) + np.float32(w.probe)).astype(np.float32)
    counts = po.sasa_counts(w.pos, w.types, (rr * rr).astype(np.float32), dots)
    s = (counts.astype(np.float32) / np.float32(dots.shape[1])).astype(np.float32)
    F = po.body_factors(w.pos, w.types, w.q, w.table, w.water_index, w.w, s)
    I_ref = po.debye(w.q, w.pos, F)
    err = np.abs(I_gpu - I_ref) / np.abs(I_ref)
    return {"max": float(err.max()), "median": float(np.median(err)), "reference": "FP64 oracle, full I(q)"}


def cpu_baseline_all_cores(w, rows=11000):
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
                      f"atoms of the {len(w.pos)}-atom workload x {nq} q = {terms:.3e} pair.q terms in {t:.1f} s",
            "cpu": cpu_model()}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the batched-ensemble, incremental-move, fit-driver, histogram and non-uniform-q "
                         "measurements")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--section", choices=("forward_50k", "hydrated_20k", "sasa", "fwd_adjoint", "incremental",
                                          "incremental_sasa", "batched", "direct"),
                    help="run only this workload section, shortened for profiling (incremental: 100 moves; "
                         "batched: 2 replays), and print its JSON")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from resaxs_amd import configs
    if args.copy_probe:
        from resaxs_amd import Engine
        e = Engine(0)
        ms = e.time_copy(args.copy_probe << 20, 3)
        log(f"bench: copy probe {args.copy_probe} MiB: {ms:.4f} ms/launch, "
            f"{2 * (args.copy_probe << 20) / ms / 1e6:.1f} GB/s (read + write)")
        e.close()
    if args.section:
        def fwd():
            w, ms, pair_ms, t, acc = forward_bench(args.steps, args.warmup)
            return {"ms_per_step": ms, "pair_kernel_ms": pair_ms, "tuning": t, "rel_err": acc}

        def hyd():
            return hydrated_bench(args.steps, args.warmup)[2]

        def sasa():
            w = configs.sasa_ellipsoid_20000()
            e = make_engine(w)
            e.calc_curve(w.bodies)
            ms = e.time_sasa(20)
            e.close()
            return {"fibonacci_256": sasa_summary(w, ms), "dodeca_icosa_32": sasa_reference_sphere_bench()}

        run = {"forward_50k": fwd, "hydrated_20k": hyd, "sasa": sasa,
               "fwd_adjoint": lambda: adjoint_bench(),
               "incremental": lambda: incremental_bench(False, n_moves=100),
               "incremental_sasa": lambda: incremental_bench(True, n_moves=100),
               "batched": lambda: batched_bench(reps=2),
               "direct": lambda: direct_bench()}[args.section]
        print(json.dumps({"section": args.section, "result": run()}), flush=True)
        return
    w, ms_per_step, pair_ms, tuning, acc = forward_bench(args.steps, args.warmup)
    kernel = ("debye_uniform_kernel", "debye_shared_kernel", "debye_rec32_kernel",
              "debye_rec32r_kernel")[tuning["variant"]]
    n, nq = len(w.pos), len(w.q)
    terms = configs.pair_terms(n, nq)
    flops = configs.forward_flops(n, nq)
    achieved = flops / (pair_ms * 1e-3) / 1e12
    line = {
        "metric": "Debye pair.q terms/s (forward, device-resident step on fresh positions)",
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
        "data": "synthetic (seeded); each step on its own jittered position set, uploaded before timing",
        "config": {"workload": f"plain-{n}x{nq}", "n_atoms": n, "n_q": nq, "q_max": float(w.q[-1]),
                   "hydration": "none (plain factors); the SASA-hydrated 20,000-atom step is under hydrated_20k"},
        "roofline": {"bound": "mfma", "pipe": "VALU (FP32 + FP64 vector; no MFMA)",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": read_summary(TRAFFIC_SUMMARY, "traffic_bytes_per_launch"),
                     "kernel": f"{kernel}<{tuning['k']}>", "col_range": tuning["col_range"],
                     "kernel_ms": round(pair_ms, 4), "flop_per_launch": flops},
        "rel_err": acc,
    }
    hw, I_hyd, hyd, sasa_ms = hydrated_bench(args.steps, args.warmup)
    hyd["sasa"] = sasa_summary(hw, sasa_ms)
    hyd["sasa_reference_sphere"] = sasa_reference_sphere_bench()
    line["hydrated_20k"] = hyd
    if not args.no_adjoint:
        log("bench: forward + adjoint on the fit configuration")
        line["fwd_adjoint"] = adjoint_bench()
    if not args.no_workloads:
        log("bench: batched ensemble")
        line["batched"] = batched_bench()
        line["incremental"] = [incremental_bench(False), incremental_bench(True)]
        line["incremental_rebaseline_10000"] = [incremental_bench(False, rebaseline=REBASELINE_DEFAULT),
                                                incremental_bench(True, rebaseline=REBASELINE_DEFAULT)]
        log("bench: ensemble fit driver")
        line["fit_driver"] = fit_driver_bench()
        log("bench: distance-histogram Debye")
        line["histogram"] = histogram_bench()
        log("bench: non-uniform q grid (direct kernel)")
        line["direct_q"] = direct_bench()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w)
        line["cpu_baseline_all_cores"] = cpu_baseline_all_cores(w)
        line["cpu_config1"] = cpu_config1()
        line["hydrated_20k"]["rel_err_vs_fp64"] = hydrated_accuracy(hw, I_hyd)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
