"""Batched ensemble I(q) and the adjoint (dchi^2/dx) kernel.

Batched: per-frame I(q) and the uniform-weight mean of
calc_saxs/calc_saxs.hpp:611-624 (avg_ensemble), each frame against the FP64
oracle.  Adjoint: against the oracle's analytic FP64 gradient and against
central finite differences (h = 1e-3 A) of the oracle's chi^2 with the scale
re-optimised at every point (envelope theorem), bar 1e-4 relative.
"""
import sys

import numpy as np
import pytest

from helpers import engine_for, oracle_iq, rel_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# (variant, K, persistent): the default run-flush kernel at K = 128 runs one wave per
# item (its work-queue form would lose occupancy); at K = 64 it takes the work queue
@pytest.mark.parametrize("variant,k,persistent", [(3, 0, True), (3, 64, True), (3, 64, False), (2, 0, True),
                                                  (2, 0, False)])
def test_batch_frames_and_mean(oracle, variant, k, persistent):
    from resaxs_amd import configs
    _say(f"case: batch of 24 frames, 100..2500 atoms x 128 q, variant {variant} K={k} persistent={persistent}")
    w = configs.ensemble_512(n_frames=24, nmin=100, nmax=2500)
    e = engine_for(w.q, w.table, w.water_index, 0.0)
    e.set_variant(variant)
    if k:
        e.set_tuning(k, 0)
    e.set_persistent(persistent)
    e.batch_set_frames(w.bodies, w.offsets)
    frames, mean = e.batch_calc()
    frames2, mean2 = e.batch_calc()          # graph replay gives the same bits
    assert np.array_equal(frames, frames2) and np.array_equal(mean, mean2)
    launch = e.batch_launch()
    assert launch["n_jobs"] > 0
    # the work queue only where asked for and where it keeps the kernel's occupancy (the
    # run-flush kernel at K = 64 does; at K = 128 its queue form would drop to two waves per
    # SIMD, and the engine launches one wave per item instead)
    _say(f"  launch: {launch}")
    if not persistent:
        assert launch["persistent_waves"] == 0
    if persistent and (variant, k) == (3, 64):
        assert launch["persistent_waves"] > 0
    refs = []
    for f in range(24):
        b = w.bodies[w.offsets[f]:w.offsets[f + 1]]
        pos = b[:, :3]
        types = b[:, 3].view(np.uint32)
        refs.append(oracle_iq(w.q, pos, types, w.table))
        assert rel_err(frames[f], refs[-1]).max() < 1e-5, f
    assert rel_err(mean, np.mean(refs, axis=0)).max() < 1e-5


def test_batch_full_size(oracle):
    """The batched workload at full size: 512 frames of 1,000-15,000 atoms x 128 q in one
    graph-captured launch.  Sampled frames (smallest, largest and four more) against the
    FP64 oracle; every frame's q = 0 value is (sum F)^2; the mean is the frames' average;
    a graph replay reproduces every bit; the persistent work-queue launch at K = 64
    agrees to 1e-5."""
    from resaxs_amd import configs
    w = configs.ensemble_512()
    nf = len(w.sizes)
    _say(f"case: batched full size, {nf} frames, {int(w.sizes.sum())} atoms x 128 q")
    e = engine_for(w.q, w.table, w.water_index, 0.0)
    e.batch_set_frames(w.bodies, w.offsets)
    frames, mean = e.batch_calc()
    frames2, mean2 = e.batch_calc()
    assert np.array_equal(frames, frames2) and np.array_equal(mean, mean2)
    assert np.allclose(mean, frames.mean(axis=0), rtol=1e-13, atol=0)
    t = w.bodies[:, 3].view(np.uint32)
    for f in range(nf):
        F0 = w.table[0][t[w.offsets[f]:w.offsets[f + 1]]].astype(np.float64).sum()
        assert abs(frames[f, 0] - F0 * F0) <= 1e-12 * F0 * F0, f
    pick = sorted({int(np.argmin(w.sizes)), int(np.argmax(w.sizes)), 1, 100, 257, 511})
    for f in pick:
        b = w.bodies[w.offsets[f]:w.offsets[f + 1]]
        ref = oracle_iq(w.q, b[:, :3], b[:, 3].view(np.uint32), w.table)
        err = rel_err(frames[f], ref).max()
        _say(f"  frame {f} ({w.sizes[f]} atoms): max rel err {err:.3e}")
        assert err < 1e-5, f
    e.set_tuning(64, 0)
    e.set_persistent(True)
    e.batch_set_frames(w.bodies, w.offsets)
    frames3, _ = e.batch_calc()
    assert e.batch_launch()["persistent_waves"] > 0
    err = rel_err(frames3, frames).max()
    _say(f"  K=64 work-queue launch vs default: max rel diff {err:.3e}")
    # two FP32 kernels, each within a few 1e-6 of FP64 at these frames' deepest minima: they
    # agree to the accuracy contract, not to each other's rounding
    assert err < 1e-5


def _frames(seed, n_frames, nmin, nmax):
    from resaxs_amd import configs
    w = configs.ensemble_512(n_frames=n_frames, nmin=nmin, nmax=nmax, seed=seed)
    return w


def test_batch_independent_of_single_structure(oracle):
    """The batch keeps a layout of its own: single-structure calls on the same engine (a
    hydrated calc_curve, which rebuilds the engine's class table, and an incremental
    update) between batch launches leave the batch's results unchanged bit for bit, whether
    they come before or after its first launch."""
    from resaxs_amd import configs
    w = _frames(5, 8, 200, 1500)
    s = configs.sasa_ellipsoid_20000(n=3000)
    _say("case: batch interleaved with a hydrated single structure")
    e = engine_for(w.q, w.table, w.water_index, 0.0)
    e.batch_set_frames(w.bodies, w.offsets)
    ref_frames, ref_mean = e.batch_calc()
    # hydrated single structure on the same engine (same q and table, SASA on)
    e.set_sasa_params(s.radii, s.probe, n_half=s.n_half)
    e.set_water_weight(s.w)
    e.initialize(len(s.pos))
    I1 = e.calc_curve(s.bodies)
    frames, mean = e.batch_calc()   # hydration is now on: the frames get exposed-dot classes
    hyd = [oracle.sasa_counts(w.bodies[w.offsets[f]:w.offsets[f + 1], :3],
                              w.bodies[w.offsets[f]:w.offsets[f + 1], 3].view(np.uint32),
                              (s.radii + np.float32(s.probe)) ** 2,
                              np.stack([oracle.fibonacci_dots(s.n_half, r, s.probe) for r in s.radii]))
           for f in range(len(w.offsets) - 1)]
    for f in (0, 3, 7):
        b = w.bodies[w.offsets[f]:w.offsets[f + 1]]
        ref = oracle_iq(w.q, b[:, :3], b[:, 3].view(np.uint32), w.table, w.water_index, s.w, hyd[f], 2 * s.n_half)
        assert rel_err(frames[f], ref).max() < 1e-5, f
    # back to plain: the hydrated layout serves w = 0 with the plain factors
    e.set_water_weight(0.0)
    frames0, mean0 = e.batch_calc()
    assert rel_err(frames0, ref_frames).max() < 1e-6
    # a new engine: single-structure work before the first batch launch, and between launches
    e2 = engine_for(w.q, w.table, w.water_index, 0.0)
    e2.batch_set_frames(w.bodies, w.offsets)
    e2.set_sasa_params(s.radii, s.probe, n_half=s.n_half)
    e2.initialize(len(s.pos))
    e2.set_water_weight(s.w)
    I2 = e2.calc_curve(s.bodies)
    e2.set_water_weight(0.0)
    f2, m2 = e2.batch_calc()
    assert np.array_equal(f2, ref_frames) and np.array_equal(m2, ref_mean)
    e2.set_water_weight(s.w)
    e2.calc_curve(s.bodies)
    moved = s.bodies.copy()
    moved[:50, :3] += np.float32(0.3)
    e2.recalc_curve(moved, 0, 50)
    e2.set_water_weight(0.0)
    f3, m3 = e2.batch_calc()
    assert np.array_equal(f3, ref_frames) and np.array_equal(m3, ref_mean)
    assert np.array_equal(I1, I2)


def test_batch_factor_swap_keeps_the_launch(oracle):
    """A fit over (c, w) on resident frames: the default atomic factors regenerated at each
    expansion factor, the water weight changed between launches.  Every launch matches the
    FP64 oracle for its own (c, w), and the captured launch (jobs, shape) is kept."""
    from resaxs_amd import configs
    rng = np.random.default_rng(12)
    q = np.linspace(0.01, 0.45, 70).astype(np.float32)
    n, nf = 900, 3
    base = configs.in_ball(rng, n, (n / 0.045 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0))
    types = rng.choice(np.array([1, 2, 3, 5, 8, 9, 13], dtype=np.uint32), size=n).astype(np.uint32)
    models = [(base + rng.normal(scale=0.4, size=base.shape)).astype(np.float32) for _ in range(nf)]
    bodies = np.concatenate([oracle.bodies(p, types) for p in models])
    offsets = np.arange(nf + 1, dtype=np.int32) * n
    radii = oracle.atomic_radii()
    avg_r = np.float32(np.mean(radii[types].astype(np.float64)))
    dots = np.stack([oracle.dodeca_icosa_dots(r, 1.8) for r in radii])
    counts = [oracle.sasa_counts(p, types, ((radii + np.float32(1.8)) ** 2).astype(np.float32), dots) for p in models]
    _say("case: batch factor swaps over (c, w), 3 frames x 900 atoms x 70 q")
    from resaxs_amd import Engine, RSX_DOTS_DODECA_ICOSA
    e = Engine(0)
    e.set_q(q)
    e.set_sasa_params(radii, 1.8, RSX_DOTS_DODECA_ICOSA, 0)
    e.set_atomic_factors(avg_r)
    e.set_water_weight(1.0)
    e.batch_set_frames(bodies, offsets)
    shape = None
    for c, ww in ((1.0, 1.0), (0.97, 1.0), (1.03, -1.5), (1.03, 0.0), (0.95, 3.2)):
        e.set_expansion_factor(c)
        e.set_water_weight(ww)
        frames, mean = e.batch_calc()
        launch = e.batch_launch()
        shape = shape or launch
        assert launch == shape
        table = oracle.atomic_ff(q, avg_r, np.float32(c))
        for f in range(nf):
            ref = oracle_iq(q, models[f], types, table, 14, ww, counts[f], 32)
            err = rel_err(frames[f], ref).max()
            assert err < 1e-5, (c, ww, f, err)
    _say("  every (c, w) point within 1e-5 of the oracle, one captured launch")


def test_batch_autotune_and_cache(oracle, tmp_path):
    """rsx_batch_autotune times the FP32 kernels over K, column range and launch form on the
    frames, keeps the fastest and appends it to the cache; a second engine takes the cached
    entry and reproduces the first one's results bit for bit; frames still meet the bar."""
    w = _frames(7, 24, 100, 2500)
    _say("case: batch autotune + cache, 24 frames")
    cache = str(tmp_path / "tune.txt")
    e = engine_for(w.q, w.table, w.water_index, 0.0)
    e.batch_set_frames(w.bodies, w.offsets)
    t = e.batch_autotune(cache)
    _say(f"  tuned: {t}")
    tok = open(cache).read().split()
    assert len(tok) == 6 and ":batch:nq128:" in tok[0]
    assert (int(tok[1]), int(tok[2]), int(tok[3])) == (t["k"], t["col_range"], t["variant"])
    if int(tok[4]) == 0:
        assert t["persistent_waves"] == 0
    frames, mean = e.batch_calc()
    for f in (0, 11, 23):
        b = w.bodies[w.offsets[f]:w.offsets[f + 1]]
        assert rel_err(frames[f], oracle_iq(w.q, b[:, :3], b[:, 3].view(np.uint32), w.table)).max() < 1e-5, f
    g = engine_for(w.q, w.table, w.water_index, 0.0)
    g.batch_set_frames(w.bodies, w.offsets)
    assert g.batch_autotune(cache) == t
    f2, m2 = g.batch_calc()
    assert np.array_equal(f2, frames) and np.array_equal(m2, mean)
    assert len(open(cache).read().split()) == 6
    e.close()
    g.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_ragged_and_empty_frames(oracle, persistent):
    """Frames of 0, 1, 2, 3, 65 and 1000 atoms, empty frames first, between and last: every
    frame's I(q) follows the FP64 oracle (an empty frame gives 0 at every q, one atom F^2),
    and the mean is the uniform average over all frames, empty ones included."""
    from resaxs_amd import configs
    rng = np.random.default_rng(21)
    q = np.linspace(0.0, 0.5, 48).astype(np.float32)
    table = configs.factor_table(rng, q)
    sizes = [0, 1, 2, 0, 3, 65, 1000, 0]
    pos = [configs.in_ball(rng, m, 12.0).astype(np.float32) if m else np.zeros((0, 3), np.float32) for m in sizes]
    typ = [configs.seeded_types(rng, m) if m else np.zeros(0, np.uint32) for m in sizes]
    bodies = np.concatenate([configs.bodies_of(p, t) if len(p) else np.zeros((0, 4), np.float32)
                             for p, t in zip(pos, typ)]).astype(np.float32)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    _say(f"case: ragged batch {sizes} x {len(q)} q, persistent={persistent}")
    e = engine_for(q, table, configs.WATER, 0.0)
    e.set_persistent(persistent)
    e.batch_set_frames(bodies, offsets)
    frames, mean = e.batch_calc()
    assert frames.shape == (len(sizes), len(q))
    for f, m in enumerate(sizes):
        if m == 0:
            assert np.all(frames[f] == 0.0), f
            continue
        ref = oracle_iq(q, pos[f], typ[f], table)
        err = rel_err(frames[f], ref).max()
        assert err < 1e-5, (f, m, err)
    assert np.allclose(mean, frames.mean(axis=0), rtol=1e-13, atol=0)
    e.close()


def test_zero_atoms_single_structure():
    """A structure of no atoms: calc_curve gives 0 at every q, plain and hydrated, and an
    empty incremental update leaves it so."""
    from resaxs_amd import configs
    rng = np.random.default_rng(22)
    q = np.linspace(0.0, 0.5, 16).astype(np.float32)
    table = configs.factor_table(rng, q)
    empty = np.zeros((0, 4), np.float32)
    for hydrate in (False, True):
        _say(f"case: zero atoms, hydrate={hydrate}")
        if hydrate:
            e = engine_for(q, table, configs.WATER, 0.5, n=0, radii=configs.RADII, probe=configs.PROBE, n_half=32)
        else:
            e = engine_for(q, table, configs.WATER, 0.0, n=0)
        assert np.all(e.calc_curve(empty) == 0.0)
        assert np.all(e.recalc_curve(empty, 0, 0) == 0.0)
        e.close()


def test_batch_rejects_bad_offsets():
    """Offsets are checked before any use: offsets[0] must be 0 and they must not decrease."""
    from resaxs_amd import RsxError
    from resaxs_amd._lib import RSX_INVALID_ARG
    w = _frames(6, 3, 100, 300)
    e = engine_for(w.q, w.table, w.water_index, 0.0)
    for bad in (np.array([1, 100, 200, 300], np.int32), np.array([0, 200, 100, 300], np.int32),
                np.array([-5, 100, 200, 300], np.int32)):
        with pytest.raises(RsxError) as ei:
            e.batch_set_frames(w.bodies, bad)
        assert ei.value.code == RSX_INVALID_ARG
    e.batch_set_frames(w.bodies, w.offsets)
    assert e.batch_calc()[0].shape == (3, len(w.q))


def _fit_case(seed=41, n=1200, nq=80):
    from resaxs_amd import configs
    rng = np.random.default_rng(seed)
    q = np.linspace(0.01, 0.5, nq).astype(np.float32)
    pos = configs.in_ellipsoid(rng, n, configs.ellipsoid_axes(n)).astype(np.float32)
    types = configs.seeded_types(rng, n)
    table = configs.factor_table(rng, q)
    return rng, q, pos, types, table


@pytest.mark.parametrize("form", ["pair", "per_lane", "fp64"])
def test_adjoint_vs_analytic_and_fd(oracle, form):
    from resaxs_amd import configs
    _say(f"case: chi2 gradient, 1200 atoms x 80 q, {form} adjoint")
    rng, q, pos, types, table = _fit_case()
    F = oracle.body_factors(pos, types, q, table)
    Iref = oracle.debye(q, pos, F)
    i_exp = (Iref * (1.0 + 0.01 * rng.normal(size=len(q)))).astype(np.float32)
    sigma = (0.01 * Iref).astype(np.float32)
    start = (pos + rng.normal(scale=0.5, size=pos.shape)).astype(np.float32)
    e = engine_for(q, table, n=len(pos))
    e.set_adjoint_tuning(form != "fp64", per_lane=form == "per_lane")
    chi2, scale, grad, iq = e.chi2_grad(configs.bodies_of(start, types), i_exp, sigma)
    assert e.adjoint_layout()["form"] == form
    xyz = start.astype(np.float64)
    Fs = oracle.body_factors(start, types, q, table)
    g_ref, I_ref, c_ref, x2_ref = oracle.chi2_grad(q, xyz, Fs, i_exp, sigma)
    assert rel_err(iq, I_ref).max() < 1e-5
    # the engine's scale and chi^2 are the oracle's formulas applied to its own I
    c_own, x2_own = oracle.scale_chi2(i_exp, sigma, iq)
    assert abs(scale - c_own) <= 1e-12 * abs(c_own) and abs(chi2 - x2_own) <= 1e-12 * x2_own
    # chi^2 vs the FP64 chi^2: first-order propagation of the 1e-5 bar on I
    r = (i_exp - c_ref * I_ref) / sigma
    dr = c_ref * 1e-5 * np.abs(I_ref) / sigma
    assert abs(chi2 - x2_ref) <= 2 * np.sum(np.abs(r) * dr) + np.sum(dr * dr)
    gnorm = np.linalg.norm(g_ref)
    d = np.linalg.norm(grad.astype(np.float64) - g_ref) / gnorm
    _say(f"  |g - g_ref| / |g_ref| = {d:.3e}")
    assert d < 1e-4
    comps = rng.choice(3 * len(pos), size=48, replace=False)
    fd = oracle.chi2_fd(q, xyz, Fs, i_exp, sigma, comps, h=1e-3)
    gg = grad.reshape(-1)[comps].astype(np.float64)
    rel = np.linalg.norm(gg - fd) / np.linalg.norm(fd)
    _say(f"  vs central differences: rel {rel:.3e}")
    assert rel < 1e-4
    # translation invariance: the gradient sums to zero
    assert np.linalg.norm(grad.astype(np.float64).sum(0)) < 1e-4 * gnorm * np.sqrt(len(pos))


def test_vjp_two_atoms_closed_form():
    """dI_k/dx for two atoms: 2 F1 F2 [cos(qr) - sinc(qr)] (x1 - x2) / r^2."""
    _say("case: vjp closed form, 2 atoms")
    q = np.linspace(0.05, 0.6, 32).astype(np.float32)
    table = np.stack([np.linspace(3, 1, 32), np.linspace(5, 2, 32)], 1).astype(np.float32)
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 3.0, -1.5]], dtype=np.float32)
    from oracle import pyoracle as po
    e = engine_for(q, table, n=2)
    wv = np.linspace(-1, 2, 32)
    g = e.vjp(po.bodies(pos, [0, 1]), wv)
    d = pos[0].astype(np.float64) - pos[1]
    r = np.linalg.norm(d)
    x = q.astype(np.float64) * r
    t = np.cos(x) - np.sin(x) / x
    coef = np.sum(wv * 2 * table[:, 0] * table[:, 1] * t) / r ** 2
    assert np.allclose(g[0], coef * d, rtol=1e-4, atol=1e-6 * abs(coef) * r)
    assert np.allclose(g[1], -coef * d, rtol=1e-4, atol=1e-6 * abs(coef) * r)


@pytest.mark.parametrize("hydrate", [False, True])
def test_batch_vjp_matches_single_structure(oracle, hydrate):
    """rsx_batch_vjp over ragged frames (700, 0, 2, 1500, 65 and 1 atoms) with weights of their
    own per frame: every frame's gradient equals the single-structure VJP of that frame to
    FP32 rounding (both FP32 adjoints; the weights enter rounded differently), sums to zero
    over the frame (translation invariance), and the two-atom frame follows the closed form
    2 F1 F2 [cos(qr) - sinc(qr)] (x1 - x2) / r^2."""
    from resaxs_amd import configs
    rng = np.random.default_rng(23)
    q = np.linspace(0.01, 0.5, 64).astype(np.float32)
    table = configs.factor_table(rng, q)
    sizes = [700, 0, 2, 1500, 65, 1]
    pos = [configs.in_ellipsoid(rng, m, configs.ellipsoid_axes(m)).astype(np.float32) if m > 2
           else configs.in_ball(rng, m, 3.0).astype(np.float32) for m in sizes]
    typ = [configs.seeded_types(rng, m) for m in sizes]
    bodies = np.concatenate([configs.bodies_of(p, t) for p, t in zip(pos, typ)]).astype(np.float32)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    w = rng.normal(size=(len(sizes), len(q)))
    kw = dict(radii=configs.RADII, probe=configs.PROBE, n_half=32) if hydrate else {}
    ww = 0.6 if hydrate else 0.0
    _say(f"case: batch VJP vs single-structure VJP, frames {sizes} x {len(q)} q, hydrate={hydrate}")
    e = engine_for(q, table, configs.WATER, ww, **kw)
    e.batch_set_frames(bodies, offsets)
    g = e.batch_vjp(w)
    assert g.shape == (int(offsets[-1]), 3)
    lay = e.adjoint_layout(batch=True)
    assert lay["form"] == ("per_lane" if hydrate else "pair")
    assert e.time_batch_vjp(2) > 0.0
    if not hydrate:
        # the per-lane form at the same K: the same products and sums
        e.set_adjoint_tuning(True, lay["k"], 0, per_lane=True)
        g_lane = e.batch_vjp(w)
        assert e.adjoint_layout(batch=True)["form"] == "per_lane"
        d = np.linalg.norm(g.astype(np.float64) - g_lane) / np.linalg.norm(g_lane.astype(np.float64))
        _say(f"  class-pair vs per-lane form, K={lay['k']}: {d:.2e}")
        assert d < 1e-6
    e.close()
    for f, m in enumerate(sizes):
        if m == 0:
            continue
        gf = g[offsets[f]:offsets[f + 1]].astype(np.float64)
        s = engine_for(q, table, configs.WATER, ww, n=m, **kw)
        ref = s.vjp(configs.bodies_of(pos[f], typ[f]), w[f]).astype(np.float64)
        s.close()
        scale = max(np.linalg.norm(ref), 1e-30)
        d = np.linalg.norm(gf - ref) / scale
        _say(f"  frame {f} ({m} atoms): |g - g_single| / |g_single| = {d:.2e}")
        assert d < 1e-5 or (m == 1 and np.all(gf == 0.0)), (f, d)
        assert np.linalg.norm(gf.sum(0)) <= 1e-5 * scale * np.sqrt(m) + 1e-30, f
    if not hydrate:
        p2, t2 = pos[2], typ[2]
        d = p2[0].astype(np.float64) - p2[1]
        r = np.linalg.norm(d)
        x = q.astype(np.float64) * r
        wk = np.where(q > 0, w[2], 0.0)
        coef = np.sum(wk * 2 * table[:, t2[0]] * table[:, t2[1]] * (np.cos(x) - np.sin(x) / x)) / r ** 2
        g2 = g[offsets[2]:offsets[3]].astype(np.float64)
        assert np.allclose(g2[0], coef * d, rtol=1e-4, atol=1e-6 * abs(coef) * r)
        assert np.allclose(g2[1], -coef * d, rtol=1e-4, atol=1e-6 * abs(coef) * r)


@pytest.mark.parametrize("n_types", [5, 40])
def test_adjoint_class_pair_form(oracle, n_types):
    """The FP32 adjoint's class-pair coefficient form (plain layouts; coefficients through
    scalar loads) against its per-lane form at the same K -- the same products, recurrences
    and sums, so the same gradient up to the rounding that contractions may move -- and at the
    automatic K = 128 (one chunk at 128 q) against the oracle's FP64 chi^2 gradient.  With 40
    types of ~12 atoms every 64-row block holds many classes and runs once per class."""
    from resaxs_amd import configs
    rng = np.random.default_rng(90 + n_types)
    q = np.linspace(0.01, 0.5, 128).astype(np.float32)
    if n_types == 5:
        n = 1500
        types = configs.seeded_types(rng, n)
        table = configs.factor_table(rng, q)
    else:
        n = 480
        types = rng.integers(0, n_types, size=n).astype(np.uint32)
        a = rng.uniform(2.0, 10.0, size=n_types)
        b = rng.uniform(0.5, 5.0, size=n_types)
        table = (a[None, :] * np.exp(-b[None, :] * q.astype(np.float64)[:, None] ** 2)).astype(np.float32)
    pos = configs.in_ellipsoid(rng, n, configs.ellipsoid_axes(n)).astype(np.float32)
    Iref = oracle.debye(q, pos, oracle.body_factors(pos, types, q, table))
    i_exp = (Iref * (1.0 + 0.01 * rng.normal(size=len(q)))).astype(np.float32)
    sigma = (0.01 * Iref).astype(np.float32)
    start = (pos + rng.normal(scale=0.5, size=pos.shape)).astype(np.float32)
    bodies = configs.bodies_of(start, types)
    _say(f"case: class-pair adjoint form, {n} atoms of {n_types} types x {len(q)} q")
    e = engine_for(q, table, n=n)
    got = {}
    for k, per_lane in ((64, False), (64, True), (100, False), (100, True), (0, False)):
        e.set_adjoint_tuning(True, k, 0, per_lane)
        _, _, g, _ = e.chi2_grad(bodies, i_exp, sigma)
        lay = e.adjoint_layout()
        assert lay["form"] == ("per_lane" if per_lane else "pair")
        got[(k, per_lane)] = (g.astype(np.float64), lay["k"])
    e.close()
    for k in (64, 100):
        gp, gl = got[(k, False)][0], got[(k, True)][0]
        d = np.linalg.norm(gp - gl) / np.linalg.norm(gl)
        _say(f"  K={k}: |g_pair - g_per_lane| / |g| = {d:.2e} (bit-equal: {bool(np.array_equal(gp, gl))})")
        assert d < 1e-6
    g_auto, k_auto = got[(0, False)]
    assert k_auto == 128
    Fs = oracle.body_factors(start, types, q, table)
    g_ref = oracle.chi2_grad(q, start.astype(np.float64), Fs, i_exp, sigma)[0]
    d = np.linalg.norm(g_auto - g_ref) / np.linalg.norm(g_ref)
    _say(f"  K=128 vs the oracle: {d:.2e}")
    assert d < 1e-4
