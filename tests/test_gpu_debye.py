"""GPU parity of the Debye pair sums (forward I(q)) against the FP64 oracle.

Bar: I(q) within 1e-5 relative of the FP64 oracle at every q (the stated
tolerance); the reference's own golden fixture (test_resaxs/test1888.*) is
matched to the fixture's stated accuracy (5e-5; it carries 1e-3 A position
rounding).
"""
import os
import sys

import numpy as np
import pytest

from helpers import engine_for, fib_dot_table, oracle_iq, r2_table, rel_err
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-5
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test1888.npz")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_golden_test1888(oracle):
    g = np.load(GOLDEN)
    _say("case: test1888 golden, 1888 bodies x 51 q, 21-type table")
    e = engine_for(g["q"], g["table"], n=len(g["pos"]))
    b = oracle.bodies(g["pos"], g["types"])
    I = e.calc_curve(b)
    ref = oracle_iq(g["q"], g["pos"], g["types"], g["table"])
    _say(f"  max rel err vs FP64 {rel_err(I, ref).max():.3e}, vs fixture {rel_err(I, g['Iq']).max():.3e}")
    assert rel_err(I, ref).max() < TOL
    assert rel_err(I, g["Iq"]).max() < 5e-5
    # q = 0: (sum F)^2 exactly (FP64 rounding only)
    F0 = g["table"][0][g["types"]].astype(np.float64).sum()
    assert abs(I[0] - F0 * F0) <= 1e-12 * F0 * F0


@pytest.mark.skipif(po.ref_lib() is None, reason="reference CPU subset (oracle/_ref) not built")
def test_golden_vs_reference_host_debye():
    """GPU I(q) against the reference's own CPU path (host_debye.cpp:54-109 compiled from
    its sources; FP64 sums rounded to FP32) on the golden inputs and a random cluster."""
    from resaxs_amd import configs
    g = np.load(GOLDEN)
    _say("case: reference host_debye, test1888 + 2500-atom cluster")
    e = engine_for(g["q"], g["table"], n=len(g["pos"]))
    I = e.calc_curve(po.bodies(g["pos"], g["types"]))
    assert rel_err(I, po.ref_host_debye(g["q"], g["pos"], g["types"], g["table"])).max() < TOL
    rng = np.random.default_rng(17)
    pos = configs.in_ball(rng, 2500, 22.0).astype(np.float32)
    types = configs.seeded_types(rng, 2500)
    q = np.linspace(0.0, 0.5, 101).astype(np.float32)
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=2500)
    I = e.calc_curve(configs.bodies_of(pos, types))
    assert rel_err(I, po.ref_host_debye(q, pos, types, table)).max() < TOL


def test_config1_plain(oracle):
    from resaxs_amd import configs
    w = configs.plain_ball_1888()
    _say("case: plain_ball_1888 1888 x 101 plain")
    e = engine_for(w.q, w.table, w.water_index, 0.0, n=len(w.pos))
    I = e.calc_curve(w.bodies)
    ref = oracle_iq(w.q, w.pos, w.types, w.table)
    err = rel_err(I, ref)
    _say(f"  max rel err {err.max():.3e}")
    assert err.max() < TOL


KS_FP64 = [8, 16, 24, 32, 40, 48, 50, 56, 64, 67]
KS_FP32 = [16, 32, 43, 48, 51, 64, 67, 80, 86, 96, 100]
KS_FP32R = [64, 86, 96, 100, 128]


@pytest.mark.parametrize("variant,k", [(0, k) for k in KS_FP64] + [(1, 32), (1, 64)] + [(2, k) for k in KS_FP32]
                         + [(3, k) for k in KS_FP32R])
def test_k_variants_agree(oracle, variant, k):
    """Every kernel variant and q-per-thread instantiation gives the oracle's I(q)."""
    from resaxs_amd import configs
    _say(f"case: variant {variant} K={k}, 1500 atoms x 120 q")
    rng = np.random.default_rng(11)
    n = 1500
    pos = configs.in_ball(rng, n, 16.0).astype(np.float32)
    types = configs.seeded_types(rng, n)
    q = np.linspace(0.0, 0.7, 120).astype(np.float32)
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=n)
    e.set_variant(variant)
    e.set_tuning(k, 128)
    I = e.calc_curve(configs.bodies_of(pos, types))
    ref = oracle_iq(q, pos, types, table)
    assert rel_err(I, ref).max() < TOL


def test_nonuniform_grid_direct_path(oracle):
    """A q grid that is not uniform takes the direct (sin per term) path."""
    from resaxs_amd import configs
    _say("case: non-uniform q, 1200 atoms x 60 q")
    rng = np.random.default_rng(12)
    n = 1200
    pos = configs.in_ball(rng, n, 15.0).astype(np.float32)
    types = configs.seeded_types(rng, n)
    q = np.sort(rng.uniform(0.0, 0.6, 60)).astype(np.float32)
    q[0] = 0.0
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=n)
    I = e.calc_curve(configs.bodies_of(pos, types))
    ref = oracle_iq(q, pos, types, table)
    assert rel_err(I, ref).max() < TOL


def test_nonuniform_grid_direct_path_full_size(oracle):
    """20,000 atoms x 256 q on the non-uniform grid q_k = 0.6 (k / 255)^1.5 (the bench's
    direct workload: one FP32 sin per term) against the full FP64 oracle."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000()
    q = (0.6 * (np.arange(256) / 255.0) ** 1.5).astype(np.float32)
    table = configs.factor_table(np.random.default_rng(77), q)
    _say("case: non-uniform q, sasa_ellipsoid_20000 plain 20000 x 256")
    e = engine_for(q, table, n=len(w.pos))
    I = e.calc_curve(w.bodies)
    assert not e.get_tuning()["recurrence"]
    ref = oracle_iq(q, w.pos, w.types, table)
    err = rel_err(I, ref)
    k = int(np.argmax(err))
    _say(f"  max rel err {err.max():.3e} at q={q[k]:.4f}, median {np.median(err):.3e}")
    assert err.max() < TOL


@pytest.mark.parametrize("radius", [60.0, 400.0])
def test_nonuniform_grid_phase_forms(oracle, radius):
    """The direct kernel forms phases q r / (2 pi) in FP32 while every one stays within 64
    revolutions (a 60 A ball at q <= 0.6: 11.5) and in FP64 beyond (a sparse 400 A ball:
    76 revolutions at the largest pairs); both against the FP64 oracle."""
    from resaxs_amd import configs
    rng = np.random.default_rng(31)
    n = 2500
    pos = configs.in_ball(rng, n, radius).astype(np.float32)
    types = configs.seeded_types(rng, n)
    q = np.sort(rng.uniform(0.0, 0.6, 96)).astype(np.float32)
    q[-1] = 0.6
    table = configs.factor_table(rng, q)
    _say(f"case: non-uniform q, {n} atoms in a {radius:.0f} A ball x {len(q)} q")
    e = engine_for(q, table, n=n)
    I = e.calc_curve(configs.bodies_of(pos, types))
    assert not e.get_tuning()["recurrence"]
    err = rel_err(I, oracle_iq(q, pos, types, table))
    _say(f"  max rel err {err.max():.3e}, median {np.median(err):.3e}")
    assert err.max() < TOL


def test_coincident_atoms_and_tiny_n(oracle):
    """r = 0 pairs give sinc = 1 (gpu_saxs_pt_wf.cl:55); N = 1, 2, 65 edge sizes."""
    _say("case: coincident atoms, N in {1, 2, 3, 65}")
    q = np.linspace(0.0, 0.5, 40).astype(np.float32)
    rng = np.random.default_rng(13)
    table = rng.uniform(1, 10, size=(40, 3)).astype(np.float32)
    for n in (1, 2, 3, 65):
        pos = rng.uniform(-5, 5, size=(n, 3)).astype(np.float32)
        if n >= 3:
            pos[2] = pos[0]
        types = rng.integers(0, 3, size=n).astype(np.uint32)
        e = engine_for(q, table, n=n)
        I = e.calc_curve(oracle.bodies(pos, types))
        ref = oracle_iq(q, pos, types, table)
        assert rel_err(I, ref).max() < TOL, n
    # closed forms: one atom -> F^2; two atoms -> F1^2 + F2^2 + 2 F1 F2 sinc(q r)
    pos = np.array([[0, 0, 0], [3.5, 0, 0]], dtype=np.float32)
    types = np.array([0, 1], dtype=np.uint32)
    e = engine_for(q, table, n=2)
    I = e.calc_curve(oracle.bodies(pos, types))
    f1, f2 = table[:, 0].astype(np.float64), table[:, 1].astype(np.float64)
    x = q.astype(np.float64) * 3.5
    sinc = np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1, x))
    assert rel_err(I, f1 * f1 + f2 * f2 + 2 * f1 * f2 * sinc).max() < TOL


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_coincident_atoms_off_diagonal(oracle, variant):
    """Coincident atoms far apart in index (off the diagonal tiles) in both pair-kernel
    variants: the zero-distance pair contributes F_i F_j at every q."""
    from resaxs_amd import configs
    _say(f"case: coincident pairs off the diagonal, 3000 atoms x 256 q, variant {variant}")
    rng = np.random.default_rng(19)
    n = 3000
    pos = configs.in_ball(rng, n, 25.0).astype(np.float32)
    for a, b in ((10, 2500), (900, 1700), (64, 65), (1234, 2999), (3, 2047)):
        pos[b] = pos[a]
    types = configs.seeded_types(rng, n)
    q = np.linspace(0.0, 0.6, 256).astype(np.float32)
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=n)
    e.set_variant(variant)
    I = e.calc_curve(configs.bodies_of(pos, types))
    ref = oracle_iq(q, pos, types, table)
    assert rel_err(I, ref).max() < TOL


@pytest.mark.parametrize("variant,k", [(2, 86), (3, 86), (3, 128)])
def test_fp32_variants_hydrated(oracle, variant, k):
    """Hydration classes (type x exposed-dot count) give many class runs per column
    range: every run flush of both FP32 kernels against the FP64 oracle."""
    from resaxs_amd import configs
    n = 4000
    w = configs.sasa_ellipsoid_20000(n=n)
    _say(f"case: hydrated 4000 x 256, variant {variant} K={k}")
    e = engine_for(w.q, w.table, w.water_index, w.w, n=n, radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.set_variant(variant)
    e.set_tuning(k, 512)
    I = e.calc_curve(w.bodies)
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    counts = oracle.sasa_counts(w.pos, w.types, r2_table(w.radii, w.probe), dots)
    ref = oracle_iq(w.q, w.pos, w.types, w.table, w.water_index, w.w, counts, 256)
    err = rel_err(I, ref)
    _say(f"  max rel err {err.max():.3e}")
    assert err.max() < TOL


def test_invariances(oracle):
    """Rigid motion and atom permutation leave I(q) unchanged; F -> a F scales I by a^2."""
    from resaxs_amd import configs
    _say("case: invariances, 3000 atoms x 64 q")
    rng = np.random.default_rng(14)
    n = 3000
    pos = configs.in_ball(rng, n, 20.0)
    types = configs.seeded_types(rng, n)
    q = np.linspace(0.0, 0.5, 64).astype(np.float32)
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=n)
    I0 = e.calc_curve(configs.bodies_of(pos.astype(np.float32), types))
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    p2 = (pos @ R.T + np.array([3.0, -2.0, 1.0])).astype(np.float32)
    perm = rng.permutation(n)
    I1 = e.calc_curve(configs.bodies_of(p2[perm], types[perm]))
    assert rel_err(I1, I0).max() < 2 * TOL
    e2 = engine_for(q, (2.0 * table).astype(np.float32), n=n)
    I2 = e2.calc_curve(configs.bodies_of(pos.astype(np.float32), types))
    assert rel_err(I2, 4.0 * I0).max() < 1e-6


def test_config2_plain_full_size(oracle):
    """20,000 atoms x 256 q without hydration against the full FP64 oracle."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000()
    _say("case: sasa_ellipsoid_20000 plain 20000 x 256")
    e = engine_for(w.q, w.table, w.water_index, 0.0, n=len(w.pos))
    I = e.calc_curve(w.bodies)
    ref = oracle_iq(w.q, w.pos, w.types, w.table)
    err = rel_err(I, ref)
    k = int(np.argmax(err))
    _say(f"  max rel err {err.max():.3e} at q={w.q[k]:.4f}, median {np.median(err):.3e}")
    assert err.max() < TOL


def _hydrated_dip_case():
    """3000 atoms, 128 q over [0, 0.5], SASA hydration with uniform radii: I(q) dips
    30x below its maximum near q index 20, where the pair terms cancel deeply."""
    from resaxs_amd import configs
    rng = np.random.default_rng(5)
    n = 3000
    q = np.linspace(0.0, 0.5, 128).astype(np.float32)
    pos = configs.in_ellipsoid(rng, n, configs.ellipsoid_axes(n)).astype(np.float32)
    types = configs.seeded_types(rng, n)
    table = configs.factor_table(rng, q)
    water = table.shape[1] - 1            # the table's last column serves as the water factor
    radii = np.full(int(types.max()) + 1, 1.6, dtype=np.float32)
    return q, pos, types, table, water, radii, configs.bodies_of(pos, types)


@pytest.mark.parametrize("variant,k,col_range", [(3, 128, 256), (3, 128, 1024), (2, 64, 256), (2, 64, 1024)])
def test_deep_minimum_accuracy(oracle, variant, k, col_range):
    """The FP32 kernels at the bottom of a deep dip of I(q): the 64 rows' run sums cancel
    there, so they are summed in FP64 (an FP32 row sum of the run-flush kernel reached
    2.3e-5 at L = 1024 on this case)."""
    _say(f"case: deep minimum, hydrated 3000 x 128, variant {variant} K={k} L={col_range}")
    q, pos, types, table, water, radii, b = _hydrated_dip_case()
    e = engine_for(q, table, water, 0.8, n=len(pos), radii=radii, probe=1.4, n_half=64)
    e.set_variant(variant)
    e.set_tuning(k, col_range)
    I = e.calc_curve(b)
    counts, _ = e.calc_rsasa(b, radii, 1.4, fib_dot_table(radii, 1.4, 64))
    e.close()
    ref = oracle_iq(q, pos, types, table, water, 0.8, counts=counts, n_dots=128)
    err = rel_err(I, ref)
    _say(f"  max rel err {err.max():.3e} at q index {int(np.argmax(err))} (I min at {int(np.argmin(ref))})")
    assert err.max() < 8e-6


def test_autotune_and_cache(oracle, tmp_path):
    """rsx_autotune: sweeps the FP32 kernels over K and column range, keeps the fastest,
    appends it to the cache file; a second engine with the same key takes the cached
    entry.  I(q) after tuning still meets the FP64 bar, and the time_sasa / time_forward
    probes report positive device times."""
    from resaxs_amd import configs
    _say("case: autotune + cache, 3000 atoms x 128 q, hydrated")
    q, pos, types, table, water, radii, b = _hydrated_dip_case()
    n = len(pos)
    cache = str(tmp_path / "tune.txt")
    e = engine_for(q, table, water, 0.8, n=n, radii=radii, probe=1.4, n_half=64)
    e.calc_curve(b)
    e.autotune(cache)
    t1 = e.get_tuning()
    assert t1["variant"] in (2, 3) and t1["k"] > 0 and t1["col_range"] > 0
    tok = open(cache).read().split()   # key K L variant ms
    assert len(tok) == 5 and int(tok[1]) == t1["k"] and int(tok[2]) == t1["col_range"] and int(tok[3]) == t1["variant"]
    I = e.calc_curve(b)
    assert e.time_forward(3)[1] > 0 and e.time_sasa(3) > 0
    counts, _ = e.calc_rsasa(b, radii, 1.4, fib_dot_table(radii, 1.4, 64))
    ref = oracle_iq(q, pos, types, table, water, 0.8, counts=counts, n_dots=128)
    err = rel_err(I, ref).max()
    _say(f"  tuned K={t1['k']} L={t1['col_range']} variant {t1['variant']}: max rel err {err:.3e}")
    assert err < TOL
    e.close()
    f = engine_for(q, table, water, 0.8, n=n, radii=radii, probe=1.4, n_half=64)
    f.calc_curve(b)
    f.autotune(cache)
    assert f.get_tuning() == t1
    assert np.array_equal(f.calc_curve(b), I)
    f.close()
    assert len(open(cache).read().split()) == 5


def _dev_forward(e, bodies):
    """I(q) from rsx_calc_curve_dev_async on a device copy of bodies (then synchronize)."""
    b = np.ascontiguousarray(bodies, dtype=np.float32)
    pb, po_ = e.dev_alloc(b.nbytes), e.dev_alloc(8 * e.nq)
    try:
        e.h2d(pb, b)
        e.calc_curve_dev_async(pb, len(b), po_)
        e.synchronize()
        return e.d2h(np.zeros(e.nq, dtype=np.float64), po_)
    finally:
        e.dev_free(pb)
        e.dev_free(po_)


@pytest.mark.parametrize("hydrate", [False, True])
def test_device_forward_after_retune(tmp_path, hydrate):
    """A device-resident forward after rsx_autotune has switched the column range, so that the
    forward job table changes size, equals the host result on the same positions: the device
    path reduces over the new table's slots.  Cache entries force 64, 1024 and again 64
    columns per job (more jobs, fewer, more)."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000(n=3000)
    wt = w.w if hydrate else 0.0
    _say(f"case: device-resident forward after retuning, 3000 x 256, hydrated={hydrate}")
    e = engine_for(w.q, w.table, w.water_index, wt, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.calc_curve(w.bodies)
    probe_cache = tmp_path / "probe.txt"
    e.autotune(str(probe_cache))
    key = probe_cache.read_text().split()[0]
    for L in (64, 1024, 64):
        forced = tmp_path / f"forced_{L}.txt"
        forced.write_text(f"{key} 128 {L} 3 1.0\n")
        e.autotune(str(forced))
        assert e.get_tuning()["col_range"] == L
        I_host = e.calc_curve(w.bodies)
        same = _dev_forward(e, w.bodies)
        d = float(np.max(np.abs(same - I_host) / np.abs(I_host)))
        _say(f"  {L} columns per job: device vs host {d:.2e}")
        assert d <= 1e-12, L
    e.close()


@pytest.mark.parametrize("hydrate", [False, True])
def test_device_resident_forward_on_new_positions(oracle, hydrate):
    """rsx_calc_curve_dev_async on positions other than the ones the host calc_curve laid
    out (every atom jittered by up to 0.3 A, types unchanged): the slot order and jobs are
    reused, the columns are gathered (and, hydrated, the SASA counts taken) on the device.
    I(q) meets the FP64 bar for the new positions."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000(n=3000)
    _say(f"case: device-resident forward on new positions, 3000 x 256, hydrated={hydrate}")
    wt = w.w if hydrate else 0.0
    e = engine_for(w.q, w.table, w.water_index, wt, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    I0 = e.calc_curve(w.bodies)
    same = _dev_forward(e, w.bodies)
    assert np.all(np.abs(same - I0) <= 1e-12 * np.abs(I0))
    rng = np.random.default_rng(31)
    centre = w.pos.astype(np.float64).mean(axis=0)
    # shrink a little first so the jitter cannot push an atom past the laid-out radius
    pos = (centre + 0.98 * (w.pos - centre) + rng.uniform(-0.3, 0.3, size=w.pos.shape)).astype(np.float32)
    got = _dev_forward(e, oracle.bodies(pos, w.types))
    counts = None
    if hydrate:
        counts = oracle.sasa_counts(pos, w.types, r2_table(w.radii, w.probe), fib_dot_table(w.radii, w.probe, w.n_half))
    ref = oracle_iq(w.q, pos, w.types, w.table, w.water_index, wt, counts, 2 * w.n_half)
    err = rel_err(got, ref).max()
    _say(f"  max rel err {err:.3e}")
    assert err < TOL
    # the engine's incremental state follows the host bodies again after device steps: an
    # update, a commit, another device step and a revert
    moved = w.bodies.copy()
    moved[:40, :3] += np.float32(0.25)
    I_inc = e.recalc_curve(moved, 0, 40)
    f = engine_for(w.q, w.table, w.water_index, wt, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    I_full = f.calc_curve(moved)
    f.close()
    assert rel_err(I_inc, I_full).max() < 1e-6
    e.commit()
    _dev_forward(e, oracle.bodies(pos, w.types))
    e.revert()
    assert np.array_equal(e.recalc_curve(moved, 0, 0), I_inc)
    moved[100:130, :3] -= np.float32(0.2)
    I_inc2 = e.recalc_curve(moved, 100, 30)
    g = engine_for(w.q, w.table, w.water_index, wt, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    assert rel_err(I_inc2, g.calc_curve(moved)).max() < 1e-6
    g.close()
    e.close()


def test_device_resident_forward_checks():
    """The device-resident forward checks its bodies on the device and reports through
    rsx_synchronize: a type outside the factor table (its pairs skipped), positions beyond the
    radius the layout was chosen for; parameters changed since the host calc_curve are refused
    up front.  A clean call after a reported failure succeeds."""
    from resaxs_amd import RsxError, configs
    w = configs.sasa_ellipsoid_20000(n=2000)
    _say("case: device-resident forward checks, 2000 x 256")
    e = engine_for(w.q, w.table, w.water_index, 0.0, n=len(w.pos))
    I0 = e.calc_curve(w.bodies)
    bad = w.bodies.copy()
    bad[5, 3] = np.array([w.table.shape[1] + 3], dtype=np.uint32).view(np.float32)[0]
    with pytest.raises(RsxError) as ei:
        _dev_forward(e, bad)
    assert ei.value.code == -3 and "type" in str(ei.value)
    far = w.bodies.copy()
    far[:, :3] *= np.float32(1.6)
    with pytest.raises(RsxError) as ei:
        _dev_forward(e, far)
    assert ei.value.code == -3 and "radius" in str(ei.value)
    again = _dev_forward(e, w.bodies)
    assert np.all(np.abs(again - I0) <= 1e-12 * np.abs(I0))
    e.set_factors(w.table * np.float32(1.01), w.water_index)
    with pytest.raises(RsxError) as ei:
        _dev_forward(e, w.bodies)
    assert ei.value.code == -5
    e.close()
