"""Incremental recalculation (recalc_curve) and the commit/revert transaction.

Semantics follow saxs_cl_gpu::recalc_curve / commit / revert
(resaxslib/src/saxs_alg_cl_gpu_pt_wf.cpp:468-620): bodies [start, start+len)
moved since the last calculation; commit checkpoints, revert restores the last
checkpoint (a no-op before any commit).
"""
import sys

import numpy as np
import pytest

from helpers import engine_for, fib_dot_table, oracle_iq, r2_table, rel_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_protein(seed, n):
    from resaxs_amd import configs
    q = np.linspace(0.0, 0.6, 128).astype(np.float32)
    rng, pos, types, table = configs._protein_like(seed, n, q)
    order = configs.morton_order(pos.astype(np.float64))
    return q, pos[order].copy(), types[order].copy(), table


def test_recalc_matches_full_recompute(oracle):
    from resaxs_amd import configs
    _say("case: 40 moves of 100 atoms on 4000 atoms, plain")
    q, pos, types, table = _small_protein(31, 4000)
    e = engine_for(q, table, n=len(pos))
    e.set_rebaseline(0)
    e.calc_curve(configs.bodies_of(pos, types))
    rng = np.random.default_rng(32)
    for m in range(40):
        start, newp = configs.random_move(rng, pos, 100)
        pos[start:start + 100] = newp
        I = e.recalc_curve(configs.bodies_of(pos, types), start, 100)
        e.commit()
    ref = oracle_iq(q, pos, types, table)
    err = rel_err(I, ref)
    _say(f"  running I after 40 moves: max rel err {err.max():.3e}")
    assert err.max() < 1e-5
    full = engine_for(q, table, n=len(pos)).calc_curve(configs.bodies_of(pos, types))
    assert rel_err(I, full).max() < 1e-5


def test_empty_and_full_moves(oracle):
    from resaxs_amd import configs
    _say("case: len=0 and len=N recalc")
    q, pos, types, table = _small_protein(33, 1500)
    e = engine_for(q, table, n=len(pos))
    b = configs.bodies_of(pos, types)
    I0 = e.calc_curve(b)
    assert np.array_equal(e.recalc_curve(b, 10, 0), I0)
    pos2 = (pos + 0.3).astype(np.float32)
    pos2[:700] += 1.0
    I2 = e.recalc_curve(configs.bodies_of(pos2, types), 0, len(pos))
    assert rel_err(I2, oracle_iq(q, pos2, types, table)).max() < 1e-5


def test_commit_revert_roundtrip():
    from resaxs_amd import configs
    _say("case: commit / revert round trip")
    q, pos, types, table = _small_protein(34, 2000)
    e = engine_for(q, table, n=len(pos))
    b0 = configs.bodies_of(pos, types)
    I0 = e.calc_curve(b0)
    e.revert()                       # no commit yet: no-op
    rng = np.random.default_rng(35)
    start, newp = configs.random_move(rng, pos, 200)
    p1 = pos.copy()
    p1[start:start + 200] = newp
    I1 = e.recalc_curve(configs.bodies_of(p1, types), start, 200)
    e.commit()
    start2, newp2 = configs.random_move(rng, p1, 300)
    p2 = p1.copy()
    p2[start2:start2 + 300] = newp2
    I2 = e.recalc_curve(configs.bodies_of(p2, types), start2, 300)
    assert not np.array_equal(I2, I1)
    e.revert()
    # after revert the state is the committed one: an empty recalc returns I1 bit for bit
    assert np.array_equal(e.recalc_curve(configs.bodies_of(p1, types), 0, 0), I1)
    # and moving back to p0 from the committed state reproduces I0 within tolerance
    I3 = e.recalc_curve(b0, start, 200)
    assert rel_err(I3, I0).max() < 1e-5


def test_state_errors():
    from resaxs_amd import Engine, RsxError
    _say("case: call-order errors")
    e = Engine(0)
    q = np.linspace(0, 0.5, 16).astype(np.float32)
    e.set_q(q)
    e.set_factors(np.ones((16, 2), np.float32))
    b = np.zeros((4, 4), np.float32)
    with pytest.raises(RsxError) as ex:
        e.calc_curve(b)
    assert ex.value.code == -5          # SAXS_ARGS_NOT_SET
    e.initialize(4)
    with pytest.raises(RsxError) as ex:
        e.recalc_curve(b, 0, 1)
    assert ex.value.code == -6          # SAXS_NO_INITIAL_CALC
    with pytest.raises(RsxError) as ex:
        e.commit()
    assert ex.value.code == -6
    with pytest.raises(RsxError) as ex:
        e.calc_curve(np.zeros((5, 4), np.float32))
    assert ex.value.code == -3          # SAXS_INVALID_ARG
    e.calc_curve(b)
    with pytest.raises(RsxError) as ex:
        e.recalc_curve(b, 3, 2)
    assert ex.value.code == -3


def test_dirty_params_force_full_recompute(oracle):
    from resaxs_amd import configs
    _say("case: water weight change forces a full recompute")
    n = 2500
    w = configs.sasa_ellipsoid_20000(n=n)
    e = engine_for(w.q, w.table, w.water_index, 0.0, n=n, radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.calc_curve(w.bodies)
    e.set_water_weight(w.w)
    I = e.recalc_curve(w.bodies, 0, 0)
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    counts = oracle.sasa_counts(w.pos, w.types, r2_table(w.radii, w.probe), dots)
    ref = oracle_iq(w.q, w.pos, w.types, w.table, w.water_index, w.w, counts, 256)
    assert rel_err(I, ref).max() < 1e-5


def test_hydrated_recalc(oracle):
    """With hydration every move recomputes SASA; atoms whose exposed-dot count
    changes join the update set.  After 10 moves the running I(q) equals the
    FP64 oracle on the final positions and final counts."""
    from resaxs_amd import configs
    _say("case: hydrated recalc, 10 moves of 150 atoms on 3000 atoms")
    n = 3000
    w = configs.incremental_50000(n=n)
    pos = w.pos.copy()
    e = engine_for(w.q, w.table, w.water_index, w.w, n=n, radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.set_rebaseline(0)
    e.calc_curve(configs.bodies_of(pos, w.types))
    rng = np.random.default_rng(36)
    for m in range(10):
        start, newp = configs.random_move(rng, pos, 150)
        pos[start:start + 150] = newp
        I = e.recalc_curve(configs.bodies_of(pos, w.types), start, 150)
        e.commit()
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    counts = oracle.sasa_counts(pos, w.types, r2_table(w.radii, w.probe), dots)
    assert np.array_equal(e.get_rsasa(), counts.astype(np.float32) / np.float32(256))
    ref = oracle_iq(w.q, pos, w.types, w.table, w.water_index, w.w, counts, 256)
    err = rel_err(I, ref)
    _say(f"  max rel err {err.max():.3e}")
    assert err.max() < 1e-5


@pytest.mark.parametrize("variant,uniform_q", [(3, False), (0, True), (2, True)])
def test_hydrated_recalc_kernels(oracle, variant, uniform_q):
    """Hydrated moves through the other pair kernels: the direct kernel of a non-uniform
    q grid, the FP64 recurrence and the per-lane-sum FP32 kernel.  Atoms whose count
    changed without moving enter as delta-class rows in each of them; after 6 moves the
    running I(q) equals the FP64 oracle on the final positions and counts."""
    from resaxs_amd import configs
    _say(f"case: hydrated recalc, variant {variant}, {'uniform' if uniform_q else 'non-uniform'} q, "
         f"6 moves of 150 atoms on 2500 atoms")
    n = 2500
    w = configs.incremental_50000(n=n)
    q = w.q
    if not uniform_q:
        rng_q = np.random.default_rng(3)
        q = np.sort(np.concatenate([[0.0], rng_q.uniform(0.005, 0.6, size=len(w.q) - 1)])).astype(np.float32)
    table = configs.factor_table(np.random.default_rng(4), q)
    pos = w.pos.copy()
    e = engine_for(q, table, w.water_index, w.w, n=n, radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.set_variant(variant)
    e.set_rebaseline(0)
    e.calc_curve(configs.bodies_of(pos, w.types))
    rng = np.random.default_rng(37)
    for m in range(6):
        start, newp = configs.random_move(rng, pos, 150)
        pos[start:start + 150] = newp
        I = e.recalc_curve(configs.bodies_of(pos, w.types), start, 150)
        e.commit()
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    counts = oracle.sasa_counts(pos, w.types, r2_table(w.radii, w.probe), dots)
    ref = oracle_iq(q, pos, w.types, table, w.water_index, w.w, counts, 256)
    err = rel_err(I, ref)
    _say(f"  max rel err {err.max():.3e}")
    assert err.max() < 1e-5


def _full_recompute(w, pos, hydrate, variant):
    """Fresh engine, full calc_curve with the given pair-kernel variant."""
    from resaxs_amd import configs
    kw = dict(radii=w.radii, probe=w.probe, n_half=w.n_half) if hydrate else {}
    f = engine_for(w.q, w.table, w.water_index, w.w if hydrate else 0.0, n=len(pos), **kw)
    f.set_variant(variant)
    I = f.calc_curve(configs.bodies_of(pos, w.types))
    counts = f.get_rsasa() if hydrate else None
    f.close()
    return I, counts


def test_incremental_full_size_plain():
    """The incremental workload at full size: 50,000 atoms x 256 q, 500-atom rigid
    moves, the running FP64 I(q) re-baselined every 200 moves.  After 199 committed
    moves (the longest run without a re-baseline) the running I(q) equals a fresh
    full recompute by the FP64 pair kernel within 1e-5; the 200th move re-baselines."""
    from resaxs_amd import configs
    _say("case: 50000 atoms x 256 q, 200 moves of 500 atoms, re-baseline every 200, plain")
    w = configs.incremental_50000()
    pos = w.pos.copy()
    e = engine_for(w.q, w.table, n=len(pos))
    e.set_rebaseline(200)
    b = configs.bodies_of(pos, w.types)
    e.calc_curve(b)
    rng = np.random.default_rng(301)
    for m in range(199):
        start, newp = configs.random_move(rng, pos, 500)
        pos[start:start + 500] = newp
        b[start:start + 500, :3] = newp
        I = e.recalc_curve(b, start, 500)
        e.commit()
    ref, _ = _full_recompute(w, pos, False, 0)
    err = rel_err(I, ref)
    _say(f"  running I after 199 moves vs FP64 full recompute: max rel err {err.max():.3e}")
    assert err.max() < 1e-5
    start, newp = configs.random_move(rng, pos, 500)
    pos[start:start + 500] = newp
    b[start:start + 500, :3] = newp
    I = e.recalc_curve(b, start, 500)
    e.commit()
    ref, _ = _full_recompute(w, pos, False, 0)
    err = rel_err(I, ref)
    _say(f"  after the re-baselining 200th move: max rel err {err.max():.3e}")
    assert err.max() < 1e-5
    e.close()


def test_incremental_full_size_hydrated(oracle):
    """The SASA variant of the incremental workload at full size: every move
    recomputes the exposed-dot counts of all 50,000 atoms (bit-exact against the
    oracle's cell-list count) and the running I(q) follows a fresh FP64 recompute."""
    from resaxs_amd import configs
    _say("case: 50000 atoms x 256 q, 20 moves of 500 atoms, SASA hydration, no re-baseline")
    w = configs.incremental_50000()
    pos = w.pos.copy()
    e = engine_for(w.q, w.table, w.water_index, w.w, n=len(pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.set_rebaseline(0)
    b = configs.bodies_of(pos, w.types)
    e.calc_curve(b)
    rng = np.random.default_rng(302)
    for m in range(20):
        start, newp = configs.random_move(rng, pos, 500)
        pos[start:start + 500] = newp
        b[start:start + 500, :3] = newp
        I = e.recalc_curve(b, start, 500)
        e.commit()
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    counts = oracle.sasa_counts(pos, w.types, r2_table(w.radii, w.probe), dots)
    assert np.array_equal(e.get_rsasa(), counts.astype(np.float32) / np.float32(256))
    ref, ref_s = _full_recompute(w, pos, True, 0)
    assert np.array_equal(ref_s, e.get_rsasa())
    err = rel_err(I, ref)
    _say(f"  running I after 20 hydrated moves vs FP64 full recompute: max rel err {err.max():.3e}")
    assert err.max() < 1e-5
    e.close()


def test_hydrated_local_recount_with_revert(oracle):
    """Hydrated moves recount SASA only near the moved segment; after a revert the
    device copy of the positions is stale and the next move uploads them all.  Counts
    stay bit-exact and I(q) follows the FP64 oracle through commit / revert / move."""
    from resaxs_amd import configs
    _say("case: hydrated local recount with commit / revert, 8000 atoms, moves of 120 atoms")
    n = 8000
    w = configs.incremental_50000(n=n)
    pos = w.pos.copy()
    e = engine_for(w.q, w.table, w.water_index, w.w, n=n, radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.set_rebaseline(0)
    b = configs.bodies_of(pos, w.types)
    e.calc_curve(b)
    rng = np.random.default_rng(37)
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    r2 = r2_table(w.radii, w.probe)
    for step in range(6):
        start, newp = configs.random_move(rng, pos, 120)
        trial = pos.copy()
        trial[start:start + 120] = newp
        bt = configs.bodies_of(trial, w.types)
        I = e.recalc_curve(bt, start, 120)
        if step % 3 == 2:
            e.revert()                     # reject: back to pos
            # the restored exposure fractions are the committed ones
            assert np.array_equal(e.get_rsasa(), counts.astype(np.float32) / np.float32(256)), step
            continue
        e.commit()
        pos = trial
        counts = oracle.sasa_counts(pos, w.types, r2, dots)
        assert np.array_equal(e.get_rsasa(), counts.astype(np.float32) / np.float32(256)), step
        ref = oracle_iq(w.q, pos, w.types, w.table, w.water_index, w.w, counts, 256)
        err = rel_err(I, ref)
        _say(f"  step {step}: max rel err {err.max():.3e}")
        assert err.max() < 1e-5


def test_hydrated_recount_fallbacks_and_size_change(oracle):
    """The local recount's other paths.  A jittered move of 9000 atoms is more than the
    captured recount's stage carries (the recount runs as eager launches) and changes 6130
    exposed-dot counts, more than the read-back block holds (every count is compared);
    small moves before and after it take the captured graph.  Then the engine takes a
    structure of another size, whose moves capture the graph again.  Counts stay bit-exact
    and I(q) follows the FP64 oracle throughout."""
    from resaxs_amd import configs

    def check(e, w, pos, what):
        dots = fib_dot_table(w.radii, w.probe, w.n_half)
        counts = oracle.sasa_counts(pos, w.types, r2_table(w.radii, w.probe), dots)
        assert np.array_equal(e.get_rsasa(), counts.astype(np.float32) / np.float32(256)), what
        return counts

    _say("case: hydrated recount fallbacks, 12000 atoms: 200-atom move, jitter of 9000 atoms, 150-atom moves")
    rng = np.random.default_rng(41)
    w = configs.incremental_50000(n=12000)
    e = engine_for(w.q, w.table, w.water_index, w.w, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    e.set_rebaseline(0)
    pos = w.pos.copy()
    e.calc_curve(configs.bodies_of(pos, w.types))
    moves = [("move 200", *configs.random_move(rng, pos, 200))]
    jitter = pos[1000:10000] + np.random.default_rng(5).normal(scale=0.4, size=(9000, 3)).astype(np.float32)
    moves.append(("jitter 9000", 1000, jitter.astype(np.float32)))
    for what, start, newp in moves + [("move 150", None, None), ("move 150", None, None)]:
        if start is None:
            start, newp = configs.random_move(rng, pos, 150)
        pos[start:start + len(newp)] = newp
        I = e.recalc_curve(configs.bodies_of(pos, w.types), start, len(newp))
        e.commit()
        counts = check(e, w, pos, what)
        err = rel_err(I, oracle_iq(w.q, pos, w.types, w.table, w.water_index, w.w, counts, 256)).max()
        _say(f"  {what}: max rel err {err:.3e}")
        assert err < 1e-5, what

    _say("case: the same engine on a structure of 7000 atoms, two 150-atom moves")
    w2 = configs.incremental_50000(seed=11, n=7000)
    pos2 = w2.pos.copy()
    e.initialize(len(pos2))
    e.calc_curve(configs.bodies_of(pos2, w2.types))
    for k in range(2):
        start, newp = configs.random_move(rng, pos2, 150)
        pos2[start:start + 150] = newp
        I = e.recalc_curve(configs.bodies_of(pos2, w2.types), start, 150)
        e.commit()
        counts = check(e, w2, pos2, f"size change move {k}")
        # the engine keeps the first workload's q and factor table
        err = rel_err(I, oracle_iq(w.q, pos2, w2.types, w.table, w.water_index, w.w, counts, 256)).max()
        _say(f"  move {k}: max rel err {err:.3e}")
        assert err < 1e-5
    e.close()


def test_hydrated_recount_slab_rebuild_and_overflow(oracle):
    """The slab cell list of the local recounts.  300 committed 40-atom moves pass its
    periodic rebuild (every 256 recounts, which clears the tombstones moves leave).  Then a
    move gathers 80 atoms into a 0.6 A ball, more than a bucket holds: that recount reports
    the overflow and the counts come from a full count instead; the next moves rebuild the
    slab, overflow again right after the build, and go on without it.  Counts stay bit-exact
    and I(q) follows the FP64 oracle."""
    from resaxs_amd import configs

    def check(e, w, pos, what, iq):
        dots = fib_dot_table(w.radii, w.probe, w.n_half)
        counts = oracle.sasa_counts(pos, w.types, r2_table(w.radii, w.probe), dots)
        assert np.array_equal(e.get_rsasa(), counts.astype(np.float32) / np.float32(256)), what
        err = rel_err(iq, oracle_iq(w.q, pos, w.types, w.table, w.water_index, w.w, counts, 256)).max()
        _say(f"  {what}: max rel err {err:.3e}")
        assert err < 1e-5, what

    _say("case: slab rebuild and overflow, 6000 atoms: 300 moves of 40 atoms, then a gathering move")
    rng = np.random.default_rng(43)
    w = configs.incremental_50000(n=6000)
    e = engine_for(w.q, w.table, w.water_index, w.w, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    pos = w.pos.copy()
    e.calc_curve(configs.bodies_of(pos, w.types))
    for k in range(300):
        start, newp = configs.random_move(rng, pos, 40)
        pos[start:start + 40] = newp
        I = e.recalc_curve(configs.bodies_of(pos, w.types), start, 40)
        e.commit()
    check(e, w, pos, "after 300 moves", I)
    # 80 atoms into a 0.7 A cube at the centre of the cell of the segment's first atom (cell
    # edge 2 max(r + p) + 0.02, as the engine's cell lists use), so all land in one bucket
    start = 2000
    cell = 2.0 * float(np.float32(max(w.radii)) + np.float32(w.probe)) + 0.02
    centre = (np.floor(pos[start].astype(np.float64) / cell) + 0.5) * cell
    gather = centre + np.random.default_rng(9).uniform(-0.35, 0.35, size=(80, 3))
    pos[start:start + 80] = gather.astype(np.float32)
    I = e.recalc_curve(configs.bodies_of(pos, w.types), start, 80)
    e.commit()
    check(e, w, pos, "gathering move", I)
    for k in range(3):
        start, newp = configs.random_move(rng, pos, 40)
        pos[start:start + 40] = newp
        I = e.recalc_curve(configs.bodies_of(pos, w.types), start, 40)
        e.commit()
        check(e, w, pos, f"move {k} after the gathering", I)
    e.close()
