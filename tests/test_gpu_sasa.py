"""GPU SASA counts (bit-exact vs the FP32 oracle) and the hydrated Debye sum.

The dot templates come from the library's own generator, which equals the
oracle's restatement of rsasa_algorithm.cpp:96-116 bit for bit (checked in the
CPU suite), so both sides test the same dots with the same FP32 predicate.
"""
import os
import sys

import numpy as np
import pytest

from helpers import engine_for, fib_dot_table, oracle_iq, padded_density_dots, r2_table, rel_err
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine():
    from resaxs_amd import Engine
    return Engine(0)


SASA_TIES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sasa_ties.npz")


def test_sasa_near_ties_bit_exact():
    """Counts on the near-tie fixture (tests/golden/make_sasa_ties.py): 40 atom pairs with a
    dot within an ulp of the 1e-4 occlusion threshold, where every fused ordering of the
    squared distance flips 14-30 of the 80 counts.  The kernel's counts equal the unfused
    FP32 predicate's bit for bit."""
    g = np.load(SASA_TIES)
    _say("case: SASA near-tie fixture, 80 atoms x 64 dots")
    e = _engine()
    c, frac = e.calc_rsasa(g["bodies"], g["radii"], g["probe"], g["dots"])
    assert np.array_equal(c, g["counts"]), np.nonzero(c != g["counts"])[0]
    assert np.array_equal(frac, g["counts"].astype(np.float32) / np.float32(64))


def test_hydration_composition_closed_forms_gpu(oracle):
    """The engine's hydrated I(q) against closed forms of F_i = f_t + (count_i / D) w f_water
    (gpu_saxs_pt_wf.cl:130-138) with the engine's own exposed-dot counts: one isolated atom of
    each type (every dot exposed) gives (f_t + w f_water)^2; two overlapping atoms of
    different types give F1^2 + F2^2 + 2 F1 F2 sin(qr)/(qr).  Every table column differs, so
    a swapped water index or a misapplied fraction fails."""
    rng = np.random.default_rng(17)
    q = np.linspace(0.0, 0.5, 64).astype(np.float32)
    water = 4
    table = rng.uniform(0.5, 8.0, size=(len(q), 5)).astype(np.float32)
    t64 = table.astype(np.float64)
    radii = np.array([1.7, 1.55, 1.52, 1.8], dtype=np.float32)
    probe, n_half, w = 1.4, 32, 0.61
    D = 2 * n_half
    qr_sinc = lambda r: np.where(q == 0, 1.0, np.sin(q.astype(np.float64) * r) / np.where(q == 0, 1.0, q * r))
    _say("case: hydration composition, closed forms")
    for t in range(4):
        e = engine_for(q, table, water, w, n=1, radii=radii, probe=probe, n_half=n_half)
        I = e.calc_curve(oracle.bodies(np.zeros((1, 3)), [t]))
        assert e.get_rsasa()[0] == 1.0
        expect = (t64[:, t] + np.float64(np.float32(w)) * t64[:, water]) ** 2
        assert rel_err(I, expect).max() < 1e-6, t
        e.close()
    r = 3.1
    pos = np.array([[0, 0, 0], [r, 0, 0]], dtype=np.float32)
    e = engine_for(q, table, water, w, n=2, radii=radii, probe=probe, n_half=n_half)
    I = e.calc_curve(oracle.bodies(pos, [1, 3]))
    counts = oracle.sasa_counts(pos, [1, 3], r2_table(radii, probe), fib_dot_table(radii, probe, n_half),
                                cells=False)
    assert 0 < counts.min() and counts.max() < D
    s = (counts.astype(np.float32) / np.float32(D)).astype(np.float64)
    assert np.array_equal(e.get_rsasa(), counts.astype(np.float32) / np.float32(D))
    wf = np.float64(np.float32(w)) * t64[:, water]
    F1, F2 = t64[:, 1] + s[0] * wf, t64[:, 3] + s[1] * wf
    assert rel_err(I, F1 ** 2 + F2 ** 2 + 2 * F1 * F2 * qr_sinc(r)).max() < 1e-5


def test_sasa_counts_config2_bit_exact(oracle):
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000()
    _say("case: SASA counts, 20000 atoms x 256 dots")
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    e = _engine()
    counts, frac = e.calc_rsasa(w.bodies, w.radii, w.probe, dots)
    ref = oracle.sasa_counts(w.pos, w.types, r2_table(w.radii, w.probe), dots)
    assert np.array_equal(counts, ref)
    assert np.array_equal(frac, (ref.astype(np.float32) / np.float32(256)))
    assert counts.max() <= 256 and counts.sum() > 0


def test_sasa_small_cases_bruteforce(oracle):
    """Isolated atom exposes every dot; a buried atom none; random clusters
    equal the brute-force oracle (no cell pruning)."""
    _say("case: SASA small cases")
    radii = np.array([1.7, 1.2], dtype=np.float32)
    probe = 1.4
    dots = fib_dot_table(radii, probe, 32)
    e = _engine()
    one = oracle.bodies(np.zeros((1, 3)), [0])
    c, _ = e.calc_rsasa(one, radii, probe, dots)
    assert c[0] == 64
    # centre atom inside a tight shell of 40 neighbours at 2.0 A
    g = oracle.fibonacci_dots(20, 0.0, 2.0)[:, :3]
    pos = np.vstack([np.zeros((1, 3)), g]).astype(np.float32)
    c, _ = e.calc_rsasa(oracle.bodies(pos, np.zeros(len(pos), np.uint32)), radii, probe, dots)
    assert c[0] == 0
    rng = np.random.default_rng(21)
    for n in (2, 7, 150):
        pos = rng.uniform(-6, 6, size=(n, 3)).astype(np.float32)
        types = rng.integers(0, 2, size=n).astype(np.uint32)
        c, _ = e.calc_rsasa(oracle.bodies(pos, types), radii, probe, dots)
        ref = oracle.sasa_counts(pos, types, r2_table(radii, probe), dots, cells=False)
        assert np.array_equal(c, ref), n


@pytest.mark.parametrize("n_half", [16, 50, 100, 256])
def test_sasa_dots_per_lane(oracle, n_half):
    """Dot counts of 32, 100, 200 and 512 dots per sphere (one, two, four and eight dots per
    lane, the last one ragged over the lanes) on a dense 1500-atom cluster, bit-exact
    against the brute-force oracle."""
    _say(f"case: SASA {2 * n_half} dots, 1500 atoms")
    radii = np.array([1.7, 1.55, 1.52, 1.8, 1.2], dtype=np.float32)
    probe = 1.4
    dots = fib_dot_table(radii, probe, n_half)
    rng = np.random.default_rng(90 + n_half)
    pos = rng.uniform(-14, 14, size=(1500, 3)).astype(np.float32)
    types = rng.integers(0, 5, size=1500).astype(np.uint32)
    e = _engine()
    c, _ = e.calc_rsasa(oracle.bodies(pos, types), radii, probe, dots)
    ref = oracle.sasa_counts(pos, types, r2_table(radii, probe), dots, cells=False)
    assert np.array_equal(c, ref)
    assert 0 < int(np.count_nonzero(c)) < 1500  # both buried and exposed atoms


def test_hydrated_debye_vs_oracle(oracle):
    """Hydration factors F = f_t + (count/D) w f_w from the GPU's own counts;
    I(q) against the FP64 oracle using the oracle's (identical) counts."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000(n=6000)
    _say("case: hydrated Debye 6000 x 256")
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    e = engine_for(w.q, w.table, w.water_index, w.w, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    I = e.calc_curve(w.bodies)
    counts = oracle.sasa_counts(w.pos, w.types, r2_table(w.radii, w.probe), dots)
    s = e.get_rsasa()
    assert np.array_equal(s, counts.astype(np.float32) / np.float32(256))
    ref = oracle_iq(w.q, w.pos, w.types, w.table, w.water_index, w.w, counts, 256)
    err = rel_err(I, ref)
    _say(f"  max rel err {err.max():.3e}")
    assert err.max() < 1e-5


def test_hydrated_full_size(oracle):
    """The hydrated forward at full size: 20,000 atoms x 256 q with SASA hydration."""
    from resaxs_amd import configs
    w = configs.sasa_ellipsoid_20000()
    _say("case: hydrated Debye 20000 x 256 (full size)")
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    e = engine_for(w.q, w.table, w.water_index, w.w, n=len(w.pos), radii=w.radii, probe=w.probe, n_half=w.n_half)
    I = e.calc_curve(w.bodies)
    counts = oracle.sasa_counts(w.pos, w.types, r2_table(w.radii, w.probe), dots)
    ref = oracle_iq(w.q, w.pos, w.types, w.table, w.water_index, w.w, counts, 256)
    err = rel_err(I, ref)
    k = int(np.argmax(err))
    _say(f"  max rel err {err.max():.3e} at q={w.q[k]:.4f}, median {np.median(err):.3e}")
    assert err.max() < 1e-5


@pytest.mark.skipif(po.ref_lib() is None, reason="reference CPU subset (oracle/_ref) not built")
def test_sasa_counts_vs_reference_cpu_algorithm():
    """GPU exposed-dot counts equal the reference's own CPU SASA (rsasa_cpu_alg.cpp:112-185,
    voxel-pruned, compiled from its sources) on its density dot spheres, the 16 atomic
    radii (ff_coef.hpp:161-184) and its 1.8 A probe."""
    from resaxs_amd import configs
    radii = po.ref_atomic_radii()
    probe = 1.8
    rng = np.random.default_rng(41)
    e = _engine()
    for density, n in ((0.5, 2500), (2.0, 6000)):
        _say(f"case: reference CPU SASA, density {density}, {n} atoms")
        dots, nd = padded_density_dots(density, radii, probe)
        pos = configs.in_ball(rng, n, (n / 0.045 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)).astype(np.float32)
        types = rng.integers(0, 16, size=n).astype(np.uint32)
        b = po.bodies(pos, types)
        counts, _ = e.calc_rsasa(b, radii, probe, dots)
        frac = po.ref_rsasa_cpu(b, density, probe)
        ref = np.rint(frac * nd[types]).astype(np.int64)
        assert np.array_equal(counts.astype(np.int64), ref)
        assert 0 < ref.sum() < nd[types].sum()
