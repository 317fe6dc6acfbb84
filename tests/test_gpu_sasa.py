"""GPU SASA counts (bit-exact vs the FP32 oracle) and the hydrated Debye sum.

The dot templates come from the library's own generator, which equals the
oracle's restatement of rsasa_algorithm.cpp:96-116 bit for bit (checked in the
CPU suite), so both sides test the same dots with the same FP32 predicate.
"""
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
