"""Distance-histogram Debye (rsx_calc_curve_hist; no counterpart in the reference,
whose saxs_hist_cl_gpu class runs the exact pt_wf kernels).  Approximate by construction: each pair's sinc is taken at its bin
centre.  Its accuracy contract, checked against the FP64 oracle:
  * I(0) exact: (sum F)^2 to 1e-12 (the histogram counts every pair once);
  * the error falls as the square of the bin width (ratio >= 3 between 0.04 and 0.02 A);
  * at 0.02 A bins, max relative error <= 1e-3 over 0 <= q <= 0.5 for a 3000-atom
    protein-like body, plain and hydrated (measured 1.2e-4 plain, 4.3e-4 hydrated; at
    0.01 A 5e-5 plain).
"""
import sys

import numpy as np
import pytest

from helpers import engine_for, fib_dot_table, oracle_iq, r2_table, rel_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _protein(seed, n, q):
    from resaxs_amd import configs
    rng, pos, types, table = configs._protein_like(seed, n, q)
    return pos, types, table


def test_hist_plain_accuracy_and_order(oracle):
    from resaxs_amd import configs
    _say("case: histogram Debye, 3000 atoms x 128 q, plain, bins 0.04 / 0.02 / 0.01 A")
    q = np.linspace(0.0, 0.5, 128).astype(np.float32)
    pos, types, table = _protein(61, 3000, q)
    e = engine_for(q, table, n=len(pos))
    b = configs.bodies_of(pos, types)
    ref = oracle_iq(q, pos, types, table)
    errs = {}
    for dr in (0.04, 0.02, 0.01):
        I = e.calc_curve_hist(b, dr)
        errs[dr] = rel_err(I, ref).max()
        _say(f"  bin {dr:.2f} A: max rel err {errs[dr]:.3e}")
        F0 = table[0][types].astype(np.float64).sum()
        assert abs(I[0] - F0 * F0) <= 1e-12 * F0 * F0
    assert errs[0.02] <= 1e-3
    assert errs[0.04] / errs[0.02] >= 3.0


def test_hist_hydrated_accuracy(oracle):
    from resaxs_amd import configs
    _say("case: histogram Debye, 3000 atoms x 128 q, SASA hydration, bins 0.02 A")
    n = 3000
    w = configs.sasa_ellipsoid_20000(n=n)
    q = np.linspace(0.0, 0.5, 128).astype(np.float32)
    table = configs.factor_table(np.random.default_rng(62), q)
    e = engine_for(q, table, w.water_index, w.w, n=n, radii=w.radii, probe=w.probe, n_half=w.n_half)
    I = e.calc_curve_hist(w.bodies, 0.02)
    dots = fib_dot_table(w.radii, w.probe, w.n_half)
    counts = oracle.sasa_counts(w.pos, w.types, r2_table(w.radii, w.probe), dots)
    ref = oracle_iq(q, w.pos, w.types, table, w.water_index, w.w, counts, 256)
    err = rel_err(I, ref).max()
    _say(f"  max rel err {err:.3e}")
    assert err <= 1e-3


def test_hist_plain_several_windows(oracle):
    """Plain histograms larger than one LDS window (16384 bins): 2000 atoms in a ball of
    radius 60 A at 0.005 A bins (about 24,000 bins, two windows).  Every pair is counted
    exactly once (I(0) exact), and the result meets the bin-width error of a fine bin."""
    from resaxs_amd import configs
    _say("case: histogram Debye, 2000 atoms in a 60 A ball x 64 q, plain, bins 0.005 A (two windows)")
    rng = np.random.default_rng(71)
    q = np.linspace(0.0, 0.5, 64).astype(np.float32)
    pos = configs.in_ball(rng, 2000, 60.0).astype(np.float32)
    types = configs.seeded_types(rng, 2000)
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=len(pos))
    I = e.calc_curve_hist(configs.bodies_of(pos, types), 0.005)
    F0 = table[0][types].astype(np.float64).sum()
    assert abs(I[0] - F0 * F0) <= 1e-12 * F0 * F0
    err = rel_err(I, oracle_iq(q, pos, types, table)).max()
    _say(f"  max rel err {err:.3e}")
    assert err <= 2e-4
