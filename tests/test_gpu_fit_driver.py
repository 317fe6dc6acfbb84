"""The calc_saxs ensemble fitting driver (include/resaxs/calc_saxs.hpp) driven by
tests/cpp/fit_test, a C++14 program written the way the calc_saxs tool uses it
(calc_saxs/main.cpp:264-296): profile file -> grid search over the expansion factor c
and the water weight w with the least-squares scale and chi
(calc_saxs/calc_saxs.hpp:420-439, 531-607).

The target profile is the FP64 oracle's ensemble average at a known (c*, w*) times
a known scale, with 1 % sigmas: the search must return that point (to its own
stopping tolerances), the evaluator's ensemble I(q) must equal the oracle's at the
returned point and at the truth, and chi at the truth must vanish.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_err
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "fit_test")
OH2 = 14


def _avg_radius(types, radii):
    acc = np.float32(0)
    for t in types:
        acc = np.float32(acc + radii[t])
    return np.float32(acc / np.float32(len(types)))


def _ensemble_oracle(q, models, types, c, w):
    """Uniform mean over the models of the hydrated I(q) the tool's evaluator computes:
    atomic factors at the first model's mean radius with expansion factor c, water row
    OH2 at weight w, SASA with the 32 dodeca-icosa dots, probe 1.8 A, atomic radii."""
    radii = po.atomic_radii()
    table = po.atomic_ff(q, _avg_radius(types, radii), c)
    dots = np.stack([po.dodeca_icosa_dots(r, 1.8) for r in radii])
    R2 = ((radii + np.float32(1.8)) ** 2).astype(np.float32)
    acc = np.zeros(len(q))
    for pos in models:
        counts = po.sasa_counts(pos, types, R2, dots)
        frac = (counts.astype(np.float32) / np.float32(32)).astype(np.float32)
        acc += po.debye(q, pos, po.body_factors(pos, types, q, table, OH2, w, frac))
    return acc / len(models)


def test_fit_ensemble_recovers_known_parameters(tmp_path):
    from resaxs_amd import configs
    print("case: calc_saxs fit, 3 models x 1200 atoms x 80 q, target c=1.023 w=1.37 scale=2.5", flush=True)
    assert os.path.exists(BIN), "build tests/cpp/fit_test first (__graft_entry__.build())"
    rng = np.random.default_rng(77)
    n, n_models = 1200, 3
    radius = (n / 0.045 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
    base = configs.in_ball(rng, n, radius)
    # atomic types: H, C, N, O, S and the CH / CH2 / CH3 / NH / OH groups
    types = rng.choice(np.array([0, 1, 2, 3, 5, 7, 8, 9, 10, 13], dtype=np.uint32), size=n,
                       p=[0.30, 0.25, 0.08, 0.1, 0.02, 0.08, 0.07, 0.04, 0.03, 0.03]).astype(np.uint32)
    models = [(base + rng.normal(scale=0.5, size=base.shape)).astype(np.float32) for _ in range(n_models)]
    q = np.linspace(0.01, 0.4, 80).astype(np.float32)
    c_true, w_true, s_true = 1.023, 1.37, 2.5
    target = s_true * _ensemble_oracle(q, models, types, c_true, w_true)
    sigma = 0.01 * target
    d = str(tmp_path)
    with open(os.path.join(d, "models.bin"), "wb") as f:
        np.array([n_models, n], dtype=np.int32).tofile(f)
        for pos in models:
            po.bodies(pos, types).tofile(f)
    with open(os.path.join(d, "profile.dat"), "w") as f:
        f.write("# synthetic profile: q I sigma\n")
        for a, b, e in zip(q, target, sigma):
            f.write(f"{a:.9e} {b:.9e} {e:.9e}\n")
    r = subprocess.run([BIN, d, repr(c_true), repr(w_true)], cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    c, w, scale, chi, evals, engines = np.fromfile(os.path.join(d, "fit.bin"), dtype=np.float64)
    assert engines == 1 and evals % n_models == 0
    assert abs(c - c_true) < 2e-3 and abs(w - w_true) < 2e-2, (c, w)
    assert abs(scale / s_true - 1) < 2e-3 and chi < 0.05, (scale, chi)
    # the evaluator against the oracle: at the returned point and at the truth
    fit_iq = np.fromfile(os.path.join(d, "fit_iq.bin"), dtype=np.float32)
    assert rel_err(fit_iq, scale * _ensemble_oracle(q, models, types, np.float32(c), np.float32(w))).max() < 1e-5
    p_scale, p_chi = np.fromfile(os.path.join(d, "point.bin"), dtype=np.float64)
    p_iq = np.fromfile(os.path.join(d, "point_iq.bin"), dtype=np.float32)
    assert abs(p_scale / s_true - 1) < 1e-5 and p_chi < 2e-3, (p_scale, p_chi)
    assert rel_err(p_iq, target).max() < 1e-5


def test_fit_ensemble_c_abi():
    """The same search through the C ABI (rsx_fit_ensemble).  A fixed point (every range
    a single value) returns the oracle's ensemble mean; a full search returns a chi no
    worse than the best point of its own first 6 x 6 round (the best only improves,
    calc_saxs.hpp:571-575), evaluated here point by point."""
    from resaxs_amd import configs, fit_ensemble
    print("case: rsx_fit_ensemble, 2 models x 900 atoms x 60 q", flush=True)
    rng = np.random.default_rng(78)
    n = 900
    radius = (n / 0.045 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
    base = configs.in_ball(rng, n, radius)
    types = rng.choice(np.array([1, 2, 3, 5, 8, 9, 13], dtype=np.uint32), size=n).astype(np.uint32)
    models = [(base + rng.normal(scale=0.4, size=base.shape)).astype(np.float32) for _ in range(2)]
    bodies = np.concatenate([po.bodies(p, types) for p in models])
    offsets = np.array([0, n, 2 * n], dtype=np.int32)
    q = np.linspace(0.02, 0.45, 60).astype(np.float32)
    c_true, w_true = np.float32(0.985), np.float32(-0.6)
    ones = np.ones(len(q), dtype=np.float32)
    c, w, s, chi, ev, iq = fit_ensemble(bodies, offsets, q, ones, ones, scale=1.0, water=(w_true, w_true),
                                        expansion=(c_true, c_true))
    assert (c, s) == (float(c_true), 1.0) and w == float(w_true) and ev == 2
    ref = _ensemble_oracle(q, models, types, c_true, w_true)
    assert rel_err(iq, ref).max() < 1e-5
    target = (3.0 * ref).astype(np.float32)
    sigma = (0.01 * target).astype(np.float32)
    c, w, s, chi, ev, iq = fit_ensemble(bodies, offsets, q, target, sigma)
    print(f"  search: c {c:.5f} w {w:.4f} scale {s:.5f} chi {chi:.3g} in {ev} evaluations", flush=True)
    first_round = min(fit_ensemble(bodies, offsets, q, target, sigma, water=(ww, ww), expansion=(ef, ef))[3]
                      for ef in np.float32(0.95) + np.arange(6, dtype=np.float32) * np.float32(0.02)
                      for ww in np.float32(-2.0) + np.arange(6, dtype=np.float32) * np.float32(1.2))
    print(f"  best chi of the first 6 x 6 round: {first_round:.3g}", flush=True)
    assert chi <= first_round * (1 + 1e-6)
