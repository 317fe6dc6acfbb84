"""rsx_calc_rsasa_noalloc: the SASA counts of rsx_calc_rsasa with the cell list in a static
device buffer (no allocation in the call), checked against rsx_calc_rsasa bit for bit on two
calls in a row, the second smaller (what the first left in the shared buffer must not leak
into it), plus its size limits and its handling of out-of-range types."""
import ctypes
import sys

import numpy as np
import pytest

from helpers import fib_dot_table
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

MAX_ATOMS = 65536


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(e, bodies, radii, probe, dots):
    """counts from rsx_calc_rsasa_noalloc on device copies of bodies and dots, launched on the
    engine's stream"""
    L = e._L
    b = np.ascontiguousarray(bodies, dtype=np.float32)
    d = np.ascontiguousarray(dots, dtype=np.float32)
    r = np.ascontiguousarray(radii, dtype=np.float32)
    n = len(b)
    pb, pd, pc = e.dev_alloc(b.nbytes), e.dev_alloc(d.nbytes), e.dev_alloc(4 * max(n, 1))
    try:
        e.h2d(pb, b)
        e.h2d(pd, d)
        rc = L.rsx_calc_rsasa_noalloc(pb, n, r.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(r),
                                      float(probe), pd, d.shape[1], pc, L.rsx_stream(e.h))
        assert rc == 0, rc
        e.synchronize()
        out = np.zeros(n, dtype=np.uint32)
        e.d2h(out, pc)
        return out
    finally:
        for p in (pb, pd, pc):
            e.dev_free(p)


def _cluster(seed, n, types):
    rng = np.random.default_rng(seed)
    radius = (n / 0.045 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
    pos = rng.uniform(-radius, radius, size=(3 * n, 3))
    pos = pos[np.linalg.norm(pos, axis=1) <= radius][:n].astype(np.float32)
    return po.bodies(pos, rng.integers(0, types, size=n).astype(np.uint32))


def test_noalloc_sasa_equals_rsx_calc_rsasa():
    from resaxs_amd import Engine
    radii = np.array([1.7, 1.55, 1.52, 1.8, 1.2], dtype=np.float32)
    probe = 1.4
    e = Engine(0)
    # first call: 6000 atoms, 200 dots (four per lane); second: 700 atoms, 64 dots
    for seed, n, n_half in ((3, 6000, 100), (4, 700, 32)):
        _say(f"case: no-alloc SASA, {n} atoms x {2 * n_half} dots")
        b = _cluster(seed, n, len(radii))
        dots = fib_dot_table(radii, probe, n_half)
        ref, _ = e.calc_rsasa(b, radii, probe, dots)
        got = _noalloc(e, b, radii, probe, dots)
        assert np.array_equal(got, ref), int(np.sum(got != ref))
        assert 0 < ref.sum() < n * 2 * n_half
    e.close()


def test_noalloc_sasa_limits_and_foreign_types():
    from resaxs_amd import Engine
    radii = np.array([1.7, 1.2], dtype=np.float32)
    probe = 1.4
    e = Engine(0)
    L = e._L
    r = radii.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    dots = fib_dot_table(radii, probe, 16)
    _say("case: no-alloc SASA size limits")
    # beyond the stated maximum: rejected before anything runs (no pointer is touched)
    fake = ctypes.c_void_p(16)
    assert L.rsx_calc_rsasa_noalloc(fake, MAX_ATOMS + 1, r, 2, probe, fake, 32, fake, None) == -3
    assert L.rsx_calc_rsasa_noalloc(fake, 100, r, 65, probe, fake, 32, fake, None) == -3
    assert L.rsx_calc_rsasa_noalloc(fake, 100, r, 2, probe, fake, 513, fake, None) == -3
    assert L.rsx_calc_rsasa_noalloc(None, 0, r, 2, probe, None, 32, None, None) == 0
    # a body of an unknown type has no dots and occludes nothing
    _say("case: no-alloc SASA with an out-of-range type")
    b = _cluster(9, 400, 2)
    bad = b.copy()
    bad[17, 3] = np.array([7], dtype=np.uint32).view(np.float32)[0]
    got = _noalloc(e, bad, radii, probe, dots)
    keep = np.arange(len(b)) != 17
    ref, _ = e.calc_rsasa(b[keep], radii, probe, dots)
    assert got[17] == 0
    assert np.array_equal(got[keep], ref)
    e.close()
