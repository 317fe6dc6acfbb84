"""The resaxslib-compatible C++ surface (include/resaxs/) driven by tests/cpp/compat_test,
a C++14 program written the way the reference's harnesses use the library
(test_resaxs/SAXSTest.cpp:276-330): factory, params.initialize, initialize,
calc_curve, recalc_curve + commit / revert, i_rsasa, host_debye, saxs_profile.

CPU: the device-free checks (device specs, errors, body bits, generators, profile
file format, scale and chi).  GPU: every engine path, its outputs checked here
against the oracle (and the compiled reference where it was built).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import padded_density_dots, rel_err
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "compat_test")
TOL = 1e-5
OH2 = 14


def _run(args, cwd, timeout):
    assert os.path.exists(BIN), "build tests/cpp/compat_test first (__graft_entry__.build())"
    r = subprocess.run([BIN] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    return r.stdout


def test_compat_surface_cpu(tmp_path):
    out = _run(["cpu"], str(tmp_path), 60)
    assert out.count("ok ") >= 20


def _avg_radius(types, radii):
    acc = np.float32(0)
    for t in types:
        acc = np.float32(acc + radii[t])
    return np.float32(acc / np.float32(len(types)))


def _rigid_move(rng, seg):
    c = seg.mean(axis=0)
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    th = np.deg2rad(rng.uniform(0, 5))
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    d = rng.normal(size=3)
    return (seg - c) @ R.T + c + d / np.linalg.norm(d) * rng.uniform(0, 1)


@pytest.mark.gpu
def test_compat_surface_gpu(tmp_path):
    from resaxs_amd import configs
    rng = np.random.default_rng(2024)
    n, nq = 3000, 101
    radius = (n / 0.045 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
    pos = configs.in_ball(rng, n, radius).astype(np.float32)
    types = rng.integers(0, 16, size=n).astype(np.uint32)
    q = np.linspace(0.0, 0.5, nq).astype(np.float32)
    bodies = po.bodies(pos, types)
    moves, cur = [], pos.copy()
    for _ in range(6):
        s = int(rng.integers(0, n - 200))
        new = _rigid_move(rng, cur[s:s + 200].astype(np.float64)).astype(np.float32)
        moves.append((s, new))
        cur = cur.copy()
        cur[s:s + 200] = new
    d = str(tmp_path)
    q.tofile(os.path.join(d, "q.bin"))
    bodies.tofile(os.path.join(d, "bodies.bin"))
    with open(os.path.join(d, "moves.bin"), "wb") as f:
        np.array([len(moves)], dtype=np.int32).tofile(f)
        for s, new in moves:
            np.array([s, len(new)], dtype=np.int32).tofile(f)
            po.bodies(new, types[s:s + len(new)]).tofile(f)
    print("case: compat C++ surface, 3000 atoms x 101 q, 6 moves of 200 atoms", flush=True)
    _run(["gpu", d], d, 240)
    ld = lambda name, dt: np.fromfile(os.path.join(d, name), dtype=dt)

    # default table: atomic factors at the bodies' mean radius, water row OH2 (ff_coef.hpp, saxs_alg:37-49)
    radii = po.atomic_radii()
    rm = _avg_radius(types, radii)
    table = po.atomic_ff(q, rm, 1.0)
    assert np.array_equal(ld("factors.bin", np.float32).reshape(nq, 16), table)
    assert ld("water_index.bin", np.int32)[0] == OH2

    def plain(p, tab=table):
        return po.debye(q, p, po.body_factors(p, types, q, tab))

    assert rel_err(ld("iq_plain.bin", np.float32), plain(pos)).max() < TOL
    # the saxs_hist_cl_gpu factory slot is the exact sum (the reference's class of that
    # name runs the pt_wf kernels); the distance-histogram engine is approximate (0.01 A bins)
    assert rel_err(ld("iq_hist_slot.bin", np.float32), plain(pos)).max() < TOL
    assert rel_err(ld("iq_hist.bin", np.float32), plain(pos)).max() < 1e-3
    assert rel_err(ld("iq_double.bin", np.float64), plain(pos)).max() < TOL

    # moves: the last committed state, the last (reverted) move, and the revert itself
    committed = ld("bodies_committed.bin", np.float32).reshape(n, 4)[:, :3]
    expect = pos.copy()
    for s, new in moves[:-1]:
        expect[s:s + len(new)] = new
    assert np.array_equal(committed, expect)
    assert rel_err(ld("iq_committed.bin", np.float32), plain(committed)).max() < TOL
    last = committed.copy()
    s, new = moves[-1]
    last[s:s + len(new)] = new
    assert rel_err(ld("iq_last_move.bin", np.float32), plain(last)).max() < TOL
    assert np.array_equal(ld("iq_reverted.bin", np.float32), ld("iq_committed.bin", np.float32))

    # hydration with the reference GPU set-up: 16 atomic radii, probe 1.8, 32 dodeca-icosa dots
    dots = np.stack([po.dodeca_icosa_dots(r, 1.8) for r in radii])
    R2 = ((radii + np.float32(1.8)) ** 2).astype(np.float32)
    counts = po.sasa_counts(committed, types, R2, dots)
    frac = (counts.astype(np.float32) / np.float32(32)).astype(np.float32)
    assert np.array_equal(ld("rsasa_hydrated.bin", np.float32), frac)
    hyd = po.debye(q, committed, po.body_factors(committed, types, q, table, OH2, 1.5, frac))
    assert rel_err(ld("iq_hydrated.bin", np.float32), hyd).max() < TOL
    table_c = po.atomic_ff(q, rm, 1.05)
    assert np.array_equal(ld("factors_expanded.bin", np.float32).reshape(nq, 16), table_c)
    hyd_c = po.debye(q, committed, po.body_factors(committed, types, q, table_c, OH2, 1.5, frac))
    assert rel_err(ld("iq_expanded.bin", np.float32), hyd_c).max() < TOL

    # host_debye with the default (c = 1) table
    t0 = po.atomic_ff(q)
    ih = ld("iq_host_debye.bin", np.float32)
    assert rel_err(ih, plain(pos, t0)).max() < TOL
    if po.ref_lib() is not None:
        assert rel_err(ih, po.ref_host_debye(q, pos, types, t0)).max() < TOL

    # i_rsasa: GPU set-up = dodeca-icosa counts / 32; CPU set-up = density rings (rsasa_cpu_alg)
    c0 = po.sasa_counts(pos, types, R2, dots)
    assert np.array_equal(ld("rsasa_gpu.bin", np.float32), (c0.astype(np.float32) / np.float32(32)))
    fc = ld("rsasa_cpu_setup.bin", np.float64)
    if po.ref_lib() is not None:
        assert np.array_equal(fc, po.ref_rsasa_cpu(bodies, 2.0, 1.8))
        dd, nd = padded_density_dots(2.0, radii, 1.8)
        cc = po.sasa_counts(pos, types, R2, dd)
        assert np.array_equal(fc, cc.astype(np.float64) / nd[types])
