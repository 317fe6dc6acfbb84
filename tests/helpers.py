"""Shared helpers for the test-suite (engine set-up, oracle factors, error metrics)."""
import numpy as np

from oracle import pyoracle as po


def rel_err(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)


def engine_for(q, table, water_index=-1, w=0.0, n=None, radii=None, probe=None, n_half=None, device=0, **tune):
    from resaxs_amd import Engine, RSX_DOTS_FIBONACCI
    e = Engine(device)
    e.set_q(q)
    if tune:
        e.set_tuning(**tune)
    e.set_factors(table, water_index)
    if radii is not None:
        e.set_sasa_params(radii, probe, RSX_DOTS_FIBONACCI, n_half)
    e.set_water_weight(w)
    if n is not None:
        e.initialize(n)
    return e


def fib_dot_table(radii, probe, n_half):
    """[n_types, 2 n_half, 4] templates, type order of radii."""
    return np.stack([po.fibonacci_dots(n_half, r, probe) for r in radii])


def r2_table(radii, probe):
    r = np.asarray(radii, dtype=np.float32) + np.float32(probe)
    return (r * r).astype(np.float32)


def oracle_iq(q, pos, types, table, water_index=-1, w=0.0, counts=None, n_dots=None):
    s = None
    if counts is not None:
        s = (counts.astype(np.float32) / np.float32(n_dots)).astype(np.float32)
    F = po.body_factors(pos, types, q, table, water_index, w, s)
    return po.debye(q, pos, F)


def padded_density_dots(density, radii, probe):
    """The reference's per-type density dot spheres (rsasa_algorithm.cpp:58-86, from the
    compiled reference) padded to one length with dots at the atom centre; such a dot is
    always buried by its own atom, so exposed counts are unchanged.  Returns
    (dots [n_types, D, 4], real dot count per type)."""
    dl = [po.ref_density_dots(density, r, probe) for r in radii]
    D = max(len(d) for d in dl)
    dots = np.zeros((len(radii), D, 4), dtype=np.float32)
    for t, d in enumerate(dl):
        dots[t, :len(d)] = d
        dots[t, len(d):, 3] = radii[t]
    return dots, np.array([len(d) for d in dl])
