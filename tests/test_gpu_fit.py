"""Differentiable fit at the fit workload's full size (10,000 atoms x 200 q in
[0.01, 0.5], target from the reference conformation with 1 % Gaussian noise,
start displaced by N(0, 0.5 A)):

  - the adjoint gradient against central finite differences (h = 1e-3 A) of the
    FP64 oracle's chi^2 (scale re-optimised at each point) on 64 random atoms x 3
    coordinates: relative error of the 192-vector <= 1e-4;
  - the torch.autograd wrappers (DebyeIntensity, Chi2Loss) give the adjoint's
    numbers through backward();
  - 100 gradient-descent (Adam) refinement steps bring chi^2 to the noise floor.
"""
import sys

import numpy as np
import pytest

from helpers import engine_for, rel_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _target(w, e):
    I_ref = e.calc_curve(w.bodies)
    rng = w.noise_rng
    i_exp = (I_ref * (1.0 + 0.01 * rng.normal(size=len(w.q)))).astype(np.float32)
    sigma = (0.01 * I_ref).astype(np.float32)
    return i_exp, sigma


@pytest.mark.parametrize("fp32", [True, False])
def test_fit_gradient_vs_fd_full_size(oracle, fp32):
    from resaxs_amd import configs
    w = configs.fit_10000()
    _say(f"case: chi2 gradient at full size, 10000 atoms x 200 q, FD on 64 atoms x 3, "
         f"{'FP32' if fp32 else 'FP64'} adjoint")
    e = engine_for(w.q, w.table, n=len(w.pos))
    e.set_adjoint_tuning(fp32)
    i_exp, sigma = _target(w, e)
    chi2, scale, grad, iq = e.chi2_grad(w.start_bodies, i_exp, sigma)
    rng = np.random.default_rng(64)
    atoms = rng.choice(len(w.pos), size=64, replace=False)
    comps = (3 * atoms[:, None] + np.arange(3)[None, :]).reshape(-1)
    xyz = w.start.astype(np.float64)
    F = oracle.body_factors(w.start, w.types, w.q, w.table)
    fd = oracle.chi2_fd(w.q, xyz, F, i_exp, sigma, comps, h=1e-3)
    g = grad.reshape(-1)[comps].astype(np.float64)
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    mx = np.max(np.abs(g - fd)) / np.max(np.abs(fd))
    _say(f"  chi2 {chi2:.6e}, |g - fd| / |fd| = {rel:.3e}, max |g - fd| / max |fd| = {mx:.3e}")
    assert rel < 1e-4 and mx < 1e-4
    # forward in the same call against the FP64 oracle
    assert rel_err(iq, oracle.debye(w.q, w.start, F)).max() < 1e-5


def test_autograd_wrappers(oracle):
    import torch

    from resaxs_amd import configs
    from resaxs_amd.torch_debye import chi2_loss, debye_intensity
    rng = np.random.default_rng(8)
    n = 1500
    q = np.linspace(0.01, 0.5, 64).astype(np.float32)
    pos = configs.in_ellipsoid(rng, n, configs.ellipsoid_axes(n)).astype(np.float32)
    types = configs.seeded_types(rng, n)
    table = configs.factor_table(rng, q)
    _say("case: autograd wrappers, 1500 atoms x 64 q")
    e = engine_for(q, table, n=n)
    x = torch.tensor(pos, requires_grad=True)
    I = debye_intensity(x, e, types)
    wv = torch.linspace(-1.0, 1.0, len(q), dtype=torch.float64)
    (I * wv).sum().backward()
    g_vjp = e.vjp(configs.bodies_of(pos, types), wv.numpy())
    assert np.array_equal(x.grad.numpy(), g_vjp)
    F = oracle.body_factors(pos, types, q, table)
    assert rel_err(I.detach().numpy(), oracle.debye(q, pos, F)).max() < 1e-5
    # chi^2 loss: backward returns the fused adjoint's gradient (times the upstream grad)
    I_ref = oracle.debye(q, pos, F)
    i_exp = (I_ref * (1 + 0.01 * rng.normal(size=len(q)))).astype(np.float32)
    sigma = (0.01 * I_ref).astype(np.float32)
    start = (pos + rng.normal(scale=0.3, size=pos.shape)).astype(np.float32)
    y = torch.tensor(start, requires_grad=True)
    loss = chi2_loss(y, e, types, i_exp, sigma)
    (2.0 * loss).backward()
    chi2, scale, grad, _ = e.chi2_grad(configs.bodies_of(start, types), i_exp, sigma)
    assert float(loss) == chi2
    assert np.array_equal(y.grad.numpy(), 2.0 * grad)
    g_ref, _, _, _ = oracle.chi2_grad(q, start.astype(np.float64), oracle.body_factors(start, types, q, table),
                                      i_exp, sigma)
    assert np.linalg.norm(grad - g_ref) / np.linalg.norm(g_ref) < 1e-4


def test_ensemble_autograd(oracle):
    """ensemble_intensity: per-frame I(q) of a ragged ensemble (an empty frame included) from
    one batched launch against the oracle; backward is the batched VJP, equal to
    Engine.batch_vjp and, frame by frame, to the single-structure VJP; a backward that runs
    after another ensemble's forward on the same engine gets its own frames' gradient."""
    import torch

    from resaxs_amd import configs
    from resaxs_amd.torch_debye import ensemble_intensity
    rng = np.random.default_rng(9)
    q = np.linspace(0.01, 0.5, 48).astype(np.float32)
    table = configs.factor_table(rng, q)
    sizes = [300, 0, 500, 120]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    pos = np.concatenate([configs.in_ellipsoid(rng, m, configs.ellipsoid_axes(m)) if m else np.zeros((0, 3))
                          for m in sizes]).astype(np.float32)
    types = np.concatenate([configs.seeded_types(rng, m) for m in sizes]).astype(np.uint32)
    _say(f"case: ensemble autograd, frames {sizes} x {len(q)} q")
    e = engine_for(q, table)
    W = rng.normal(size=(len(sizes), len(q)))
    x = torch.tensor(pos, requires_grad=True)
    I = ensemble_intensity(x, e, types, off)
    assert I.shape == (len(sizes), len(q))
    for f, m in enumerate(sizes):
        a, b = off[f], off[f + 1]
        if m == 0:
            assert np.all(I[f].detach().numpy() == 0.0)
            continue
        ref = oracle.debye(q, pos[a:b], oracle.body_factors(pos[a:b], types[a:b], q, table))
        assert rel_err(I[f].detach().numpy(), ref).max() < 1e-5, f
    # another ensemble's forward on the same engine before this backward
    y = torch.tensor(pos * np.float32(0.97), requires_grad=True)
    ensemble_intensity(y, e, types, off)
    (I * torch.as_tensor(W)).sum().backward()
    e.batch_set_frames(configs.bodies_of(pos, types), off)
    assert np.array_equal(x.grad.numpy(), e.batch_vjp(W))
    for f, m in enumerate(sizes):
        if m < 2:
            continue
        a, b = off[f], off[f + 1]
        s = engine_for(q, table, n=m)
        g1 = s.vjp(configs.bodies_of(pos[a:b], types[a:b]), W[f]).astype(np.float64)
        s.close()
        d = np.linalg.norm(x.grad.numpy()[a:b] - g1) / np.linalg.norm(g1)
        _say(f"  frame {f} ({m} atoms): |g - g_single| / |g_single| = {d:.2e}")
        assert d < 1e-5
    e.close()


def test_refinement_100_steps():
    from resaxs_amd import configs
    from resaxs_amd.torch_debye import refine
    w = configs.fit_10000()
    _say("case: 100 Adam steps on chi2, 10000 atoms x 200 q")
    e = engine_for(w.q, w.table, n=len(w.pos))
    i_exp, sigma = _target(w, e)
    pos, hist = refine(e, w.types, w.start, i_exp, sigma, steps=100, lr=0.01)
    _say(f"  chi2: start {hist[0]:.4e}, after 10 {hist[9]:.4e}, after 100 {hist[-1]:.4e}")
    assert np.all(np.isfinite(hist)) and np.all(np.isfinite(pos))
    # the start (0.5 A random displacements) already fits to chi^2 / n_q ~ 2; 1 % noise puts
    # the floor near chi^2 = n_q: the refinement must get there
    assert hist[-1] < hist[0] and hist[-1] < 1.25 * len(w.q)


def test_device_gradients_equal_host(oracle):
    """rsx_debye_vjp_dev_async and rsx_chi2_grad_dev_async on device-resident bodies against
    the host entry points: the same layout gives the same gradient, and on moved positions
    the device path matches a host call on those positions."""
    from resaxs_amd import configs
    rng = np.random.default_rng(81)
    n = 2000
    q = np.linspace(0.01, 0.5, 64).astype(np.float32)
    pos = configs.in_ellipsoid(rng, n, configs.ellipsoid_axes(n)).astype(np.float32)
    types = configs.seeded_types(rng, n)
    table = configs.factor_table(rng, q)
    _say("case: device-resident VJP and chi2 gradient, 2000 atoms x 64 q")
    e = engine_for(q, table, n=n)
    b = configs.bodies_of(pos, types)
    I_ref = oracle.debye(q, pos, oracle.body_factors(pos, types, q, table))
    i_exp = (I_ref * (1 + 0.01 * rng.normal(size=len(q)))).astype(np.float32)
    sigma = (0.01 * I_ref).astype(np.float32)
    chi2_h, scale_h, grad_h, iq_h = e.chi2_grad(b, i_exp, sigma)   # host call: gradient layout
    w = np.linspace(-1.0, 1.0, len(q))
    g_vjp_h = e.vjp(b, w)
    bufs = {k: e.dev_alloc(v) for k, v in (("b", b.nbytes), ("w", w.nbytes), ("g", n * 12), ("ie", i_exp.nbytes),
                                           ("sg", sigma.nbytes), ("out", 16), ("iq", len(q) * 8))}
    try:
        e.h2d(bufs["b"], b)
        e.h2d(bufs["w"], w)
        e.h2d(bufs["ie"], i_exp)
        e.h2d(bufs["sg"], sigma)
        e.vjp_dev_async(bufs["b"], n, bufs["w"], bufs["g"])
        e.synchronize()
        g = e.d2h(np.zeros((n, 3), np.float32), bufs["g"])
        assert np.array_equal(g, g_vjp_h)
        e.chi2_grad_dev_async(bufs["b"], n, bufs["ie"], bufs["sg"], bufs["out"], bufs["g"], bufs["iq"])
        e.synchronize()
        out = e.d2h(np.zeros(2), bufs["out"])
        g = e.d2h(np.zeros((n, 3), np.float32), bufs["g"])
        iq = e.d2h(np.zeros(len(q)), bufs["iq"])
        assert abs(out[0] / chi2_h - 1) < 1e-9 and abs(out[1] / scale_h - 1) < 1e-12
        assert np.all(np.abs(iq - iq_h) <= 1e-12 * np.abs(iq_h))
        assert np.linalg.norm(g - grad_h) / np.linalg.norm(grad_h) < 1e-6
        # moved positions: the device path against host calls on them
        moved = (pos + rng.uniform(-0.2, 0.2, pos.shape)).astype(np.float32)
        bm = configs.bodies_of(0.98 * moved, types)
        e.h2d(bufs["b"], bm)
        e.chi2_grad_dev_async(bufs["b"], n, bufs["ie"], bufs["sg"], bufs["out"], bufs["g"], None)
        e.synchronize()
        out = e.d2h(np.zeros(2), bufs["out"])
        g = e.d2h(np.zeros((n, 3), np.float32), bufs["g"])
        chi2_m, scale_m, grad_m, _ = e.chi2_grad(bm, i_exp, sigma)
        assert abs(out[0] / chi2_m - 1) < 1e-6
        assert np.linalg.norm(g - grad_m) / np.linalg.norm(grad_m) < 1e-5
    finally:
        for p in bufs.values():
            e.dev_free(p)
    e.close()


def test_device_vjp_type_outside_table():
    """A device-resident VJP on bodies that hold a type outside the factor table (the gather
    gives its record no class): every adjoint form reads no table entry for it and adds nothing
    for its pairs, rsx_synchronize reports the type, and the next clean call gives the host
    gradient again."""
    from resaxs_amd import RsxError, configs
    rng = np.random.default_rng(82)
    n = 600
    q = np.linspace(0.01, 0.5, 48).astype(np.float32)
    pos = configs.in_ellipsoid(rng, n, configs.ellipsoid_axes(n)).astype(np.float32)
    types = configs.seeded_types(rng, n)
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=n)
    b = configs.bodies_of(pos, types)
    bad = b.copy()
    bad[7, 3] = np.array([table.shape[1] + 3], dtype=np.uint32).view(np.float32)[0]
    w = np.linspace(-1.0, 1.0, len(q))
    for form in ("pair", "per_lane", "fp64"):
        _say(f"case: device-resident VJP with a type outside the table, {n} atoms x {len(q)} q, {form}")
        e.set_adjoint_tuning(form != "fp64", 0, 0, form == "per_lane")
        g_host = e.vjp(b, w)
        assert e.adjoint_layout()["form"] == form
        pb, pw, pg = e.dev_alloc(b.nbytes), e.dev_alloc(w.nbytes), e.dev_alloc(n * 12)
        try:
            e.h2d(pw, w)
            e.h2d(pb, bad)
            e.vjp_dev_async(pb, n, pw, pg)
            with pytest.raises(RsxError) as ei:
                e.synchronize()
            assert ei.value.code == -3 and "type" in str(ei.value)
            e.h2d(pb, b)
            e.vjp_dev_async(pb, n, pw, pg)
            e.synchronize()
            g = e.d2h(np.zeros((n, 3), np.float32), pg)
            assert np.array_equal(g, g_host)
        finally:
            for p in (pb, pw, pg):
                e.dev_free(p)
    e.close()


def test_autograd_wrappers_on_device():
    """The torch wrappers with positions on the GPU (device-resident entry points, no host
    copies after the first layout) against the same wrappers on the CPU, and a short
    refinement on the device.  Runs tests/torch_device_check.py in a child process that
    imports torch before the library, as a torch program does: the library then shares the
    HIP runtime torch ships (in this process the library was loaded first)."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    _say("case: autograd wrappers on the device (child process), 1500 atoms x 64 q")
    r = subprocess.run([sys.executable, os.path.join(here, "torch_device_check.py")], capture_output=True, text=True,
                       timeout=180)
    _say(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch device check: ok" in r.stdout
