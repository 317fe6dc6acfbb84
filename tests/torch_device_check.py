"""The torch wrappers (resaxs_amd/torch_debye.py) with positions on the GPU, checked against
the same wrappers on CPU tensors and against the FP64 oracle; run by
tests/test_gpu_fit.py::test_autograd_wrappers_on_device as a child process.  torch is
imported before the library is loaded, so both share torch's HIP runtime.  Prints
"torch device check: ok" and exits 0 when every check holds."""
import os
import sys

import torch  # noqa: F401  (first: the library binds to the HIP runtime torch loaded)

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

from helpers import engine_for, rel_err  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def main():
    from resaxs_amd import configs
    from resaxs_amd.torch_debye import chi2_loss, debye_intensity, refine
    assert torch.cuda.is_available(), "torch sees no GPU"
    rng = np.random.default_rng(9)
    n = 1500
    q = np.linspace(0.01, 0.5, 64).astype(np.float32)
    pos = configs.in_ellipsoid(rng, n, configs.ellipsoid_axes(n)).astype(np.float32)
    types = configs.seeded_types(rng, n)
    table = configs.factor_table(rng, q)
    e = engine_for(q, table, n=n)
    wv = torch.linspace(-1.0, 1.0, len(q), dtype=torch.float64)
    xc = torch.tensor(pos, requires_grad=True)
    Ic = debye_intensity(xc, e, types)
    (Ic * wv).sum().backward()
    xd = torch.tensor(pos, device="cuda", requires_grad=True)
    Id = debye_intensity(xd, e, types)
    assert Id.is_cuda
    (Id * wv.cuda()).sum().backward()
    I_ref = po.debye(q, pos, po.body_factors(pos, types, q, table))
    err = rel_err(Id.detach().cpu().numpy(), I_ref).max()
    print(f"I(q) on the device vs FP64 oracle: {err:.3e}", flush=True)
    assert err < 1e-5
    assert rel_err(Id.detach().cpu().numpy(), Ic.detach().numpy()).max() < 1e-5
    gerr = np.linalg.norm(xd.grad.cpu().numpy() - xc.grad.numpy()) / np.linalg.norm(xc.grad.numpy())
    print(f"VJP device vs host: {gerr:.3e}", flush=True)
    assert gerr < 1e-6
    i_exp = (I_ref * (1 + 0.01 * rng.normal(size=len(q)))).astype(np.float32)
    sigma = (0.01 * I_ref).astype(np.float32)
    start = (pos + rng.normal(scale=0.3, size=pos.shape)).astype(np.float32)
    yc = torch.tensor(start, requires_grad=True)
    lc = chi2_loss(yc, e, types, i_exp, sigma)
    lc.backward()
    yd = torch.tensor(start, device="cuda", requires_grad=True)
    ld = chi2_loss(yd, e, types, i_exp, sigma)
    ld.backward()
    print(f"chi2 host {float(lc):.9e} device {float(ld):.9e}", flush=True)
    assert abs(float(ld) / float(lc) - 1) < 1e-9
    assert np.linalg.norm(yd.grad.cpu().numpy() - yc.grad.numpy()) / np.linalg.norm(yc.grad.numpy()) < 1e-6
    _, h_c = refine(e, types, start, i_exp, sigma, steps=15, lr=0.01)
    _, h_d = refine(e, types, start, i_exp, sigma, steps=15, lr=0.01, device="cuda")
    print(f"chi2 after 15 Adam steps: host {h_c[-1]:.5e}, device {h_d[-1]:.5e}", flush=True)
    assert h_d[-1] < h_d[0] and abs(h_d[-1] / h_c[-1] - 1) < 1e-3
    e.close()
    print("torch device check: ok", flush=True)


if __name__ == "__main__":
    main()
