"""The float64 torch autograd restatement agrees with the oracle's analytic
FP64 adjoint (both are then the references of the GPU adjoint)."""
import numpy as np
import torch

from daxtoolkit_amd import fields
from mc_autograd import torch_reference_loss
from oracle import oracle as orc


def test_torch_autograd_matches_oracle_adjoint():
    dims = (20, 17, 15)
    p = fields.gaussian_mixture_params(21, dims, n_gauss=3, sigma=(3.0, 6.0))
    f = fields.gaussian_mixture_numpy(dims, p)
    iso, org, sp = 0.45, (1.0, -2.0, 0.5), (0.5, 1.0, 1.5)
    r = orc.generate(f, iso, origin=org, spacing=sp)
    U = r["n_pts"]
    assert U > 50
    rng = np.random.default_rng(0)
    dldx = rng.standard_normal((U, 3)).astype(np.float32)
    dldn = rng.standard_normal((U, 3)).astype(np.float32)
    f64 = torch.tensor(f, dtype=torch.float64, requires_grad=True)
    L = torch_reference_loss(f64, r["first"], r["second"], iso, dldx, dldn, org, sp)
    L.backward()
    ref = orc.backward_f64(f, iso, r["first"], r["second"], dldx, dldn, origin=org, spacing=sp)
    got = f64.grad.numpy().reshape(-1)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-9 * scale
