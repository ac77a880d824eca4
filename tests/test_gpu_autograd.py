"""The discrete-adjoint rhs as a torch autograd backward (tests/torch_rhs.py):
vector-Jacobian products agree with central finite differences of the forward rhs,
and with the oracle's adjoint rhs."""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from magudi_amd import cases
from torch_rhs import MagudiRhs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["monopole", "walled_viscous"])
def test_vjp_matches_finite_differences(name):
    from magudi_amd.solver import Region
    c = (cases.acoustic_monopole(n=41, nsteps=1) if name == "monopole"
         else cases.walled_monopole(n=41, viscous=True))
    R = Region(c)
    rng = np.random.default_rng(8)
    q0 = c.target + 0.02 * rng.standard_normal((c.nv, c.npoints))
    q = torch.tensor(q0, dtype=torch.float64, requires_grad=True)
    v = torch.tensor(rng.standard_normal(q0.shape))
    out = MagudiRhs.apply(q, R, 0.7)
    (g,) = torch.autograd.grad(out, q, grad_outputs=v)
    dq = rng.standard_normal(q0.shape)
    pred = float((g.numpy() * dq).sum())
    errs = []
    for h in (1e-4, 1e-5):
        Rp = MagudiRhs.apply(torch.tensor(q0 + h * dq), R, 0.7).numpy()
        Rm = MagudiRhs.apply(torch.tensor(q0 - h * dq), R, 0.7).numpy()
        errs.append(abs(float((v.numpy() * (Rp - Rm)).sum()) / (2 * h) - pred) / abs(pred))
    assert min(errs) < 1e-7, errs
    # the same product from the oracle's adjoint rhs
    O = ob.Region(c)
    norm = O.grid(2)[0]
    ref = -norm * O.rhs(-1, q0, v.numpy() / norm, factor=0.0)
    assert np.abs(g.numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
