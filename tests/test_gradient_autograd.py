"""dE/dJ through torch.autograd (test harness only).

`DmrgEnergy` is a torch.autograd.Function whose forward converges the ground
state of a Heisenberg chain with per-bond couplings J on the engine and keeps
a device copy of the converged driver (MPS and environments) as the
checkpoint; its backward evaluates the engine's adjoint gradient
(sadmrg_dmrg_coupling_gradient) on that copy.  The engine is the only thing
that computes; torch supplies the graph.
"""
import numpy as np
import pytest

import sadmrg_testlib as tl

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _chain(L, J):
    hmap = np.zeros((L, L), dtype=np.int32)
    coup = np.zeros(L + 1)
    for i in range(L - 1):
        hmap[i, i + 1] = hmap[i + 1, i] = i + 1
        coup[i + 1] = J[i]
    return hmap, coup


class DmrgEnergy(torch.autograd.Function):
    """E(J) of the open Heisenberg chain, J = per-bond couplings (float64)."""

    @staticmethod
    def forward(fctx, J, engine, D, sweeps):
        sad, ctx = engine
        L = J.numel() + 1
        hmap, coup = _chain(L, J.detach().cpu().numpy())
        sad.lib().sadmrg_rng_seed(42)
        d = sad.DMRG(ctx, 1, 0, 1, L, D)
        d.set_couplings(hmap, coup)
        d.set_max_dcut(D)
        d.idmrg()
        for _ in range(sweeps):
            d.sweep()
        fctx.checkpoint = d.clone()  # device-resident state for the backward pass
        return torch.tensor(d.energy(), dtype=torch.float64)

    @staticmethod
    def backward(fctx, grad_e):
        g, _ = fctx.checkpoint.coupling_gradient()  # dE/dJ_c for c = 0..ns; bond i has c = i + 1
        nb = fctx.checkpoint.ns - 1
        fctx.checkpoint = None
        return grad_e * torch.from_numpy(np.asarray(g[1:nb + 1], dtype=np.float64)), None, None, None


def test_autograd_gradient_matches_finite_differences(ctx, sad):
    """torch.autograd.grad of E(J) equals central finite differences
    (h = 1e-5, each side re-converged) and the exact bond energies of the
    chain (exact diagonalisation, L = 10)."""
    L, D, sweeps = 10, 64, 3
    rng = np.random.default_rng(11)
    J = torch.tensor(rng.uniform(0.5, 1.5, L - 1), dtype=torch.float64, requires_grad=True)
    E = DmrgEnergy.apply(J, (sad, ctx), D, sweeps)
    (g,) = torch.autograd.grad(E, J)
    hmap, coup = _chain(L, J.detach().numpy())
    e_ed, bond = tl.heisenberg_bond_expectations(hmap, coup)
    assert abs(float(E.detach()) - e_ed) <= 1e-9 * abs(e_ed)
    for i in range(L - 1):
        assert abs(float(g[i]) - bond[(i, i + 1)]) <= 1e-7, (i, float(g[i]), bond[(i, i + 1)])
    h = 1e-5
    for i in (0, 4, L - 2):
        Jp = J.detach().clone()
        Jm = J.detach().clone()
        Jp[i] += h
        Jm[i] -= h
        ep = float(DmrgEnergy.apply(Jp, (sad, ctx), D, sweeps))
        em = float(DmrgEnergy.apply(Jm, (sad, ctx), D, sweeps))
        fd = (ep - em) / (2 * h)
        assert abs(float(g[i]) - fd) <= 1e-6, (i, float(g[i]), fd)
    # a scalar loss through the graph: d(E^2)/dJ = 2 E dE/dJ
    J2 = J.detach().clone().requires_grad_(True)
    loss = DmrgEnergy.apply(J2, (sad, ctx), D, sweeps) ** 2
    loss.backward()
    assert torch.allclose(J2.grad, 2 * E.detach() * g, rtol=1e-9, atol=1e-12)
