"""The torch.autograd harness on the CPU, with the oracle standing in for the GPU context.

The stand-in offers the four Context methods the harness calls (p, set_lambda,
set_state, energy_adjoint) from tcpo_adjoint, so this covers the harness logic:
gradients through both Functions, checkpointed snapshot losses and the LBFGS fit.
"""
import numpy as np
import pytest
import torch

from oracle import cpu
from tcp_allpair_amd import autograd as ag
from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import KELBG, RECIP_EWALD, REG_EE


class OracleContext:
    def __init__(self, p):
        self.p = p
        self.x = None
        self.calls = 0

    def set_lambda(self, lam):
        if not lam > 0:
            raise ValueError("lambda must be positive")
        self.p = self.p.replace(rs=float(lam))

    def set_state(self, x=None, v=None, f=None):
        if x is not None:
            self.x = np.array(x, dtype=np.float64)

    def energy_adjoint(self, q=None, with_dfdl=True):
        self.calls += 1
        U, dx, dq, dl, fl = cpu.adjoint(self.p, self.x, q)
        return U, dx, dq, dl, fl if with_dfdl else None


def small(seed=2, lam=0.45):
    p, x, v = make_plasma(12, 12, seed=seed, recip=RECIP_EWALD, kmax=3, lam=lam)
    return p.replace(flags=KELBG | REG_EE), x, v


def test_energy_gradients_by_gradcheck():
    p, x, _ = small()
    c = OracleContext(p)
    xt = torch.from_numpy(x).requires_grad_()
    qt = ag.species_charges(p).requires_grad_()
    lt = torch.tensor(0.45, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b, l: ag.energy(c, a, b, l), (xt, qt, lt),
                                    eps=1e-6, atol=1e-6, rtol=1e-6)


def test_forces_gradient_in_lambda_only():
    p, x, _ = small()
    c = OracleContext(p)
    x0 = torch.from_numpy(x)
    q0 = ag.species_charges(p)
    lt = torch.tensor(0.45, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda l: ag.forces(c, x0, q0, l), (lt,), eps=1e-6,
                                    atol=1e-7, rtol=1e-6)
    F = ag.forces(c, x0, q0, lt.detach())
    np.testing.assert_allclose(F.numpy(), -cpu.adjoint(c.p, x)[1], rtol=0, atol=0)
    with pytest.raises(RuntimeError):
        ag.forces(c, x0.clone().requires_grad_(), q0, lt)


def test_checkpointed_loss_matches_plain_and_recomputes():
    p, x, v = small()
    c = OracleContext(p)
    rng = np.random.default_rng(0)
    snaps = [torch.from_numpy(x + 0.05 * rng.standard_normal(x.shape)) for _ in range(3)]
    q = ag.species_charges(p)
    c.set_lambda(0.5)
    targets = []
    for s in snaps:
        c.set_state(x=s.numpy())
        targets.append(torch.from_numpy(-c.energy_adjoint()[1]))
    grads, calls = [], []
    for use_ck in (False, True):
        lam = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
        c.calls = 0
        loss = ag.force_matching_loss(c, snaps, targets, q, lam, use_checkpoint=use_ck)
        loss.backward()
        grads.append((float(loss.detach()), float(lam.grad)))
        calls.append(c.calls)
    assert grads[0] == grads[1]
    # checkpointing re-runs each snapshot's forward in the backward pass
    assert calls == [3, 6]


def test_fit_lambda_recovers_generating_value():
    p, x, _ = small(lam=0.5)
    c = OracleContext(p)
    rng = np.random.default_rng(1)
    snaps = [torch.from_numpy(x + 0.05 * rng.standard_normal(x.shape)) for _ in range(4)]
    c.set_lambda(0.5)
    targets = []
    for s in snaps:
        c.set_state(x=s.numpy())
        targets.append(torch.from_numpy(-c.energy_adjoint()[1]))
    lam, hist = ag.fit_lambda(c, snaps, targets, 0.72)
    assert abs(lam - 0.5) <= 1e-7
    assert hist[-1][1] <= 1e-12
