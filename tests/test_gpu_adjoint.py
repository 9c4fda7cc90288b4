"""Energy derivatives on the GPU (tcp_energy_adjoint) and the torch.autograd harness.

Tolerances:
  GPU adjoint vs the oracle (FP64):   U relative <= 1e-12; dU/dx, dU/dq, dF/dlambda per-particle
                                      relative <= 1e-10 (floor 1e-3 of the RMS); dU/dlambda <= 1e-10
  dU/dx vs the FP64 force path:       per-particle relative <= 1e-10
  central differences (h = 1e-6):     relative <= 1e-6 (floor 1 eV/A, 1 eV/e)
  lambda fit by force matching:       |lambda - 0.5| <= 1e-6
"""
import sys

import numpy as np
import pytest
import torch

from oracle import cpu
from tcp_allpair_amd import autograd as ag
from tcp_allpair_amd import forcematch
from tcp_allpair_amd.gpu import Context, TcpError
from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import (FP64, KELBG, MIXED, RECIP_EWALD, RECIP_NONE, RECIP_SPME,
                                    REG_EE)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_rows(a, ref):
    ref = ref.reshape(len(ref), -1)
    a = a.reshape(len(a), -1)
    nref = np.linalg.norm(ref, axis=1)
    floor = 1e-3 * np.sqrt(np.mean(nref ** 2))
    return np.linalg.norm(a - ref, axis=1) / np.maximum(nref, floor)


VARIANTS = [
    (RECIP_SPME, 4, 0), (RECIP_SPME, 6, KELBG), (RECIP_SPME, 4, KELBG | REG_EE),
    (RECIP_EWALD, 4, REG_EE), (RECIP_NONE, 4, 0),
]


@pytest.mark.parametrize("recip,order,flags", VARIANTS)
@pytest.mark.parametrize("charges", ["species", "random"])
def test_adjoint_vs_oracle(recip, order, flags, charges):
    say("case adjoint", recip, order, flags, charges)
    p, x, _ = make_plasma(150, 130, seed=11, K=(16, 16, 16), order=order, recip=recip, kmax=5,
                          lam=0.4)
    p = p.replace(flags=flags)
    q = None
    if charges == "random":
        q = np.where(np.arange(p.n) < p.ne, p.q[0], p.q[1]) * \
            (1.0 + 0.2 * np.random.default_rng(3).standard_normal(p.n))
    U0, dx0, dq0, dl0, fl0 = cpu.adjoint(p, x, q)
    with Context(p, FP64) as c:
        c.set_state(x=x)
        U, dx, dq, dl, fl = c.energy_adjoint(q)
    assert abs(U - U0) <= 1e-12 * abs(U0)
    assert rel_rows(dx, dx0).max() <= 1e-10
    assert rel_rows(dq, dq0).max() <= 1e-10
    assert abs(dl - dl0) <= 1e-10 * abs(dl0)
    assert rel_rows(fl, fl0).max() <= 1e-10


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_adjoint_matches_force_path(precision):
    """dU/dx = -F of the FP64 force path and U = eps (Ur + Uo) + Um, on either context."""
    say("case adjoint vs force path", precision)
    p, x, _ = make_plasma(500, 500, seed=4, K=(32, 32, 32), order=6, lam=0.45)
    p = p.replace(flags=KELBG)
    with Context(p, FP64) as c:
        c.set_state(x=x, f=np.zeros_like(x))
        ur = c.force_direct()
        um = c.force_recip()
        _, uo = c.ewald_self()[1:]
        f = c.get_state()[2]
    with Context(p, precision) as c:
        c.set_state(x=x)
        U, dx, _, _, _ = c.energy_adjoint()
        # the context's own forces are untouched
        assert not c.get_state()[2].any()
    assert rel_rows(dx, -f).max() <= 1e-10
    Uref = p.eps * (ur + uo) + um
    assert abs(U - Uref) <= 1e-12 * abs(Uref)


def test_adjoint_arguments():
    say("case adjoint arguments")
    p, x, _ = make_plasma(20, 20, seed=1, recip=RECIP_NONE)
    with Context(p, FP64) as c:
        c.set_state(x=x)
        with pytest.raises(TcpError):
            c.set_lambda(-1.0)
        with pytest.raises(TcpError):
            c.set_lambda(float("nan"))
        with pytest.raises(ValueError):
            c.energy_adjoint(np.ones(3))
        q = np.ones(p.n)
        q[3] = np.inf
        with pytest.raises(TcpError):
            c.energy_adjoint(q)
        c.set_lambda(0.3)
        assert c.energy_adjoint()[3] != 0.0


def test_autograd_gradcheck():
    say("case gradcheck")
    p, x, _ = make_plasma(24, 24, seed=8, recip=RECIP_EWALD, kmax=4, lam=0.4)
    p = p.replace(flags=KELBG | REG_EE)
    with Context(p, FP64) as c:
        xt = torch.from_numpy(x).requires_grad_()
        qt = ag.species_charges(p).requires_grad_()
        lt = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
        assert torch.autograd.gradcheck(lambda a, b, l: ag.energy(c, a, b, l), (xt, qt, lt),
                                        eps=1e-6, atol=1e-5, rtol=1e-6)
        x0, q0 = torch.from_numpy(x), ag.species_charges(p)
        assert torch.autograd.gradcheck(lambda l: ag.forces(c, x0, q0, l), (lt,), eps=1e-6,
                                        atol=1e-5, rtol=1e-6)
        with pytest.raises(RuntimeError):
            ag.forces(c, xt, q0, lt)


def test_differentiation_workload():
    """16384 particles, SPME 64^3: central differences at h = 1e-6, -F, and the lambda fit
    on 50 snapshots of a lambda = 0.5 trajectory with checkpointed snapshot losses."""
    say("case differentiation workload")
    res = forcematch.run()
    say(res)
    assert res["err_dudx_max"] <= 1e-6
    assert res["err_dudq_max"] <= 1e-6
    assert res["err_dudl"] <= 1e-6
    assert res["err_lambda"] <= 1e-6
