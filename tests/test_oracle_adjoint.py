"""Pins of the oracle's energy derivatives (tcpo_adjoint, no reference counterpart).

U = eps (Ur + Uo) + Um is tied to the reference-pinned pieces (force_allpair,
ewald_self, spme / ewald_kspace) at the species charges; the derivatives are
pinned by finite differences of that U and of the pinned forces, by U being
a quadratic form in the charges (central differences are exact up to rounding,
and Euler's identity sum_i q_i dU/dq_i = 2U holds), and by dU/dx = -F.
"""
import numpy as np
import pytest

from oracle import cpu
from tcp_allpair_amd.inputs import make_plasma
from tcp_allpair_amd.params import KELBG, RECIP_EWALD, RECIP_NONE, RECIP_SPME, REG_EE

VARIANTS = [
    (RECIP_SPME, 0), (RECIP_EWALD, 0), (RECIP_NONE, 0),
    (RECIP_SPME, KELBG), (RECIP_SPME, REG_EE), (RECIP_EWALD, KELBG | REG_EE),
]


def system(recip, flags, seed=5, ne=24, ni=20):
    p, x, _ = make_plasma(ne, ni, seed=seed, K=(16, 16, 16), order=6, recip=recip, kmax=6,
                          lam=0.4)
    return p.replace(flags=flags), x


def total_energy(p, x, q=None):
    return cpu.adjoint(p, x, q)[0]


def pinned_forces(p, x):
    f, ur, _ = cpu.force_allpair(p, x)
    um, fr = 0.0, np.zeros_like(x)
    if p.recip == RECIP_SPME:
        fr, um = cpu.spme(p, x)
    elif p.recip == RECIP_EWALD:
        fr, um = cpu.ewald_kspace(p, x)
    _, _, uo = cpu.ewald_self(p)
    return f + fr, p.eps * (ur + uo) + um


@pytest.mark.parametrize("recip,flags", VARIANTS)
def test_energy_and_position_gradient_match_pinned_forces(recip, flags):
    p, x = system(recip, flags)
    U, dudx, dudq, _, _ = cpu.adjoint(p, x)
    F, Uref = pinned_forces(p, x)
    assert abs(U - Uref) <= 1e-12 * abs(Uref) + 1e-12
    np.testing.assert_allclose(dudx, -F, rtol=1e-11, atol=1e-11 * np.abs(F).max())


@pytest.mark.parametrize("recip,flags", VARIANTS)
def test_charge_gradient_by_central_differences(recip, flags):
    p, x = system(recip, flags)
    q = np.where(np.arange(p.n) < p.ne, p.q[0], p.q[1]).astype(float)
    q *= 1.0 + 0.1 * np.random.default_rng(1).standard_normal(p.n)
    U, _, dudq, _, _ = cpu.adjoint(p, x, q)
    # U is a quadratic form in q: Euler's identity for degree-2 homogeneity
    assert abs(q @ dudq - 2.0 * U) <= 1e-10 * np.abs(q * dudq).sum()
    h = 1e-3
    for i in (0, 5, p.ne - 1, p.ne, p.n - 1):
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (total_energy(p, x, qp) - total_energy(p, x, qm)) / (2 * h)
        assert abs(fd - dudq[i]) <= 1e-8 * (abs(dudq[i]) + 1.0), (i, fd, dudq[i])


def test_charge_gradient_species_charges_match_default():
    p, x = system(RECIP_SPME, 0)
    q = np.where(np.arange(p.n) < p.ne, p.q[0], p.q[1]).astype(float)
    a = cpu.adjoint(p, x)
    b = cpu.adjoint(p, x, q)
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[2], b[2])


@pytest.mark.parametrize("flags", [0, KELBG, REG_EE, KELBG | REG_EE])
def test_lambda_derivatives_by_central_differences(flags):
    p, x = system(RECIP_SPME, flags)
    lam = p.rs
    _, _, _, dudl, dfdl = cpu.adjoint(p, x)
    h = 1e-5 * lam
    pp, pm = p.replace(rs=lam + h), p.replace(rs=lam - h)
    fd = (total_energy(pp, x) - total_energy(pm, x)) / (2 * h)
    assert abs(fd - dudl) <= 1e-7 * abs(dudl), (fd, dudl)
    Fp, _ = pinned_forces(pp, x)
    Fm, _ = pinned_forces(pm, x)
    fdF = (Fp - Fm) / (2 * h)
    np.testing.assert_allclose(dfdl, fdF, rtol=1e-5, atol=1e-6 * np.abs(fdF).max())
    # lambda only enters the regularised pairs: sum of the reactions vanishes
    assert np.abs(dfdl.sum(axis=0)).max() <= 1e-10 * np.abs(dfdl).max()


def test_lambda_derivative_vanishes_without_regularised_pairs():
    p, x = system(RECIP_SPME, 0, ne=30, ni=0)
    _, _, _, dudl, dfdl = cpu.adjoint(p, x)
    assert dudl == 0.0
    assert not dfdl.any()
