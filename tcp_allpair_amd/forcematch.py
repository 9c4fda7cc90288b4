"""Differentiation workload: derivative checks and a lambda fit by force matching.

16384 particles (8192 electrons, 8192 ions, charges -1/+1, masses 1/100) at
reduced density 1 (L = 25.3984 A), Kelbg regularisation of the electron-ion
and electron-electron pairs, SPME 64^3 at order 4, FP64.  lambda for the
derivative checks is drawn uniform in [0.3, 0.8] from the seed.  A lambda = 0.5
NVE trajectory generated by this library gives 50 snapshots and their forces;
lambda is then fitted by force matching from the drawn value.

    python -m tcp_allpair_amd.forcematch [--ne 8192] [--snapshots 50] [--stride 10]

prints one JSON line with the derivative errors, the fitted lambda and timings.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import autograd as ag
from .gpu import Context
from .inputs import make_plasma
from .params import FP64, KELBG, NOTEMP, REG_EE, RECIP_SPME, reduced_dt


def workload(ne=8192, seed=2031, K=64, order=4):
    """(params, x, v, lambda_drawn) of the differentiation workload."""
    p, x, v = make_plasma(ne, ne, seed=seed, K=(K, K, K), order=order, recip=RECIP_SPME,
                          lam=0.5)
    lam = float(np.random.Generator(np.random.PCG64(seed + 1)).uniform(0.3, 0.8))
    return p.replace(flags=KELBG | REG_EE), x, v, lam


def derivative_checks(c: Context, x, lam, h=1e-6, picks=(0, 1, 4097, 8192, 12000, 16383)):
    """Central differences of U (step h) against the adjoint at lambda = lam."""
    n = c.p.n
    picks = [i for i in picks if i < n]
    q = np.where(np.arange(n) < c.p.ne, c.p.q[0], c.p.q[1]).astype(float)
    c.set_lambda(lam)
    c.set_state(x=x)
    U, dudx, dudq, dudl, _ = c.energy_adjoint(q, with_dfdl=False)

    def energy(xx=x, qq=q, ll=lam):
        c.set_lambda(ll)
        c.set_state(x=xx)
        return c.energy_adjoint(qq, with_dfdl=False)[0]

    ex, eq = [], []
    for i in picks:
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd = (energy(xx=xp) - energy(xx=xm)) / (2 * h)
            ex.append(abs(fd - dudx[i, k]) / max(abs(dudx[i, k]), 1.0))
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (energy(qq=qp) - energy(qq=qm)) / (2 * h)
        eq.append(abs(fd - dudq[i]) / max(abs(dudq[i]), 1.0))
    fdl = (energy(ll=lam + h) - energy(ll=lam - h)) / (2 * h)
    c.set_lambda(lam)
    c.set_state(x=x)
    return dict(U=U, dudl=dudl, fd_dudl=fdl, err_dudl=abs(fdl - dudl) / abs(dudl),
                err_dudx_max=max(ex), err_dudq_max=max(eq), picks=list(picks))


def trajectory(c: Context, x, v, lam, nsnap, stride, dt):
    """nsnap snapshots, stride NVE steps apart, of a lambda = lam run from (x, v)."""
    c.set_lambda(lam)
    c.set_state(x=x)
    snaps = []
    for _ in range(nsnap):
        c.md_steps(stride, NOTEMP, dt, trace=False)
        snaps.append(c.get_state()[0])
    return snaps


def target_forces(c: Context, snaps, lam):
    c.set_lambda(lam)
    out = []
    for s in snaps:
        c.set_state(x=s)
        out.append(-c.energy_adjoint(with_dfdl=False)[1])
    return out


def run(ne=8192, nsnap=50, stride=10, seed=2031, lam_true=0.5, checks=True):
    p, x, v, lam_draw = workload(ne, seed)
    res = dict(workload="differentiation", n=p.n, box=p.box[0], K=p.K[0], order=p.order,
               flags="KELBG|REG_EE", lambda_drawn=lam_draw, lambda_true=lam_true,
               snapshots=nsnap, stride=stride)
    with Context(p, FP64) as c:
        t0 = time.perf_counter()
        if checks:
            res.update(derivative_checks(c, x, lam_draw))
        t1 = time.perf_counter()
        # forces before the first kick
        c.set_state(v=v, f=target_forces(c, [x], lam_true)[0])
        snaps = trajectory(c, x, v, lam_true, nsnap, stride, reduced_dt(0.005))
        targets = target_forces(c, snaps, lam_true)
        t2 = time.perf_counter()
        xs = [torch.from_numpy(s) for s in snaps]
        ts = [torch.from_numpy(t) for t in targets]
        lam_fit, hist = ag.fit_lambda(c, xs, ts, lam_draw)
        t3 = time.perf_counter()
    res.update(lambda_fit=lam_fit, fit_evaluations=len(hist), loss_final=hist[-1][1],
               err_lambda=abs(lam_fit - lam_true), seconds_checks=t1 - t0,
               seconds_trajectory=t2 - t1, seconds_fit=t3 - t2)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ne", type=int, default=8192)
    ap.add_argument("--snapshots", type=int, default=50)
    ap.add_argument("--stride", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2031)
    a = ap.parse_args()
    print(json.dumps(run(a.ne, a.snapshots, a.stride, a.seed)), flush=True)


if __name__ == "__main__":
    main()
