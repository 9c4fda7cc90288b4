"""Write the golden fixtures of the reference's own routines.

Runs oracle/_ref/ref_harness -- data.f90, force.f90 and vverlet.f90 compiled
from the reference sources by oracle/Makefile -- on seeded inputs and stores
inputs and full-precision REAL8 outputs as compressed .npz (no pickles).

    python tests/golden/make_golden.py [case ...]   # needs oracle/_ref/ref_harness

Cases (reference routine, file:line):
  force_tiny   Force_direct_self + EWALD_SELF, 24 e + 20 i        force.f90:4,52
  force_c0     the same at 1000 e + 1000 i, L = 12.599 A          force.f90:52
  force_bins   1000+1000 with narrow RDF bins (bins past 501 skipped)
  force_aniso  orthorhombic box, 150 e + 90 i, rc < L/2           force.f90:18-24
  psi          EWALD_SELF's PME branch, K = (8, 6, 10)            force.f90:27-42
  ext          Force_direct_ext / Force_direct_sum, own 40+30     force.f90:183,346
  rigid        REMOVE_RIGIDMTN                                     force.f90:974
  vv1_notemp / vv1_stoch       VVerletNotemp1 / VVerletStoch1     vverlet.f90:132,9
  vv2_notemp / vv2_berend / vv2_isokin                             vverlet.f90:174,227,301
  traj_nve     40 steps of main.f90:87-153 without the reciprocal part, 250+250
  traj_isokin  20 steps, isokinetic thermostat
  traj_c0      200 NVE steps at dt = 0.005 tau* of 1000 e + 1000 i (C0's size,
               step and run length) without the reciprocal part
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import refharness as rh  # noqa: E402
from tcp_allpair_amd.inputs import make_plasma  # noqa: E402
from tcp_allpair_amd.params import BEREND, ISOKIN, NOTEMP, STOCH, reduced_dt  # noqa: E402


def params_dict(p):
    return dict(ne=p.ne, ni=p.ni, box=np.array(p.box), q=np.array(p.q), xm=np.array(p.xm), rc=p.rc,
                rs=p.rs, dx=np.array(p.dx), Te=p.Te, Ti=p.Ti, K=np.array(p.K), tau=p.tau, fric=p.fric)


# outputs each mode leaves unchanged (the inputs already hold them)
UNCHANGED = {rh.MODE_FORCE: ("x", "v", "vdf"), rh.MODE_PSI: ("x", "v", "f", "rdf", "vdf"),
             rh.MODE_EXT: ("x", "v", "vdf"), rh.MODE_RIGID: ("x", "v", "rdf", "vdf"),
             rh.MODE_VV1: ("f", "rdf", "vdf"), rh.MODE_VV2: ("x", "rdf"), rh.MODE_TRAJ: ()}


def save(name, p, inputs, out, mode, **extra):
    d = {f"p_{k}": v for k, v in params_dict(p).items()}
    d.update({f"in_{k}": v for k, v in inputs.items() if v is not None})
    for k, v in out.items():
        if k not in UNCHANGED[mode]:
            d[f"out_{k}"] = np.asarray(v)
    d.update(extra)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
    print("wrote", name)


def traj_c0():
    """C0's NVE run (1000 e + 1000 i at reduced density 1, dt = 0.005 tau*,
    200 steps) through main.f90:87-153 without the reciprocal part (the
    reference's SPME needs FFTE, which is not here)."""
    dt = reduced_dt(0.005)
    p, x, v = make_plasma(1000, 1000, seed=29)
    f0 = rh.run(rh.MODE_FORCE, p, x, v)["f"]
    out = rh.run(rh.MODE_TRAJ, p, x, v, f0, dt=dt, nsteps=200, thermo=NOTEMP)
    save("traj_c0", p, dict(x=x, v=v, f=f0), out, rh.MODE_TRAJ, dt=dt, thermo=NOTEMP, nsteps=200)


def main():
    # Force_direct_self
    for name, ne, ni, seed, kw in [("force_tiny", 24, 20, 11, {}), ("force_c0", 1000, 1000, 7, {}),
                                    ("force_bins", 1000, 1000, 8, dict(nbins=800))]:
        p, x, v = make_plasma(ne, ni, seed=seed, **kw)
        save(name, p, dict(x=x), rh.run(rh.MODE_FORCE, p, x, v), rh.MODE_FORCE)
    p, x, v = make_plasma(150, 90, seed=5)
    L = p.box[0]
    p = p.replace(box=(L * 1.2, L * 0.9, L * 0.95), rc=0.4 * L, rs=0.37)
    x = x * np.array([1.2, 0.9, 0.95])
    save("force_aniso", p, dict(x=x), rh.run(rh.MODE_FORCE, p, x, v), rh.MODE_FORCE)
    # PSI
    p, x, v = make_plasma(10, 10, seed=2, K=(8, 6, 10))
    p = p.replace(box=(p.box[0] * 1.1, p.box[1], p.box[2] * 0.8), rc=0.39 * p.box[0])
    save("psi", p, dict(), rh.run(rh.MODE_PSI, p, x, v), rh.MODE_PSI)
    # ext / sum
    p, x, v = make_plasma(90, 75, seed=13)
    save("ext", p, dict(x=x), rh.run(rh.MODE_EXT, p, x, v, nown=(40, 30)), rh.MODE_EXT, nown=np.array([40, 30]))
    # rigid
    p, x, v = make_plasma(60, 50, seed=17)
    f = np.random.Generator(np.random.PCG64(3)).standard_normal(x.shape) * 10.0
    save("rigid", p, dict(f=f), rh.run(rh.MODE_RIGID, p, x, v, f), rh.MODE_RIGID)
    # integrators
    dt = reduced_dt(0.005)
    p, x, v = make_plasma(70, 80, seed=19)
    f = np.random.Generator(np.random.PCG64(4)).standard_normal(x.shape) * 30.0
    # a few particles pushed over the box face so the DNINT wrap is exercised
    x[:5, 0] = 0.5 * p.box[0] - 1e-4
    v[:5, 0] = 50.0
    for name, thermo, fric in [("vv1_notemp", NOTEMP, 0.0), ("vv1_stoch", STOCH, 0.7)]:
        q = p.replace(fric=fric)
        save(name, q, dict(x=x, v=v, f=f), rh.run(rh.MODE_VV1, q, x, v, f, dt=dt, thermo=thermo), rh.MODE_VV1,
             dt=dt, thermo=thermo)
    for name, thermo in [("vv2_notemp", NOTEMP), ("vv2_berend", BEREND), ("vv2_isokin", ISOKIN)]:
        q = p.replace(tau=0.05, Te=p.Te * 1.3, Ti=p.Ti * 0.8)
        save(name, q, dict(v=v, f=f), rh.run(rh.MODE_VV2, q, x, v, f, dt=dt, thermo=thermo), rh.MODE_VV2,
             dt=dt, thermo=thermo)
    # trajectories: forces of step 0 from the harness itself (Finit-free restart start)
    for name, thermo, nsteps in [("traj_nve", NOTEMP, 40), ("traj_isokin", ISOKIN, 20)]:
        p, x, v = make_plasma(250, 250, seed=23)
        f0 = rh.run(rh.MODE_FORCE, p, x, v)["f"]
        out = rh.run(rh.MODE_TRAJ, p, x, v, f0, dt=dt, nsteps=nsteps, thermo=thermo)
        save(name, p, dict(x=x, v=v, f=f0), out, rh.MODE_TRAJ, dt=dt, thermo=thermo, nsteps=nsteps)


if __name__ == "__main__":
    cases = sys.argv[1:]
    if not cases:
        main()
        traj_c0()
    elif cases == ["traj_c0"]:
        traj_c0()
    else:
        sys.exit("cases: none (all) or traj_c0")
