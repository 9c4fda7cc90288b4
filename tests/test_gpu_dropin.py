"""The unchanged reference driver (main.f90, init.f90, data.f90, result.f90)
linked with the GPU drop-in for force.f90/vverlet.f90, against the oracle's
step on the same restart state.

`mpirun -np 2`: rank 0 is the idle PME rank, rank 1 the only direct rank with
every particle on its context.  `-np 3` and `-np 4`: two and three direct
ranks run the reference's real-space ring (the even case with its antipodal
exchange), the PME rank computes the reciprocal part of every gathered
particle on its own context, and BEREND/ISOKIN reduce their kinetic sums
over the direct ranks.  `-np 10` and `-np 12`: two PME groups
(data.f90:151) of four and five direct ranks; the PME ranks exchange their
groups' positions and each computes the whole mesh.

Tolerances: the reference's own writers print ES15.8 energies and ES14.6
positions (result.f90:324, :131), so the files are compared at 2e-8 and 2e-6
relative respectively.  The RDF is compared where the reference counts each
pair once (one direct rank, or an odd number of them); with an even number
the antipodal blocks are counted by both partners (main.f90:121-130).
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import cpu
from tcp_allpair_amd.inputs import (make_plasma, read_restart_xyz, write_input_xyz, write_restart_xyz,
                                   write_tcp_prm)
from tcp_allpair_amd.params import BEREND, ISOKIN, NBIN, NOTEMP, PI, STOCH, TPS, reduced_dt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tcp_allpair_amd", "fortran", "_build", "tcp_pbc_gpu")
MPIRUN = "/opt/conda/bin/mpirun"
THERMO_NAME = {NOTEMP: "NOTEMP", BEREND: "BEREND", ISOKIN: "ISOKIN"}


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in driver not built (needs the reference sources)")
@pytest.mark.parametrize("nranks,thermo,sync", [(2, NOTEMP, False), (2, NOTEMP, True),
                                                (3, NOTEMP, False), (4, NOTEMP, False),
                                                (4, BEREND, False), (3, ISOKIN, False),
                                                (2, ISOKIN, False), (10, NOTEMP, False),
                                                (12, BEREND, False)])
def test_dropin_driver(tmp_path, nranks, thermo, sync):
    """sync: TCP_DROPIN_SYNC=1, every subroutine downloads what it writes (with one
    direct rank the step's results otherwise stay on the device until VVerlet*2)."""
    print("case dropin", nranks, THERMO_NAME[thermo], "sync" if sync else "", flush=True)
    nsteps = 10
    p, x, v = make_plasma(250, 251, seed=51, K=(32, 32, 32), order=4)
    p = p.replace(tau=0.5, fric=0.0, Te=p.Te * 1.2, Ti=p.Ti * 0.8)
    f0, _, _ = cpu.force_allpair(p, x)
    fr, _ = cpu.spme(p, x)
    f0 = cpu.remove_rigidmtn(p, f0 + fr)
    dt = reduced_dt(0.005)
    dt_fs = dt * TPS
    write_tcp_prm(tmp_path / "tcp.prm", p, tmax_fs=nsteps * dt_fs, trest_fs=nsteps * dt_fs,
                  tdump_fs=5 * dt_fs, tsamp_fs=dt_fs, dt_fs=dt_fs, thermo=THERMO_NAME[thermo],
                  rest="ON")
    write_restart_xyz(tmp_path / "restart.xyz", p, x, v, f0)
    env = dict(os.environ, TCP_PRECISION="FP64")
    if sync:
        env["TCP_DROPIN_SYNC"] = "1"
    r = subprocess.run([MPIRUN, "-launcher", "fork", "-n", str(nranks), BIN], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    # oracle on the same state (restart.xyz holds %.17g values, read back exactly)
    xr, vr, frr = read_restart_xyz(tmp_path / "restart.xyz")
    dt_int = dt_fs / TPS
    rows = []
    rdf = np.zeros(3 * NBIN)
    vdf = np.zeros(2 * NBIN)
    for s in range(nsteps):
        o = cpu.step(p, thermo, dt_int, xr, vr, frr, rdf, vdf, fast=True)
        ep = o["Epot"] + o["Um"]
        rows.append([(s + 1) * dt_int * TPS, o["mv2el"] / (3 * p.ne), o["mv2ion"] / (3 * p.ni), ep,
                     0.5 * (o["mv2el"] + o["mv2ion"]) + ep])
        if s == 4:
            rdf5, vdf5 = rdf.copy(), vdf.copy()
    ener = np.loadtxt(tmp_path / "ener00.dat")
    np.testing.assert_allclose(ener, np.array(rows), rtol=2e-8)
    # REST01.xyz at ES14.6
    xs, vs, fs = read_restart_xyz(tmp_path / "REST01.xyz")
    np.testing.assert_allclose(xs, xr, rtol=2e-6, atol=1e-6 * p.box[0])
    np.testing.assert_allclose(vs, vr, rtol=2e-6, atol=1e-6 * np.abs(vr).max())
    np.testing.assert_allclose(fs, frr, rtol=2e-6, atol=1e-6 * np.abs(frr).max())
    # VDF001.dat: counts of the first 5 steps
    ntau = 5.0
    vcols = np.loadtxt(tmp_path / "VDF001.dat")
    hv = vdf5.reshape(2, NBIN)
    np.testing.assert_allclose(vcols[:, 1], hv[0, :NBIN - 1] / ntau, rtol=2e-6)
    np.testing.assert_allclose(vcols[:, 3], hv[1, :NBIN - 1] / ntau, rtol=2e-6)
    direct = nranks - ((nranks - 1) // 8 + 1)  # ranks minus PME ranks (data.f90:151)
    if direct % 2 == 0:  # antipodal blocks counted by both partners (main.f90:121-130)
        return
    # RDF001.dat: counts of the first 5 steps, normalised as result.f90:244-252
    V = p.volume
    cols = np.loadtxt(tmp_path / "RDF001.dat")
    i = np.arange(1, NBIN)
    shell = (i ** 3 - (i - 1) ** 3).astype(float)
    norms = [p.ne ** 2 * ntau * p.dx[0] ** 3 * 4 * PI / 3 / V, p.n ** 2 * ntau * p.dx[1] ** 3 * 4 * PI / 3 / V,
             p.ni ** 2 * ntau * p.dx[2] ** 3 * 4 * PI / 3 / V]
    h = rdf5.reshape(3, NBIN)
    for t in range(3):
        np.testing.assert_allclose(cols[:, 2 * t + 1], h[t, :NBIN - 1] / norms[t] / shell, rtol=2e-6, atol=1e-12)


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in driver not built (needs the reference sources)")
@pytest.mark.parametrize("nranks", [2, 3])
def test_dropin_fresh_start(tmp_path, nranks):
    """rest = 'OFF': main.f90 reads input.xyz (main.f90:347-379), Vinit draws the
    velocities (init.f90:26-112) and Finit computes the forces at x - dt v through the
    drop-in's Force_direct_self / Force_direct_ext with their extra PRE_data argument
    (init.f90:195, 202-203, 215-216) before the steps.  Vinit's stream is the
    compiler's unseeded RANDOM_NUMBER, so Te = Ti = 1e-30 eV scales its velocities to
    about 1e-15 (electrons): the start is then x0 at rest up to 1e-12 of the first
    kick, and the oracle's steps from (x0, v = 0, F(x0)) must reproduce ener00.dat
    (ES15.8), REST01.xyz (ES14.6) and the RDF counts.  The VDF bins (width 4 T/500)
    count nothing at such a temperature."""
    print("case dropin fresh start", nranks, flush=True)
    nsteps = 8
    p, x, _ = make_plasma(240, 250, seed=53, K=(32, 32, 32), order=4)
    p = p.replace(Te=1e-30, Ti=1e-30)
    dt = reduced_dt(0.005)
    dt_fs = dt * TPS
    write_tcp_prm(tmp_path / "tcp.prm", p, tmax_fs=nsteps * dt_fs, trest_fs=nsteps * dt_fs,
                  tdump_fs=4 * dt_fs, tsamp_fs=dt_fs, dt_fs=dt_fs, thermo="NOTEMP", rest="OFF")
    write_input_xyz(tmp_path / "input.xyz", p, x)
    env = dict(os.environ, TCP_PRECISION="FP64")
    r = subprocess.run([MPIRUN, "-launcher", "fork", "-n", str(nranks), BIN], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    f0, _, _ = cpu.force_allpair(p, x)
    fr, _ = cpu.spme(p, x)
    xo, vo = x.copy(), np.zeros_like(x)
    fo = cpu.remove_rigidmtn(p, f0 + fr)
    dt_int = dt_fs / TPS
    rows = []
    rdf = np.zeros(3 * NBIN)
    vdf = np.zeros(2 * NBIN)
    for s in range(nsteps):
        o = cpu.step(p, NOTEMP, dt_int, xo, vo, fo, rdf, vdf, fast=True)
        ep = o["Epot"] + o["Um"]
        rows.append([(s + 1) * dt_int * TPS, o["mv2el"] / (3 * p.ne), o["mv2ion"] / (3 * p.ni), ep,
                     0.5 * (o["mv2el"] + o["mv2ion"]) + ep])
        if s == 3:
            rdf4 = rdf.copy()
    ener = np.loadtxt(tmp_path / "ener00.dat")
    np.testing.assert_allclose(ener, np.array(rows), rtol=2e-8)
    xs, vs, fs = read_restart_xyz(tmp_path / "REST01.xyz")
    np.testing.assert_allclose(xs, xo, rtol=2e-6, atol=1e-6 * p.box[0])
    np.testing.assert_allclose(vs, vo, rtol=2e-6, atol=1e-6 * np.abs(vo).max())
    np.testing.assert_allclose(fs, fo, rtol=2e-6, atol=1e-6 * np.abs(fo).max())
    assert not np.loadtxt(tmp_path / "VDF001.dat")[:, [1, 3]].any()
    if nranks == 2:  # one direct rank counts each pair once
        V = p.volume
        cols = np.loadtxt(tmp_path / "RDF001.dat")
        i = np.arange(1, NBIN)
        shell = (i ** 3 - (i - 1) ** 3).astype(float)
        norms = [p.ne ** 2 * 4 * p.dx[0] ** 3 * 4 * PI / 3 / V,
                 p.n ** 2 * 4 * p.dx[1] ** 3 * 4 * PI / 3 / V,
                 p.ni ** 2 * 4 * p.dx[2] ** 3 * 4 * PI / 3 / V]
        h = rdf4.reshape(3, NBIN)
        for t in range(3):
            np.testing.assert_allclose(cols[:, 2 * t + 1], h[t, :NBIN - 1] / norms[t] / shell,
                                       rtol=2e-6, atol=1e-12)


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in driver not built (needs the reference sources)")
def test_dropin_fresh_start_thermal(tmp_path):
    """rest = 'OFF' at T* = 1: Vinit's velocities are the compiler's, so the run is
    checked through what they cannot change.  ener00.dat's first row holds the
    temperatures after one step of Finit's forces; the forces in REST01.xyz are the
    oracle's forces at REST01.xyz's printed positions (RMS relative 1e-4: the ES14.6
    positions move close pairs by parts in 1e6 of their separation)."""
    print("case dropin fresh start thermal", flush=True)
    nsteps = 6
    p, x, _ = make_plasma(240, 250, seed=54, K=(32, 32, 32), order=4)
    dt_fs = reduced_dt(0.005) * TPS
    write_tcp_prm(tmp_path / "tcp.prm", p, tmax_fs=nsteps * dt_fs, trest_fs=nsteps * dt_fs,
                  tdump_fs=nsteps * dt_fs, tsamp_fs=dt_fs, dt_fs=dt_fs, thermo="NOTEMP", rest="OFF")
    write_input_xyz(tmp_path / "input.xyz", p, x)
    env = dict(os.environ, TCP_PRECISION="FP64")
    r = subprocess.run([MPIRUN, "-launcher", "fork", "-n", "2", BIN], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    ener = np.loadtxt(tmp_path / "ener00.dat")
    assert ener.shape == (nsteps, 5) and np.isfinite(ener).all()
    # one step from T = Te, Ti moves the temperatures by far less than they are
    assert 0.5 * p.Te < ener[0, 1] < 2.0 * p.Te and 0.5 * p.Ti < ener[0, 2] < 2.0 * p.Ti
    # NVE: the total energy holds to the integrator's fluctuation (against the kinetic energy)
    assert np.ptp(ener[:, 4]) <= 1e-3 * 1.5 * (p.ne * p.Te + p.ni * p.Ti)
    xs, _, fs = read_restart_xyz(tmp_path / "REST01.xyz")
    f0, _, _ = cpu.force_allpair(p, xs)
    fr, _ = cpu.spme(p, xs)
    fo = cpu.remove_rigidmtn(p, f0 + fr)
    err = np.sqrt(((fs - fo) ** 2).sum() / (fo ** 2).sum())
    print("REST01 forces rms rel err", err, flush=True)
    assert err <= 1e-4


RUN = os.path.join(ROOT, "tcp_allpair_amd", "tcp_run")


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in driver not built (needs the reference sources)")
@pytest.mark.parametrize("thermo", [NOTEMP, BEREND, STOCH])
def test_standalone_driver_matches_dropin_bytes(tmp_path, thermo):
    """tcp_run (C++ host I/O over the C ABI, steps in tcp_md_steps) and the unchanged
    Fortran driver over the drop-in write the same files byte for byte in FP64."""
    name = {NOTEMP: "NOTEMP", BEREND: "BEREND", STOCH: "STOCH"}[thermo]
    print("case standalone vs dropin", name, flush=True)
    nsteps = 12
    p, x, v = make_plasma(300, 301, seed=77, K=(32, 32, 32), order=4)
    p = p.replace(tau=0.5, fric=0.2, Te=p.Te * 1.1, Ti=p.Ti * 0.9)
    f0, _, _ = cpu.force_allpair(p, x)
    fr, _ = cpu.spme(p, x)
    f0 = cpu.remove_rigidmtn(p, f0 + fr)
    dt_fs = reduced_dt(0.005) * TPS
    dirs = {}
    for who in ("dropin", "standalone"):
        d = tmp_path / who
        d.mkdir()
        write_tcp_prm(d / "tcp.prm", p, tmax_fs=nsteps * dt_fs, trest_fs=6 * dt_fs,
                      tdump_fs=4 * dt_fs, tsamp_fs=2 * dt_fs, dt_fs=dt_fs, thermo=name, rest="ON")
        write_restart_xyz(d / "restart.xyz", p, x, v, f0)
        dirs[who] = d
    env = dict(os.environ, TCP_PRECISION="FP64")
    r = subprocess.run([MPIRUN, "-launcher", "fork", "-n", "2", BIN], cwd=dirs["dropin"], env=env,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([RUN], cwd=dirs["standalone"], env=env, capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    outs = sorted(n for n in os.listdir(dirs["dropin"]) if n not in ("tcp.prm", "restart.xyz"))
    assert outs == sorted(n for n in os.listdir(dirs["standalone"]) if n not in ("tcp.prm", "restart.xyz"))
    assert {"ener00.dat", "ener01.dat", "ener02.dat", "REST01.xyz", "REST02.xyz", "RDF001.dat",
            "RDF003.dat", "VDF002.dat"} <= set(outs)
    for n in outs:
        assert (dirs["dropin"] / n).read_bytes() == (dirs["standalone"] / n).read_bytes(), n
