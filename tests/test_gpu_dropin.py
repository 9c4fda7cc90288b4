"""The unchanged reference driver (main.f90, init.f90, data.f90, result.f90)
linked with the GPU drop-in for force.f90/vverlet.f90, against the oracle's
step on the same restart state.

`mpirun -np 2`: rank 0 is the idle PME rank, rank 1 the only direct rank with
every particle on its context.  `-np 3` and `-np 4`: two and three direct
ranks run the reference's real-space ring (the even case with its antipodal
exchange), the PME rank computes the reciprocal part of every gathered
particle on its own context, and BEREND/ISOKIN reduce their kinetic sums
over the direct ranks.

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
from tcp_allpair_amd.inputs import make_plasma, read_restart_xyz, write_restart_xyz, write_tcp_prm
from tcp_allpair_amd.params import BEREND, ISOKIN, NBIN, NOTEMP, PI, STOCH, TPS, reduced_dt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tcp_allpair_amd", "fortran", "_build", "tcp_pbc_gpu")
MPIRUN = "/opt/conda/bin/mpirun"
THERMO_NAME = {NOTEMP: "NOTEMP", BEREND: "BEREND", ISOKIN: "ISOKIN"}


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in driver not built (needs the reference sources)")
@pytest.mark.parametrize("nranks,thermo", [(2, NOTEMP), (3, NOTEMP), (4, NOTEMP), (4, BEREND),
                                           (3, ISOKIN)])
def test_dropin_driver(tmp_path, nranks, thermo):
    print("case dropin", nranks, THERMO_NAME[thermo], flush=True)
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
    if (nranks - 1) % 2 == 0 and nranks > 2:
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
