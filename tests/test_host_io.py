"""The host-side readers and writers of the reference's text files (libtcp_io.so,
shared with the standalone driver tcp_run).

The writers are pinned byte for byte against the reference's own DUMP, SAMPLING
and Restart (result.f90, compiled from its sources into oracle/_ref) on the same
data; Fortran's ES edit descriptor against strings the flang runtime printed for
the edge cases (three-digit exponents drop the letter E, negative zero); the
readers by round trips of the files the tests write for the Fortran driver.
"""
import os

import numpy as np
import pytest

from oracle import refharness
from tcp_allpair_amd import hostio
from tcp_allpair_amd.inputs import make_plasma, write_restart_xyz, write_tcp_prm
from tcp_allpair_amd.params import NBIN

# strings the flang runtime wrote for FORMAT(ES14.6) and FORMAT(ES15.8)
FLANG_ES = [
    (1.5, 14, 6, "  1.500000E+00"), (-2.25, 14, 6, " -2.250000E+00"), (0.0, 14, 6, "  0.000000E+00"),
    (1e100, 14, 6, "  1.000000+100"), (1e-100, 14, 6, "  1.000000-100"),
    (-0.0, 14, 6, " -0.000000E+00"), (123456.75, 14, 6, "  1.234568E+05"),
    (9.9999995, 14, 6, "  9.999999E+00"), (1e10, 14, 6, "  1.000000E+10"),
    (1e-10, 14, 6, "  1.000000E-10"), (1.234567891, 15, 8, " 1.23456789E+00"),
    (-2.0, 15, 8, "-2.00000000E+00"), (1e-300, 15, 8, " 1.00000000-300"),
    (7e99, 15, 8, " 7.00000000E+99"), (1e100, 15, 8, " 1.00000000+100"),
]


@pytest.mark.parametrize("v,w,d,s", FLANG_ES)
def test_format_es_matches_flang(v, w, d, s):
    assert hostio.format_es(v, w, d) == s


def test_prm_reader_round_trip(tmp_path):
    p, _, _ = make_plasma(123, 77, seed=3, K=(32, 16, 8))
    p = p.replace(tau=0.25, fric=0.125, Te=11.5, Ti=3.25)
    write_tcp_prm(tmp_path / "tcp.prm", p, tmax_fs=1.5, trest_fs=0.5, tdump_fs=0.25, tsamp_fs=0.125,
                  dt_fs=0.015625, thermo="STOCH", rest="ON")
    r = hostio.read_prm(tmp_path / "tcp.prm")
    assert (r.tmax, r.trest, r.tdump, r.tsamp, r.dt) == (1.5, 0.5, 0.25, 0.125, 0.015625)
    assert tuple(r.box) == tuple(p.box) and (r.Te, r.Ti) == (11.5, 3.25)
    assert (r.NTel, r.NTion) == (123, 77) and (r.rc, r.rs) == (p.rc, p.rs)
    assert tuple(r.xm) == tuple(p.xm) and tuple(r.q) == tuple(p.q)
    assert r.thermo == b"STOCH " and (r.tau, r.alpha) == (0.25, 0.125)
    assert r.rest == b"ON" and r.trace == b"OFF"
    assert tuple(r.dx) == tuple(p.dx) and tuple(r.K) == (32, 16, 8)


def test_prm_reader_errors(tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        hostio.read_prm(tmp_path / "missing.prm")
    (tmp_path / "bad.prm").write_text("# title\n# times\n1 2 3\n")
    with pytest.raises(RuntimeError, match="malformed"):
        hostio.read_prm(tmp_path / "bad.prm")


def test_restart_reader_round_trip(tmp_path):
    p, x, v = make_plasma(40, 35, seed=9)
    f = np.random.default_rng(1).standard_normal(x.shape) * 1e3
    write_restart_xyz(tmp_path / "restart.xyz", p, x, v, f)
    xr, vr, fr, ids = hostio.read_restart(tmp_path / "restart.xyz", p.n)
    np.testing.assert_array_equal(xr, x)
    np.testing.assert_array_equal(vr, v)
    np.testing.assert_array_equal(fr, f)
    np.testing.assert_array_equal(ids, np.r_[np.arange(1, 41), np.arange(1, 36)])
    with pytest.raises(RuntimeError, match="malformed"):
        hostio.read_restart(tmp_path / "restart.xyz", p.n + 1)


@pytest.mark.skipif(not refharness.available(), reason="reference harness not built (needs the reference sources)")
def test_writers_match_reference_writers(tmp_path, monkeypatch):
    rng = np.random.default_rng(7)
    p, x, v = make_plasma(37, 41, seed=5)
    p = p.replace(Te=17.25, Ti=4.5)
    f = rng.standard_normal(x.shape) * 10.0 ** rng.integers(-3, 4, size=x.shape)
    f[0, 0], f[1, 1], v[2, 2] = 1e-120, -0.0, 0.0
    ids = rng.permutation(p.n) + 1
    rdf = 2.0 * rng.integers(0, 400, size=(3, NBIN))
    vdf = 1.0 * rng.integers(0, 50, size=(2, NBIN))
    ndump, nrest = 25, 50
    nloop = 150
    tnow, mv2el, mv2ion, epot, um = 1.2345678901, 987.654321, 1234.5, -4321.123456789, 12.75
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    ref_dir.mkdir()
    our_dir.mkdir()
    refharness.writers(str(ref_dir), p, x, v, f, ids, ndump, nloop, nrest, tnow, mv2el, mv2ion,
                       epot, um, rdf, vdf)
    monkeypatch.chdir(our_dir)
    hostio.write_dump(nloop // ndump, p.ne, p.ni, ndump, p.box, p.dx, p.dv[0], p.dv[1], rdf, vdf)
    hostio.ener_open("ener00.dat")
    tps = 10.180505306645389
    hostio.ener_row("ener00.dat", tnow * tps, mv2el / (3 * p.ne), mv2ion / (3 * p.ni), um + epot,
                    0.5 * (mv2el + mv2ion) + (um + epot))
    hostio.write_rest(nloop // nrest, p.ne, p.ni, x, v, f, ids)
    hostio.ener_open(f"ener{nloop // nrest:02d}.dat")
    names = ["RDF006.dat", "VDF006.dat", "REST03.xyz", "ener03.dat"]
    for name in names:
        assert (ref_dir / name).read_bytes() == (our_dir / name).read_bytes(), name
    # the harness opens ener00.dat without the header main.f90 writes
    ours = (our_dir / "ener00.dat").read_text().splitlines()
    assert (ref_dir / "ener00.dat").read_text().splitlines() == ours[1:]
    assert sorted(os.listdir(ref_dir)) == sorted(os.listdir(our_dir))
