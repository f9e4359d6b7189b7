"""The ./nuosc drop-in (nuosc_amd/bin/nuosc) on the GPU: analysis.dat against
the oracle's lines, the reference's stdout lines, P3 snapshot files in the
reference's binary format (nuosc_snapshot.cpp:82-122), renormalisation, and
the unknown-flag behaviour (nuosc.cpp:135-137)."""
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "nuosc_amd", "bin", "nuosc")

ARGS = ["--zmax", "1.6", "--xmax", "0.8", "--dz", "0.1", "--nv", "3", "--mu", "1", "--pmo", "1", "--ko", "0.1",
        "--cfl", "0.4", "--ipt", "0", "--eps0", "1e-3", "--sigma", "1", "--alpha", "0.9", "--lnue", "0.6",
        "--lnueb", "0.5", "--END_STEP", "30", "--ANA_EVERY", "10"]


def run_cli(tmp_path, args, timeout=120):
    return subprocess.run([CLI] + args, cwd=tmp_path, capture_output=True, text=True, timeout=timeout)


def test_cli_analysis_dat_matches_oracle(tmp_path, oracle_mod):
    p = run_cli(tmp_path, ARGS)
    assert p.returncode == 0, p.stderr
    lines = open(tmp_path / "analysis.dat").read().splitlines()
    assert lines[0].startswith("### [ phy_time,")
    rows = [l for l in lines[1:] if l.strip()]
    assert len(rows) == 4
    o = oracle_mod.OracleSim(nn=3, x0=-0.8, x1=0.8, z0=-1.6, z1=1.6, dx=0.1, dz=0.1, cfl=0.4, ko=0.1, mu=1.0,
                             pmo=1.0)
    o.fill_init(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    want = []
    for k in range(4):
        if k:
            o.step(10)
        out, _ = o.analysis()
        want.append(o.analysis_line(out))
    for got, exp in zip(rows, want):
        g, e = got.split(), exp.split()
        assert len(g) == len(e) == 9
        assert g[0] == e[0]                       # phy_time, same stream precision
        for a, b in zip(g[1:], e[1:]):
            a, b = float(a), float(b)
            assert abs(a - b) <= 1e-10 * abs(b) + 1e-13, (got, exp)
    assert rows[0].endswith(" ") and "T=" in p.stdout and "[Summ]" in p.stdout
    assert "ns/step-grid" in p.stdout


def test_cli_p3_snapshot_format(tmp_path):
    p = run_cli(tmp_path, ARGS + ["--DUMP_EVERY", "15", "--snap_p3", "1"])
    assert p.returncode == 0, p.stderr
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("P3_"))
    assert files == ["P3_000000.bin", "P3_000000.bin.meta", "P3_000015.bin", "P3_000030.bin"]
    meta = open(tmp_path / "P3_000000.bin.meta").read().splitlines()
    dt, sx, nz, sv = meta[0].split()
    assert (int(sx), int(nz), int(sv)) == (3, 32, 3)
    assert len(meta[3].split()) == 32 and len(meta[4].split()) == 3
    raw = open(tmp_path / "P3_000015.bin", "rb").read()
    t, = struct.unpack_from("<i", raw, 0)
    ptime, = struct.unpack_from("<d", raw, 4)
    data = np.frombuffer(raw[12:], dtype="<f8")
    assert t == 15 and abs(ptime - 15 * 0.04) < 1e-12
    assert data.size == 3 * 32 * 3
    assert np.all(np.abs(data) <= 1.0 + 1e-9)


def test_cli_renorm_matches_oracle(tmp_path, oracle_mod):
    p = run_cli(tmp_path, ARGS + ["--renorm", "1"])
    assert p.returncode == 0, p.stderr
    rows = [l for l in open(tmp_path / "analysis.dat").read().splitlines()[1:] if l.strip()]
    o = oracle_mod.OracleSim(nn=3, x0=-0.8, x1=0.8, z0=-1.6, z1=1.6, dx=0.1, dz=0.1, cfl=0.4, ko=0.1, mu=1.0,
                             pmo=1.0, renorm=1)
    o.fill_init(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    o.step(30)
    out, _ = o.analysis()
    g = [float(x) for x in rows[-1].split()]
    for a, b in zip(g[1:], out[1:]):
        assert abs(a - b) <= 1e-10 * abs(b) + 1e-13


def test_cli_unknown_flag(tmp_path):
    p = run_cli(tmp_path, ["--bogus", "1"])
    assert p.returncode == 0
    assert "Unreconganized parameters --bogus!" in p.stdout


def test_cli_open_boundary_and_growth_file(tmp_path, oracle_mod):
    """--bc open --zc: the 1-D open column with an entering pulse; analysis.dat
    against the oracle and growth.dat holding |rho_ex| per analysis call."""
    # one x cell: --xmax after --zmax, --dx after --dz (which sets dx too)
    args = ["--zmax", "6.4", "--xmax", "0.05", "--dz", "0.1", "--dx", "0.1", "--nv", "2", "--mu", "1", "--pmo", "0",
            "--ko", "0.1", "--cfl", "0.4", "--ipt", "0", "--eps0", "1e-3", "--sigma", "1", "--alpha", "0.9",
            "--lnue", "0.6", "--lnueb", "0.5", "--END_STEP", "40", "--ANA_EVERY", "20", "--bc", "open",
            "--zc", "-8"]
    p = run_cli(tmp_path, args)
    assert p.returncode == 0, p.stderr
    assert "z open" in p.stdout
    rows = [l for l in open(tmp_path / "analysis.dat").read().splitlines()[1:] if l.strip()]
    grow = [l for l in open(tmp_path / "growth.dat").read().splitlines()[1:] if l.strip()]
    assert len(rows) == len(grow) == 3
    o = oracle_mod.OracleSim(nn=2, x0=-0.05, x1=0.05, z0=-6.4, z1=6.4, dx=0.1, dz=0.1, cfl=0.4, ko=0.1, mu=1.0,
                             pmo=0.0, bc=1, zc=-8.0)
    assert o.shape == (1, 128, 42)
    o.fill_init(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    for k in range(3):
        if k:
            o.step(20)
        out, _ = o.analysis()
        exp = o.analysis_line(out).split()
        got = rows[k].split()
        for a, b in zip(got[1:], exp[1:]):
            a, b = float(a), float(b)
            assert abs(a - b) <= 1e-10 * abs(b) + 1e-13, (rows[k], exp)
        st = o.get_state()
        rex = np.sqrt(o.s.ds_L * np.sum(o.vgrid()[3][None, None, :] * (st[2] ** 2 + st[3] ** 2)))
        t, r = map(float, grow[k].split())
        assert t == float(exp[0]) or abs(t - float(exp[0])) <= 1e-12
        assert abs(r - rex) <= 1e-9 * rex


def ostream_g(x):
    """A double as std::ostream writes it with the default flags and
    precision 6: the %g conversion (C and Python agree on it)."""
    return "%g" % x


def test_cli_p3_snapshot_parity(tmp_path, oracle_mod):
    """P3 snapshots against the oracle and the reference's writer
    (nuosc_snapshot.cpp:13-41, 82-122):
      * .meta byte for byte: 'dt sx nz sv', 'x0 x1', 'z0 z1', then Z[0..nz),
        X[xidx], vx[vidx], vy[vidx], vz[vidx], each value followed by a blank,
        every number as std::ostream writes it at precision 6;
      * .bin: int32 step, f64 phy_time (bitwise the oracle's accumulated
        time), then f64[sx][nz][sv] = (ee - mm)/G0 at x = 0, nx/2, nx-1 and
        the directions ((nn-1)/2) nn + v, v < nn (nuosc.cpp:169-175), within
        1e-10 of P3 from the oracle's state after the same steps."""
    p = run_cli(tmp_path, ARGS + ["--DUMP_EVERY", "15", "--snap_p3", "1"])
    assert p.returncode == 0, p.stderr
    o = oracle_mod.OracleSim(nn=3, x0=-0.8, x1=0.8, z0=-1.6, z1=1.6, dx=0.1, dz=0.1, cfl=0.4, ko=0.1, mu=1.0,
                             pmo=1.0)
    o.fill_init(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    nx, nz, nv = o.shape
    xidx = [0, nx // 2, nx - 1]
    vidx = [((3 - 1) // 2) * 3 + v for v in range(3)]
    vx, vy, vz, _ = o.vgrid()
    X = -0.8 + (np.arange(nx) + 0.5) * 0.1
    Z = -1.6 + (np.arange(nz) + 0.5) * 0.1
    lines = [" ".join(ostream_g(v) for v in (o.s.dt,)) + " " + " ".join(str(n) for n in (len(xidx), nz, len(vidx))),
             f"{ostream_g(-0.8)} {ostream_g(0.8)}", f"{ostream_g(-1.6)} {ostream_g(1.6)}",
             "".join(ostream_g(z) + " " for z in Z), "".join(ostream_g(X[i]) + " " for i in xidx),
             "".join(ostream_g(vx[v]) + " " for v in vidx), "".join(ostream_g(vy[v]) + " " for v in vidx),
             "".join(ostream_g(vz[v]) + " " for v in vidx)]
    want_meta = "".join(line + "\n" for line in lines).encode()
    assert open(tmp_path / "P3_000000.bin.meta", "rb").read() == want_meta
    G0 = o.G0
    for t in (0, 15, 30):
        if t:
            o.step(15)
        raw = open(tmp_path / f"P3_{t:06d}.bin", "rb").read()
        assert len(raw) == 4 + 8 + 8 * len(xidx) * nz * len(vidx)
        step, = struct.unpack_from("<i", raw, 0)
        ptime, = struct.unpack_from("<d", raw, 4)
        assert step == t and ptime == o.phy_time
        got = np.frombuffer(raw[12:], dtype="<f8").reshape(len(xidx), nz, len(vidx))
        st = o.get_state()
        p3 = (st[0] - st[1]) / G0[None, None, :]
        want = p3[xidx][:, :, vidx]
        assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(want)), t
