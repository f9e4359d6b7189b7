"""GPU parity for the discretisation variants (the reference's compile-time
KO_ORD_3 / 5th-order KO and ADVEC_* options, nuosc_class.cpp:112-183), run by
the radius-3 instantiation of the stage kernel, against the oracle:

  * one calRHS, FP64: relative L-inf <= 1e-12 per field;
  * 30 RK4 steps, FP64: relative L-inf <= 1e-10 per field;
  * calRHS of the FP32 form against FP64 on a state exactly representable in
    FP32: within 64 eps32 max|field| / h;
  * an open-z column with the 5th-order KO (gz = 3), 20 steps: <= 1e-10.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCHEMES = [(3, 1), (3, 2), (3, 3), (3, 4), (3, 5), (5, 0), (5, 3), (5, 4)]   # (ko_order, adv)


def cfg_for(nx, nz, nn=2, dx=0.1, ko=0.1, pmo=0.5, **over):
    c = dict(nv=nn, x0=-nx * dx / 2, x1=nx * dx / 2, z0=-nz * dx / 2, z1=nz * dx / 2, dx=dx, dz=dx, CFL=0.4,
             ko=ko, mu=1.0, pmo=pmo)
    c.update(over)
    return c


def oracle_sim(O, c, ko_order, adv, **over):
    return O.OracleSim(nn=c["nv"], gx=3, gz=3, x0=c["x0"], x1=c["x1"], z0=c["z0"], z1=c["z1"], dx=c["dx"],
                       dz=c["dz"], cfl=c["CFL"], ko=c["ko"], mu=c["mu"], pmo=c["pmo"], ko_order=ko_order, adv=adv,
                       **over)


def rel_linf(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def rough_state(o, seed):
    o.fill_init(ipt=5, alpha=0.9, eps0=0.3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=seed)
    base = o.get_state()
    rng = np.random.default_rng(seed)
    return base * (1 + 0.05 * rng.standard_normal(base.shape))


@pytest.mark.parametrize("ko_order,adv", SCHEMES)
def test_calrhs_schemes_fp64(oracle_mod, nuosc_mod, ko_order, adv):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(13, 11)
    o = oracle_sim(O, c, ko_order, adv)
    g = N.NuOsc(**c, ko_order=ko_order, adv=adv)
    st = rough_state(o, 31 + 7 * adv + ko_order)
    ro, rg = o.calrhs(st), g.calRHS(st)
    for f in range(8):
        assert rel_linf(rg[f], ro[f]) <= 1e-12, (f, rel_linf(rg[f], ro[f]))
    g.close()
    o.close()


@pytest.mark.parametrize("ko_order,adv", SCHEMES)
def test_rk4_schemes_30_steps(oracle_mod, nuosc_mod, ko_order, adv):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(18, 20, ko=0.05)
    o = oracle_sim(O, c, ko_order, adv)
    g = N.NuOsc(**c, ko_order=ko_order, adv=adv)
    init = dict(ipt=0, alpha=0.9, eps0=1e-3, sigma=0.4, lnue=0.6, lnueb=0.5)
    o.fill_init(**init)
    g.fillInitValue(**init)
    o.step(30)
    g.step_rk4(30)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(8):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, (f, rel_linf(gs[f], os_[f]))
    g.close()
    o.close()


@pytest.mark.parametrize("ko_order,adv", SCHEMES)
def test_calrhs_schemes_fp32(oracle_mod, nuosc_mod, ko_order, adv):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(9, 14)
    o = oracle_sim(O, c, ko_order, adv)
    a = N.NuOsc(**c, ko_order=ko_order, adv=adv)
    b = N.NuOsc(**c, ko_order=ko_order, adv=adv, precision=N.PREC_FP32)
    st = rough_state(o, 5 + adv).astype(np.float32).astype(np.float64)
    r64, r32 = a.calRHS(st), b.calRHS(st)
    eps32 = np.finfo(np.float32).eps
    for f in range(8):
        bound = 64 * eps32 * np.max(np.abs(st[f])) / c["dx"]
        assert np.max(np.abs(r32[f] - r64[f])) <= bound, f
    a.close()
    b.close()
    o.close()


def test_open_column_ko5(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    c = dict(nv=2, x0=0.0, x1=0.1, z0=-6.4, z1=6.4, dx=0.1, dz=0.1, CFL=0.4, ko=0.05, mu=1.0, pmo=0.0)
    o = O.OracleSim(nn=2, gx=3, gz=3, x0=0.0, x1=0.1, z0=-6.4, z1=6.4, dx=0.1, dz=0.1, cfl=0.4, ko=0.05, mu=1.0,
                    pmo=0.0, bc=1, zc=-8.0, ko_order=5, adv=3)
    g = N.NuOsc(**c, gx=3, gz=3, bc=N.BC_OPEN_Z, zc=-8.0, ko_order=5, adv=N.ADV_LOPSIDED)
    init = dict(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    o.fill_init(**init)
    g.fillInitValue(**init)
    o.step(20)
    g.step_rk4(20)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(8):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, (f, rel_linf(gs[f], os_[f]))
    g.close()
    o.close()


def test_scheme_arguments_checked(nuosc_mod):
    N = nuosc_mod
    c = cfg_for(8, 8)
    with pytest.raises(N.NuOscError):
        N.NuOsc(**c, ko_order=4)
    with pytest.raises(N.NuOscError):
        N.NuOsc(**c, adv=9)
    col = dict(nv=2, x0=0.0, x1=0.1, z0=-1.6, z1=1.6, dx=0.1, dz=0.1, CFL=0.4, ko=0.1, mu=1.0, pmo=0.0)
    with pytest.raises(N.NuOscError):   # radius 3 needs gz = 3 at open z ends
        N.NuOsc(**col, gz=2, bc=N.BC_OPEN_Z, ko_order=5)
    b = N.Batch(precision=N.PREC_FP64)
    g = N.NuOsc(**c, adv=N.ADV_UPWIND2)
    with pytest.raises(N.NuOscError):
        b.add(g)
    g.close()
    b.close()


def test_cli_scheme_flags(tmp_path):
    exe = os.path.join(ROOT, "nuosc_amd", "bin", "nuosc")
    p = subprocess.run([exe, "--zmax", "0.8", "--dz", "0.1", "--nv", "2", "--ko", "0.05", "--END_STEP", "10",
                        "--ANA_EVERY", "5", "--ko_order", "5", "--adv", "lopsided"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "5-th order KO" in p.stdout and "lopsided" in p.stdout
    rows = [l for l in open(tmp_path / "analysis.dat").read().splitlines()[1:] if l.strip()]
    assert len(rows) == 3
