"""Grid shapes that exercise the edges of the stage kernel's decomposition.

A block of the stage kernel covers 4 waves x R x rows (FP64 R = 4, FP32
R = 2) of one direction chunk (16 FP64 / 32 FP32 directions).  These shapes
put rows past nx into the last wave (nx % R != 0), whole waves past the last
quad into the last block, periodic ghost rows that wrap more than once (nx <
2) and chunks only partly filled with directions (42 and 162 directions):

  * calRHS against the oracle, FP64: relative L-inf <= 1e-12 per field;
  * calRHS of the FP32 form against FP64 on a state exactly representable in
    FP32: within 64 eps32 max|field| / h per field (FP32 rounding of the
    stencil);
  * 30 RK4 steps against the oracle, FP64: relative L-inf <= 1e-10 per field.
  * one x row: the stage kernel's NOX form (alone) and its ordinary form (in a
    batch) give the same bits over 20 RK4 steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 5), (2, 7), (3, 9), (5, 6), (13, 11), (18, 4), (33, 3)]


def cfg_for(nx, nz, nn, dx=0.1, pmo=1.0):
    return dict(nv=nn, x0=-nx * dx / 2, x1=nx * dx / 2, z0=-nz * dx / 2, z1=nz * dx / 2, dx=dx, dz=dx, CFL=0.4,
                ko=0.1, mu=1.0, pmo=pmo)


def oracle_sim(O, c):
    return O.OracleSim(nn=c["nv"], x0=c["x0"], x1=c["x1"], z0=c["z0"], z1=c["z1"], dx=c["dx"], dz=c["dz"],
                       cfl=c["CFL"], ko=c["ko"], mu=c["mu"], pmo=c["pmo"])


def rel_linf(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def rough_state(o, seed):
    o.fill_init(ipt=5, alpha=0.9, eps0=0.3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=seed)
    base = o.get_state()
    rng = np.random.default_rng(seed)
    return base * (1 + 0.05 * rng.standard_normal(base.shape))


@pytest.mark.parametrize("nn", [2, 4])
@pytest.mark.parametrize("nx,nz", SHAPES)
def test_calrhs_shapes_fp64(oracle_mod, nuosc_mod, nx, nz, nn):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(nx, nz, nn)
    o = oracle_sim(O, c)
    g = N.NuOsc(**c)
    assert (g.nx, g.nz) == (nx, nz) and o.shape[:2] == (nx, nz)
    st = rough_state(o, 100 * nx + nz)
    ro, rg = o.calrhs(st), g.calRHS(st)
    for f in range(8):
        assert rel_linf(rg[f], ro[f]) <= 1e-12, (f, rel_linf(rg[f], ro[f]))
    g.close()
    o.close()


@pytest.mark.parametrize("nn", [2, 4])
@pytest.mark.parametrize("nx,nz", SHAPES)
def test_calrhs_shapes_fp32(oracle_mod, nuosc_mod, nx, nz, nn):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(nx, nz, nn)
    o = oracle_sim(O, c)
    a = N.NuOsc(**c)
    b = N.NuOsc(**c, precision=N.PREC_FP32)
    st = rough_state(o, 7 * nx + nz).astype(np.float32).astype(np.float64)
    r64, r32 = a.calRHS(st), b.calRHS(st)
    eps32 = np.finfo(np.float32).eps
    for f in range(8):
        bound = 64 * eps32 * np.max(np.abs(st[f])) / c["dx"]
        err = np.max(np.abs(r32[f] - r64[f]))
        assert err <= bound, (f, err, bound)
    a.close()
    b.close()
    o.close()


@pytest.mark.parametrize("nx,nz", [(1, 5), (3, 9), (5, 6), (18, 4)])
def test_rk4_shapes_30_steps(oracle_mod, nuosc_mod, nx, nz):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(nx, nz, 2, pmo=0.5)
    o = oracle_sim(O, c)
    g = N.NuOsc(**c)
    init = dict(ipt=0, alpha=0.9, eps0=1e-3, sigma=0.3, lnue=0.6, lnueb=0.5)
    o.fill_init(**init)
    g.fillInitValue(**init)
    o.step(30)
    g.step_rk4(30)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(8):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, (f, rel_linf(gs[f], os_[f]))
    g.close()
    o.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_one_row_form_bitwise(nuosc_mod, prec):
    """A grid of one x row runs the stage kernel's NOX form alone (x
    neighbours taken as the point itself) and the ordinary form inside a
    batch (x neighbours from the periodic ghost rows, copies of the row):
    the two give the same bits."""
    N = nuosc_mod
    p = N.PREC_FP64 if prec == "fp64" else N.PREC_FP32
    c = cfg_for(1, 40, 4, pmo=0.5)
    init = dict(ipt=5, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5, seed=3)
    a = N.NuOsc(**c, precision=p)
    b = N.NuOsc(**c, precision=p)
    for g in (a, b):
        g.fillInitValue(**init)
    bt = N.Batch(precision=p)
    bt.add(b)
    a.step_rk4(20)
    bt.step(20)
    assert np.array_equal(a.get_state(), b.get_state())
    bt.close()
    a.close()
    b.close()
