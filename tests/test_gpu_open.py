"""GPU parity of the open z boundary (x periodic) against the CPU oracle.

The oracle's open boundary is pinned on the CPU by the free-streaming
closed form (third-order convergence with inflow and outflow active), the
exactness of the outflow extrapolation on quadratics and the telescoped
boundary flux (tests/test_oracle_open.py).  Bars here:
  * calRHS at a nonzero time: relative L-inf <= 1e-12 per field;
  * ghosted download at a nonzero time equals the oracle's fill to 1e-14;
  * 200 RK4 steps (mu = 1, a pulse entering through z0): relative L-inf
    <= 1e-10 per field; Pee, |M0| relative 1e-10; |P|-1 absolute 1e-12 +
    relative 1e-10;
  * graph replay and eager launches give identical bits (the boundary time is
    read on the device);
  * the 4096-cell, 642-direction column with gz = 3: 20 steps against the
    oracle at the same bar;
  * FP32 form: Pee within 1e-6 and max |P|-1 below 1e-5 of the FP64 run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COL = dict(nv=2, x0=0.0, x1=0.1, z0=-6.4, z1=6.4, dx=0.1, dz=0.1, CFL=0.4, ko=0.1, mu=1.0, pmo=0.0)
INIT = dict(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)


def oracle_sim(O, c, zc, gz=2):
    return O.OracleSim(nn=c["nv"], gz=gz, x0=c["x0"], x1=c["x1"], z0=c["z0"], z1=c["z1"], dx=c["dx"],
                       dz=c["dz"], cfl=c["CFL"], ko=c["ko"], mu=c["mu"], pmo=c["pmo"], bc=1, zc=zc)


def gpu_sim(N, c, zc, gz=2, **kw):
    return N.NuOsc(gz=gz, bc=N.BC_OPEN_Z, zc=zc, **dict(c, **kw))


def rel_linf(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def test_open_calrhs_and_ghosts(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    c = dict(COL, x1=0.4)          # 4 x 128: x wrap and open z together
    o, g = oracle_sim(O, c, -7.0), gpu_sim(N, c, -7.0)
    init = dict(INIT, eps0=0.3)
    o.fill_init(**init)
    g.fillInitValue(**init)
    o.step(30)
    st = o.get_state()
    t = o.phy_time
    g.set_state(st)
    g.set_phy_time(t)
    r_o = o.calrhs(st)
    r_g = g.calRHS(st)
    for f in range(8):
        assert rel_linf(r_g[f], r_o[f]) <= 1e-12, f
    o.fill_ghosts(0, t=t)
    go = o.ghosted(0)
    for f in range(8):
        gg = g.get_field_ghosted(f)
        assert gg.shape == go[f].shape
        assert np.allclose(gg, go[f], rtol=1e-14, atol=1e-300), f


@pytest.mark.parametrize("nx_len", [0.1, 0.8])
def test_open_200_steps_match_oracle(oracle_mod, nuosc_mod, nx_len):
    O, N = oracle_mod, nuosc_mod
    c = dict(COL, x1=nx_len)
    o, g = oracle_sim(O, c, -8.0), gpu_sim(N, c, -8.0)
    o.fill_init(**INIT)
    g.fillInitValue(**INIT)
    a0 = np.max(np.abs(o.get_state()[2]))
    for blk in range(10):
        o.step(20)
        g.step_rk4(20)
        ao, so = o.analysis()
        ag = g.analysis()
        assert so == 0 and ag.status == 0
        assert ag.phy_time == ao[0]
        assert abs(ag.surv - ao[2]) <= 1e-10 * abs(ao[2])
        assert abs(ag.aM0 - ao[6]) <= 1e-10 * abs(ao[6]) + 1e-15
        assert abs(ag.maxdP - ao[1]) <= 1e-12 + 1e-10 * abs(ao[1])
    gs, os_ = g.get_state(), o.get_state()
    for f in range(8):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, f
    # the pulse entered: the coherence grew well above its initial tail
    assert np.max(np.abs(os_[2])) > 3 * a0


def test_open_graph_equals_eager(nuosc_mod):
    N = nuosc_mod
    a = gpu_sim(N, COL, -8.0, use_graph=True)
    b = gpu_sim(N, COL, -8.0, use_graph=False)
    for s in (a, b):
        s.fillInitValue(**INIT)
        s.step_rk4(7)
        s.step_rk4(13)
    assert a.phy_time == b.phy_time
    assert np.array_equal(a.get_state(), b.get_state())


def test_open_column_4096_gz3(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    c = dict(nv=8, x0=-0.025, x1=0.025, z0=-102.4, z1=102.4, dx=0.05, dz=0.05, CFL=0.4, ko=0.1, mu=1.0,
             pmo=0.0)
    zc = -102.4 - 3.0
    o, g = oracle_sim(O, c, zc, gz=3), gpu_sim(N, c, zc, gz=3)
    assert g.shape == o.shape == (1, 4096, 642)
    init = dict(INIT, eps0=1e-6)
    o.fill_init(**init)
    g.fillInitValue(**init)
    o.step(20)
    g.step_rk4(20)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(8):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, f


def test_open_fp32_against_fp64(nuosc_mod):
    N = nuosc_mod
    a = gpu_sim(N, COL, -8.0)
    b = gpu_sim(N, COL, -8.0, precision=N.PREC_FP32)
    for s in (a, b):
        s.fillInitValue(**INIT)
    for blk in range(10):
        a.step_rk4(20)
        b.step_rk4(20)
        x, y = a.analysis(), b.analysis()
        assert abs(y.surv - x.surv) <= 1e-6
        assert y.maxdP <= 1e-5
