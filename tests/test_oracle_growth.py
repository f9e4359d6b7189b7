"""Pins the coupled interaction + advection + Kreiss-Oliger + RK4 path of the
oracle in the linear regime: started on the fastest-growing eigenvector of a
z Fourier mode (tests/growth.py builds the linearised operator of
nuosc_class.cpp:190-200 with the semi-discrete advection/KO symbols), the
measured growth rate of |rho_ex| must equal the eigenvalue's real part."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import growth as G  # noqa: E402


@pytest.mark.parametrize("m,mu", [(3, 1.0), (5, 1.0), (2, 0.7)])
def test_linear_growth_rate(oracle_mod, m, mu):
    o = oracle_mod.OracleSim(nn=4, x0=0, x1=0.1, z0=-3.2, z1=3.2, dx=0.1, dz=0.1, cfl=0.4, ko=0.1, mu=mu,
                             pmo=0.0)
    o.fill_init(ipt=3, alpha=0.9, eps0=0.0, sigma=1.0, lnue=0.6, lnueb=0.5)
    vg = o.vgrid()
    nz = o.shape[1]
    Z = (np.arange(nz) + 0.5) * 0.1 - 3.2
    st, gamma = G.mode_eigenstate(vg, o.G0, o.G0b, mu, m, Z, 0.1, 0.1, 1e-6)
    assert gamma > 0.02          # an unstable mode (fast flavour instability)
    o.set_state(st)
    ts, rs = [], []
    for _ in range(12):
        s = o.get_state()
        ts.append(o.phy_time)
        rs.append(np.sum(vg[3] * (s[2] ** 2 + s[3] ** 2)))
        o.step(25)
    rate = G.fitted_rate(ts, rs)
    assert abs(rate - gamma) <= 1e-5 * gamma, (rate, gamma)


def test_growth_needs_the_interaction(oracle_mod):
    """With mu = 0 the same state only decays (KO) -- the growth above comes
    from the interaction term, not from the stencil."""
    o = oracle_mod.OracleSim(nn=4, x0=0, x1=0.1, z0=-3.2, z1=3.2, dx=0.1, dz=0.1, cfl=0.4, ko=0.1, mu=0.0,
                             pmo=0.0)
    o.fill_init(ipt=3, alpha=0.9, eps0=0.0, sigma=1.0, lnue=0.6, lnueb=0.5)
    vg = o.vgrid()
    Z = (np.arange(64) + 0.5) * 0.1 - 3.2
    st, _ = G.mode_eigenstate(vg, o.G0, o.G0b, 1.0, 3, Z, 0.1, 0.1, 1e-6)
    o.set_state(st)
    s0 = o.get_state()
    o.step(100)
    s1 = o.get_state()
    r0 = np.sum(vg[3] * (s0[2] ** 2 + s0[3] ** 2))
    r1 = np.sum(vg[3] * (s1[2] ** 2 + s1[3] ** 2))
    assert r1 <= r0
