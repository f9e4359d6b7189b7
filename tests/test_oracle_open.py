"""The oracle's open z boundary, pinned by what the mathematics fixes.

The reference's open boundary (updateInjetOpenBoundary,
nuosc_boundary.cpp:54-57) is a stub, so nothing of the reference pins this
path; these checks do:
  * mu = 0, pmo = 0: every field free-streams, f(z, t) = f0(z - vz t), and the
    inflow ghosts carry exactly that continuation of the ipt-0 profile.  With
    a pulse that enters through z0 and leaves through z1 (and one leaving
    through z0), the interior error against the exact solution converges at
    third order in dz (the order of the Kreiss-Oliger term), which a wrong
    inflow time, sign or extrapolation weight breaks;
  * quadratic profiles are reproduced exactly by the outflow ghosts;
  * the z-sum of the mu = 0 right-hand side telescopes to the boundary flux
    of the stencil (the ghost and edge cells only);
  * a flavour-unstable open run (mu = 1) stays finite and near |P| = 1.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle import oracle as O  # noqa: E402

COL = dict(nn=2, gx=2, x0=0.0, x1=0.1, dx=0.1, cfl=0.4, ko=0.1, pmo=0.0, bc=1)


def exact_state(o, eps0, sigma, zc, t):
    """Free-streaming ipt-0 state at time t, interior layout (8, 1, nz, nv)."""
    vx, vy, vz, vw = o.vgrid()
    Z = np.ctypeslib.as_array(o.s.Z, (o.s.nz,)).copy()
    zz = Z[:, None] - vz[None, :] * t - zc
    a = eps0 * np.exp(-zz * zz / (2.0 * sigma * sigma))
    p3o = np.sqrt(1.0 - a * a)
    G0, G0b = o.G0[None, :], o.G0b[None, :]
    st = np.zeros((8, 1) + a.shape)
    st[0, 0] = 0.5 * G0 * (1 + p3o)
    st[1, 0] = 0.5 * G0 * (1 - p3o)
    st[2, 0] = 0.5 * G0 * a
    st[4, 0] = 0.5 * G0b * (1 + p3o)
    st[5, 0] = 0.5 * G0b * (1 - p3o)
    st[6, 0] = 0.5 * G0b * a
    return st


def free_stream_error(dz, T=8.0, eps0=0.5, sigma=1.0, zc=-3.0, L=6.4):
    o = O.OracleSim(z0=-L, z1=L, dz=dz, mu=0.0, zc=zc, gz=2, **COL)
    o.fill_init(ipt=0, alpha=0.9, eps0=eps0, sigma=sigma, lnue=0.6, lnueb=0.5)
    n = int(round(T / o.s.dt))
    o.step(n)
    err = np.max(np.abs(o.get_state() - exact_state(o, eps0, sigma, zc, o.phy_time)))
    o.close()
    return err


@pytest.mark.parametrize("zc", [-3.0, -9.0])
def test_open_free_streaming_converges_third_order(zc):
    """zc = -3: the pulse starts inside and leaves through both ends
    (outflow); zc = -9: it starts outside and enters through z0 (inflow)."""
    errs = [free_stream_error(dz, zc=zc) for dz in (0.2, 0.1, 0.05)]
    rates = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    print("open-boundary free-streaming errors", errs, "rates", rates)
    assert errs[-1] < 2e-4
    assert min(rates) > 2.6


def test_open_wrong_inflow_time_is_detected():
    """Planted mistake (on the CPU): boundary data frozen at t = 0 keeps the
    entering pulse out, an O(eps0 G0) error, so the convergence check above
    has teeth."""
    o = O.OracleSim(z0=-6.4, z1=6.4, dz=0.1, mu=0.0, zc=-9.0, gz=2, **COL)
    o.fill_init(ipt=0, alpha=0.9, eps0=0.5, sigma=1.0, lnue=0.6, lnueb=0.5)
    # RK4 by hand on calrhs, whose boundary fill uses phy_time (= 0 here)
    dt = o.s.dt
    u = o.get_state()
    t0 = 0.0
    for _ in range(200):
        k1 = o.calrhs(u)
        k2 = o.calrhs(u + 0.5 * dt * k1)
        k3 = o.calrhs(u + 0.5 * dt * k2)
        k4 = o.calrhs(u + dt * k3)
        u = u + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        t0 += dt
    err = np.max(np.abs(u - exact_state(o, 0.5, 1.0, -9.0, t0)))
    assert err > 1e-2
    o.close()


def test_open_outflow_ghosts_reproduce_quadratics():
    o = O.OracleSim(z0=-1.6, z1=1.6, dz=0.1, mu=0.0, zc=0.0, gz=3, **COL)
    o.fill_init(ipt=3, alpha=0.9, eps0=0.0, sigma=1.0, lnue=0.6, lnueb=0.5)
    nx, nz, nv = o.shape
    vz = o.vgrid()[2]
    j = np.arange(nz, dtype=np.float64)
    rng = np.random.default_rng(3)
    c = rng.uniform(-1, 1, (8, 3, nv))
    st = np.zeros((8, 1, nz, nv))
    for f in range(8):
        st[f, 0] = c[f, 0] + c[f, 1] * j[:, None] + c[f, 2] * (j[:, None] ** 2)
    o.set_state(st, buf=0)
    o.fill_ghosts(0, t=0.3)
    g = o.ghosted(0)           # (8, nx+2gx, nz+2gz, nv)
    gz = 3
    for q in range(1, gz + 1):
        for side, jj in ((0, -q), (1, nz - 1 + q)):
            out = (vz <= 0) if side == 0 else (vz >= 0)
            want = c[:, 0] + c[:, 1] * jj + c[:, 2] * jj * jj
            got = g[:, 2, jj + gz, :]
            assert np.allclose(got[:, out], want[:, out], rtol=0, atol=1e-11 * max(1.0, jj * jj))
    o.close()


def test_open_rhs_sum_is_boundary_flux():
    o = O.OracleSim(z0=-1.6, z1=1.6, dz=0.1, mu=0.0, zc=-2.0, gz=2, **COL)
    o.fill_init(ipt=0, alpha=0.9, eps0=0.3, sigma=0.7, lnue=0.6, lnueb=0.5)
    o.step(3)
    nx, nz, nv = o.shape
    st = o.get_state() + 1e-3 * np.random.default_rng(5).uniform(-1, 1, (8, nx, nz, nv))
    rhs = o.calrhs(st)
    g = o.ghosted(O.OracleSim.BUF_PRE)[:, 2]   # (8, nz+4, nv): ghosts filled by calrhs
    vz = o.vgrid()[2]
    dz, ko = o.s.dz, o.s.ko
    fz = -vz / (12 * dz)
    kz = -ko / dz / 16.0
    F = lambda j: g[:, j + 2, :]  # noqa: E731
    n = nz
    s_adv = (F(-2) + F(-1) + F(0) + F(1)) - (F(n - 2) + F(n - 1) + F(n) + F(n + 1)) \
        - 8.0 * ((F(-1) + F(0)) - (F(n - 1) + F(n)))
    s_ko = (F(n + 1) - 3 * F(n) + 3 * F(n - 1) - F(n - 2)) - (F(1) - 3 * F(0) + 3 * F(-1) - F(-2))
    want = fz[None, :] * s_adv + kz * s_ko
    got = rhs[:, 0].sum(axis=1)
    assert np.allclose(got, want, rtol=1e-9, atol=1e-12)
    o.close()


@pytest.mark.parametrize("gz", [2, 3])
def test_open_unstable_run_stays_bounded(gz):
    o = O.OracleSim(z0=-6.4, z1=6.4, dz=0.1, mu=1.0, zc=-8.0, gz=gz, **COL)
    o.fill_init(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    for _ in range(10):
        o.step(40)
        ana, status = o.analysis()
        assert status == 0
        assert np.isfinite(ana).all()
        assert ana[1] < 1e-4   # max |P|-1
    o.close()


def test_open_rejects_short_columns():
    with pytest.raises(ValueError):
        O.OracleSim(z0=0.0, z1=0.2, dz=0.1, mu=0.0, gz=2, **COL)
    with pytest.raises(ValueError):
        O.OracleSim(z0=-1.0, z1=1.0, dz=0.1, mu=0.0, gz=4, **COL)


def test_ipt0_profile_centre():
    """zc shifts the ipt-0 profile; zc = 0 is the reference's profile."""
    a = O.OracleSim(z0=-1.6, z1=1.6, dz=0.1, mu=0.0, zc=0.0, gz=2, **COL)
    b = O.OracleSim(z0=-1.6, z1=1.6, dz=0.1, mu=0.0, zc=0.5, gz=2, **COL)
    for o in (a, b):
        o.fill_init(ipt=0, alpha=0.9, eps0=0.2, sigma=0.4, lnue=0.6, lnueb=0.5)
    sa, sb = a.get_state(), b.get_state()
    # Z[16] - 0.5 = Z[11]: the profile moves by 5 rows
    assert np.allclose(sa[2, 0, 11], sb[2, 0, 16], rtol=1e-12, atol=0)
    assert not np.allclose(sa[2, 0, 11], sb[2, 0, 11], rtol=1e-3, atol=0)
    a.close()
    b.close()
