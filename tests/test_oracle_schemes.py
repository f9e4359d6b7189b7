"""The oracle's discretisation variants (the reference's compile-time options
KO_ORD_3 / 5th-order KO and ADVEC_*, nuosc_class.cpp:112-183), pinned by what
the mathematics fixes rather than by restating their stencils:

  * 5th-order Kreiss-Oliger: -ko/(64h) delta^6 annihilates polynomials of
    degree <= 5 and maps j^6 to 6! (-ko/(64h)); on the Nyquist mode it has
    the sign of +ko/h (it amplifies: the reference's sign), where the 3rd-order
    default -ko/(16h) delta^4 has -ko/h (it damps);
  * advection schemes: observed order of accuracy of -v d/dz on a sine (CFD2
    and UPWIND2 2, LOPSIDED and the default 4), exact zero for ADVEC_OFF;
  * upwinding side: the one-sided schemes are dissipative, <f, L f> <= 0 for
    random f in every direction (the downwind choice would be >= 0);
  * MUSCL with the superbee limiter: on linear data the x term is exact for
    either sign of vx, and so is the z term for vz > 0; for vz < 0 the z term
    as the reference writes it (nuosc_class.cpp:155-157: the last limited
    slope enters with - where the x term has +) cancels to 0, which the
    mirrored x form would not give; and one value worked by hand from the
    limiter's definition that minmod would not give.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle import oracle as O  # noqa: E402


def column(nz, dz, ko_order=3, adv=0, ko=0.0, nn=2, gz=3):
    """One x cell, nz z cells, mu = 0, no vacuum term: the RHS is advection +
    KO only, acting along z."""
    return O.OracleSim(nn=nn, gx=3, gz=gz, x0=0.0, x1=0.1, z0=0.0, z1=nz * dz, dx=0.1, dz=dz, mu=0.0,
                       pmo=0.0, ko=ko, ko_order=ko_order, adv=adv)


def row(nx, dx, adv=0, nn=4):
    """nx x cells, 4 z cells, mu = 0, no vacuum term, no KO: with data
    constant in z the RHS is the x advection term alone."""
    return O.OracleSim(nn=nn, gx=3, gz=3, x0=0.0, x1=nx * dx, z0=0.0, z1=0.4, dx=dx, dz=0.1, mu=0.0,
                       pmo=0.0, ko=0.0, adv=adv)


def profile_state(o, prof):
    """Every field and direction carries the z profile prof[j]."""
    nx, nz, nv = o.shape
    return np.broadcast_to(prof[None, None, :, None], (8, nx, nz, nv)).copy()


def profile_state_x(o, prof):
    """Every field and direction carries the x profile prof[i]."""
    nx, nz, nv = o.shape
    return np.broadcast_to(prof[None, :, None, None], (8, nx, nz, nv)).copy()


@pytest.mark.parametrize("p", range(7))
def test_ko5_annihilates_low_degree_polynomials(p):
    nz, dz, ko = 24, 0.1, 0.3
    o = column(nz, dz, ko_order=5, adv=O.OracleSim.ADV_OFF, ko=ko)
    j = np.arange(nz, dtype=np.float64)
    st = profile_state(o, (j - 11.5) ** p)
    r = o.calrhs(st)[0, 0]        # (nz, nv)
    inner = r[3:nz - 3]           # away from the periodic wrap
    want = 720.0 * (-ko / dz / 64.0) if p == 6 else 0.0
    scale = max(1.0, np.max(np.abs(st)) * ko / dz)
    assert np.max(np.abs(inner - want)) <= 1e-9 * scale, (p, inner[0, 0], want)
    o.close()


@pytest.mark.parametrize("ko_order,sign", [(3, -1.0), (5, +1.0)])
def test_ko_sign_on_the_nyquist_mode(ko_order, sign):
    nz, dz, ko = 16, 0.1, 0.2
    o = column(nz, dz, ko_order=ko_order, adv=O.OracleSim.ADV_OFF, ko=ko)
    f = (-1.0) ** np.arange(nz)
    r = o.calrhs(profile_state(o, f))[0, 0]
    # delta^4 of (-1)^j = 16 (-1)^j, delta^6 = -64 (-1)^j: both give ko/h in size
    assert np.allclose(r, sign * ko / dz * f[:, None], rtol=1e-12, atol=0)
    o.close()


ORDERS = [(O.OracleSim.ADV_CFD4, 4), (O.OracleSim.ADV_CFD2, 2), (O.OracleSim.ADV_UPWIND2, 2),
          (O.OracleSim.ADV_LOPSIDED, 4)]


@pytest.mark.parametrize("adv,order", ORDERS)
def test_advection_order_of_accuracy(adv, order):
    L = 6.4
    errs = []
    for nz in (64, 128, 256):
        dz = L / nz
        o = column(nz, dz, adv=adv)
        z = (np.arange(nz) + 0.5) * dz
        k = 2 * np.pi / L
        r = o.calrhs(profile_state(o, np.sin(k * z)))[0, 0]
        vz = o.vgrid()[2]
        exact = -vz[None, :] * k * np.cos(k * z)[:, None]
        errs.append(np.max(np.abs(r - exact)))
        o.close()
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > order - 0.15, (adv, errs, rates)
    assert max(rates) < order + 0.6, (adv, errs, rates)


def test_advection_off_is_zero():
    o = column(20, 0.1, adv=O.OracleSim.ADV_OFF)
    rng = np.random.default_rng(3)
    st = rng.standard_normal((8,) + o.shape)
    assert not np.any(o.calrhs(st))
    o.close()


@pytest.mark.parametrize("adv", [O.OracleSim.ADV_UPWIND2, O.OracleSim.ADV_LOPSIDED, O.OracleSim.ADV_MUSCL])
def test_one_sided_schemes_are_upwind(adv):
    """<f, L f> <= 0 per direction: the scheme takes its extra points on the
    side the flow comes from.  Along x for every moving direction; along z
    for every moving direction, except MUSCL's vz < 0 branch, whose z form as
    written is not a flux difference (test_muscl_linear_data)."""
    n, h = 40, 0.1
    rng = np.random.default_rng(8)
    f = rng.standard_normal(n)
    o = column(n, h, adv=adv, nn=4)
    r = o.calrhs(profile_state(o, f))[0, 0]    # (nz, nv)
    vz = o.vgrid()[2]
    energy = np.einsum("j,jv->v", f, r)
    moving = (vz > 1e-12) if adv == O.OracleSim.ADV_MUSCL else (np.abs(vz) > 1e-12)
    assert np.all(energy[moving] < 0), energy[moving].max()
    o.close()
    o = row(n, h, adv=adv)
    r = o.calrhs(profile_state_x(o, f))[0, :, 0]   # (nx, nv)
    vx = o.vgrid()[0]
    energy = np.einsum("i,iv->v", f, r)
    moving = np.abs(vx) > 1e-12
    assert np.all(energy[moving] < 0), energy[moving].max()
    o.close()


def test_muscl_linear_data():
    """Slope s: the x term is -vx s for either sign of vx (the limited slopes
    of a linear profile are s); the z term is -vz s for vz > 0, and for vz < 0
    it is 0 as the reference writes it: x1 - phi(s, s)/2 - x0 - phi(s, s)/2
    = s - s/2 - s/2 (the x form has + phi/2 there and gives s)."""
    n, h, slope = 30, 0.1, 0.7
    o = column(n, h, adv=O.OracleSim.ADV_MUSCL, nn=4)
    r = o.calrhs(profile_state(o, slope * np.arange(n) * h))[0, 0]
    vz = o.vgrid()[2]
    want = np.where(vz > 0, -vz * slope, 0.0)
    assert np.allclose(r[3:n - 3], want[None, :], rtol=1e-12, atol=1e-12)
    assert np.any(vz < -0.5)   # the vz < 0 branch is exercised
    o.close()
    o = row(n, h, adv=O.OracleSim.ADV_MUSCL)
    r = o.calrhs(profile_state_x(o, slope * np.arange(n) * h))[0, :, 0]
    vx = o.vgrid()[0]
    assert np.allclose(r[3:n - 3], -vx[None, :] * slope, rtol=1e-12, atol=1e-12)
    assert np.any(vx < -0.5)
    o.close()


def test_muscl_superbee_value():
    """z profile ... 0, 1, 2, 5 ... around cell j = 10 (left slopes 1, 1; right
    slope 3).  For v > 0 the upwind difference is
      x0 + phi(x0 - x_-1, x_1 - x0)/2 - x_-1 - phi(x_-1 - x_-2, x0 - x_-1)/2
      = 2 + phi(1, 3)/2 - 1 - phi(1, 1)/2,
    with superbee phi(a, b) = max(min(2a, b), min(a, 2b)) (a, b > 0):
    phi(1, 3) = 2, phi(1, 1) = 1, so 1.5 (minmod would give 1)."""
    nz, dz = 24, 0.1
    o = column(nz, dz, adv=O.OracleSim.ADV_MUSCL, nn=4)
    prof = np.zeros(nz)
    prof[8:12] = [0.0, 1.0, 2.0, 5.0]
    prof[12:] = 5.0
    r = o.calrhs(profile_state(o, prof))[0, 0]
    vz = o.vgrid()[2]
    pos = vz > 1e-12
    assert np.allclose(r[10, pos], -vz[pos] / dz * 1.5, rtol=1e-12, atol=0)
    o.close()


def test_scheme_rejects_narrow_ghosts():
    with pytest.raises(ValueError):
        O.OracleSim(nn=2, gx=2, gz=2, ko_order=5)
    with pytest.raises(ValueError):
        O.OracleSim(nn=2, gx=2, gz=2, adv=O.OracleSim.ADV_LOPSIDED)
    O.OracleSim(nn=2, gx=2, gz=2, adv=O.OracleSim.ADV_MUSCL).close()
