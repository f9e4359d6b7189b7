"""Pins the oracle's calRHS, ghost fill and RK4 against formulations that do
not share its code:

* calRHS vs an independent numpy formulation: 2x2 complex density matrices,
  H(v) = mu sum_k w_k (1 - v.v_k)(rho_k - conj(rhobar_k)) + pmo (sin2t s1 - cos2t s3),
  drho/dt = -i[H, rho] (the same H acts on rhobar), plus -v.grad with the
  textbook 4th-order central derivative and Kreiss-Oliger dissipation
  -eps/(16h) delta^4 (order 2r = 4), both written with periodic np.roll;
* a single Fourier mode: the semi-discrete symbol and the exact RK4
  amplification G(z) = 1 + z + z^2/2 + z^3/6 + z^4/24;
* vacuum precession: RK4 = the 4th-order Taylor polynomial of the linear
  commutator map, applied n times;
* conservation of sum_xz (ee + mm) per direction (periodic, any mu);
* ghost band: brute-force modular-index check.
"""
import numpy as np
import pytest

THETA = 37 * np.pi / 180.0
CT, ST = np.cos(2 * THETA), np.sin(2 * THETA)
S1 = np.array([[0, 1], [1, 0]], dtype=complex)
S3 = np.array([[1, 0], [0, -1]], dtype=complex)


def small_sim(O, nx=8, nz=10, nn=2, dx=0.1, dz=0.1, ko=0.1, mu=1.0, pmo=0.0, cfl=0.4):
    return O.OracleSim(nn=nn, x0=-nx * dx / 2, x1=nx * dx / 2, z0=-nz * dz / 2, z1=nz * dz / 2, dx=dx, dz=dz,
                       cfl=cfl, ko=ko, mu=mu, pmo=pmo)


def to_rho(s):
    ee, mm, emr, emi, bee, bmm, bemr, bemi = s
    rho = np.empty(ee.shape + (2, 2), complex)
    rho[..., 0, 0], rho[..., 1, 1] = ee, mm
    rho[..., 0, 1], rho[..., 1, 0] = emr + 1j * emi, emr - 1j * emi
    rhob = np.empty_like(rho)
    rhob[..., 0, 0], rhob[..., 1, 1] = bee, bmm
    rhob[..., 0, 1], rhob[..., 1, 0] = bemr + 1j * bemi, bemr - 1j * bemi
    return rho, rhob


def from_rho(rho, rhob):
    return np.stack([rho[..., 0, 0].real, rho[..., 1, 1].real, rho[..., 0, 1].real, rho[..., 0, 1].imag,
                     rhob[..., 0, 0].real, rhob[..., 1, 1].real, rhob[..., 0, 1].real, rhob[..., 0, 1].imag])


def d1_central4(f, axis, h):
    return (8 * (np.roll(f, -1, axis) - np.roll(f, 1, axis)) - (np.roll(f, -2, axis) - np.roll(f, 2, axis))) / (12 * h)


def ko_delta4(f, axis, h, eps):
    d4 = np.roll(f, 2, axis) + np.roll(f, -2, axis) - 4 * (np.roll(f, 1, axis) + np.roll(f, -1, axis)) + 6 * f
    return -eps / (16 * h) * d4


def numpy_rhs(state, vg, dx, dz, ko, mu, pmo):
    vx, vy, vz, vw = vg
    rho, rhob = to_rho(state)
    D = rho - np.conj(rhob)                     # (x, z, k, 2, 2)
    m0 = np.einsum("k,xzkab->xzab", vw, D)
    mx = np.einsum("k,xzkab->xzab", vw * vx, D)
    my = np.einsum("k,xzkab->xzab", vw * vy, D)
    mz = np.einsum("k,xzkab->xzab", vw * vz, D)
    H = mu * (m0[:, :, None] - vx[None, None, :, None, None] * mx[:, :, None]
              - vy[None, None, :, None, None] * my[:, :, None] - vz[None, None, :, None, None] * mz[:, :, None])
    H = H + pmo * (ST * S1 - CT * S3)
    out = from_rho(-1j * (H @ rho - rho @ H), -1j * (H @ rhob - rhob @ H))
    for f in range(8):
        out[f] += -vx[None, None, :] * d1_central4(state[f], 0, dx) - vz[None, None, :] * d1_central4(state[f], 1, dz)
        out[f] += ko_delta4(state[f], 0, dx, ko) + ko_delta4(state[f], 1, dz, ko)
    return out


def random_state(rng, shape, amp=0.3):
    """A positive-ish random state with O(1) diagonal and O(amp) coherences."""
    st = np.empty((8,) + shape)
    st[0] = 1 + 0.2 * rng.random(shape)
    st[1] = 0.3 * rng.random(shape)
    st[4] = 0.9 + 0.2 * rng.random(shape)
    st[5] = 0.3 * rng.random(shape)
    for f in (2, 3, 6, 7):
        st[f] = amp * rng.standard_normal(shape)
    return st


@pytest.mark.parametrize("mu,pmo,ko", [(1.0, 0.0, 0.1), (0.7, 1.0, 0.0), (0.0, -1.0, 0.3), (2.0, 0.5, 0.05)])
def test_calrhs_vs_numpy_commutator_form(oracle_mod, mu, pmo, ko):
    o = small_sim(oracle_mod, nx=6, nz=7, nn=2, dx=0.13, dz=0.07, ko=ko, mu=mu, pmo=pmo)
    rng = np.random.default_rng(5)
    st = random_state(rng, o.shape)
    got = o.calrhs(st)
    want = numpy_rhs(st, o.vgrid(), o.s.dx, o.s.dz, ko, mu, pmo)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-13 * scale


def test_calrhs_mutation_is_caught(oracle_mod):
    """A plausible slip (the bar mm row using +Ibee) would be visible."""
    o = small_sim(oracle_mod, nx=4, nz=5, nn=1, mu=1.0, pmo=0.0)
    rng = np.random.default_rng(1)
    st = random_state(rng, o.shape)
    want = numpy_rhs(st, o.vgrid(), o.s.dx, o.s.dz, 0.1, 1.0, 0.0)
    got = o.calrhs(st)
    bad = got.copy()
    bad[5] = 2 * got[4] - got[5]   # flip the sign of the interaction term only
    assert np.max(np.abs(bad - want)) > 1e-3


@pytest.mark.parametrize("axis,k,method", [(1, 3, "rk4"), (0, 2, "rk4"), (1, 1, "rk4"),
                                           (1, 3, "srk3"), (0, 2, "srk3")])
def test_fourier_mode_rk4_amplification(oracle_mod, axis, k, method):
    """A Fourier mode of every field is an exact eigenvector of ADV+KO; one
    step multiplies it by the stability polynomial of the integrator:
    RK4 1 + z + z^2/2 + z^3/6 + z^4/24, SSP-RK3 1 + z + z^2/2 + z^3/6."""
    nx, nz = 12, 16
    ko, dx, dz, cfl = 0.2, 0.1, 0.1, 0.4
    o = small_sim(oracle_mod, nx=nx, nz=nz, nn=2, dx=dx, dz=dz, ko=ko, mu=0.0, pmo=0.0, cfl=cfl)
    vx, vy, vz, vw = o.vgrid()
    n = (nx, nz)[axis]
    h = (dx, dz)[axis]
    vel = (vx, vz)[axis]
    kk = 2 * np.pi * k / (n * h)
    th = kk * h
    coord = (np.arange(n) + 0.5) * h
    # every field carries the mode with its own phase
    st = np.zeros((8, nx, nz, o.shape[2]))
    for f in range(8):
        ph = 0.3 * f
        prof = np.cos(kk * coord + ph)
        st[f] = prof[:, None, None] if axis == 0 else prof[None, :, None]
    o.set_state(st)
    nsteps = 25
    if method == "rk4":
        o.step(nsteps)
    else:
        o.step_srk3(nsteps)
    lam = -1j * vel * (8 * np.sin(th) - np.sin(2 * th)) / (6 * h) - (ko / h) * np.sin(th / 2) ** 4
    z = lam * o.s.dt
    poly = 1 + z + z ** 2 / 2 + z ** 3 / 6 + (z ** 4 / 24 if method == "rk4" else 0)
    G = poly ** nsteps
    got = o.get_state()
    for f in range(8):
        ph = 0.3 * f
        e = np.exp(1j * (kk * coord + ph))
        want = (e[:, None] * G[None, :]).real
        want = want[:, None, :] if axis == 0 else want[None, :, :]
        assert np.max(np.abs(got[f] - want)) < 1e-13, f
    assert abs(o.phy_time - nsteps * o.s.dt) < 1e-15 * nsteps


def test_vacuum_precession_rk4_polynomial(oracle_mod):
    pmo, nsteps = 1.3, 40
    o = small_sim(oracle_mod, nx=2, nz=3, nn=1, mu=0.0, pmo=pmo, ko=0.5)
    nv = o.shape[2]
    rng = np.random.default_rng(9)
    # homogeneous state: advection and dissipation vanish
    x0 = rng.standard_normal((8, nv))
    st = np.broadcast_to(x0[:, None, None, :], (8,) + o.shape).copy()
    o.set_state(st)
    o.step(nsteps)

    def L(vec4):  # -i[H_vac, rho] on (ee, mm, emr, emi)
        rho = np.array([[vec4[0], vec4[2] + 1j * vec4[3]], [vec4[2] - 1j * vec4[3], vec4[1]]])
        H = pmo * (ST * S1 - CT * S3)
        d = -1j * (H @ rho - rho @ H)
        return np.array([d[0, 0].real, d[1, 1].real, d[0, 1].real, d[0, 1].imag])

    A = np.stack([L(e) for e in np.eye(4)], axis=1)
    hA = o.s.dt * A
    P = np.eye(4) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    Pn = np.linalg.matrix_power(P, nsteps)
    got = o.get_state()
    for base in (0, 4):
        want = Pn @ x0[base:base + 4]
        for q in range(4):
            assert np.max(np.abs(got[base + q] - want[q][None, None, :])) < 1e-13


@pytest.mark.parametrize("pmo", [0.0, 1.0])
def test_trace_conserved_per_direction(oracle_mod, pmo):
    o = small_sim(oracle_mod, nx=10, nz=12, nn=2, mu=1.0, pmo=pmo, ko=0.1)
    rng = np.random.default_rng(2)
    st = random_state(rng, o.shape, amp=0.05)
    o.set_state(st)
    tr0 = (st[0] + st[1]).sum(axis=(0, 1))
    trb0 = (st[4] + st[5]).sum(axis=(0, 1))
    o.step(10)
    s = o.get_state()
    assert np.max(np.abs((s[0] + s[1]).sum(axis=(0, 1)) - tr0)) < 1e-12 * np.max(tr0)
    assert np.max(np.abs((s[4] + s[5]).sum(axis=(0, 1)) - trb0)) < 1e-12 * np.max(trb0)


@pytest.mark.parametrize("nx,nz,g", [(5, 7, 2), (3, 4, 3), (1, 6, 2), (6, 1, 2)])
def test_ghost_fill_brute_force(oracle_mod, nx, nz, g):
    O = oracle_mod
    o = O.OracleSim(nn=1, gx=g, gz=g, x0=0, x1=nx * 0.1, z0=0, z1=nz * 0.1, dx=0.1, dz=0.1)
    assert o.shape[:2] == (nx, nz)
    rng = np.random.default_rng(4)
    st = rng.standard_normal((8,) + o.shape)
    o.set_state(st)
    o.fill_ghosts(0)
    gh = o.ghosted(0)
    for i in range(-g, nx + g):
        for j in range(-g, nz + g):
            xin, zin = 0 <= i < nx, 0 <= j < nz
            cell = gh[:, i + g, j + g, :]
            if xin or zin:
                assert np.array_equal(cell, st[:, i % nx, j % nz, :]), (i, j)
            else:
                assert not cell.any(), (i, j)
