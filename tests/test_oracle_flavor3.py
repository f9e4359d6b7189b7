"""The oracle's 3-flavour right-hand side (the reference's compile-time
N_FLAVOR_3: 18 fields, nuosc_class.cpp:28-82, 189-277, 543-569), pinned by
formulations that share none of its code.  The reference's own N_FLAVOR_3
build stops at the preprocessor (`#ifdef 0`, nuosc_class.cpp:207), so no run
of it exists to compare with.

* interaction: an independent numpy formulation with 3x3 complex density
  matrices rho = [[ee, emr + i emi, conj(ter + i tei)],
                  [emr - i emi, mm, mtr + i mti],
                  [ter + i tei, mtr - i mti, tt]]
  (rhobar likewise from the b fields), H(v) = mu sum_k w_k (1 - v.v_k)
  (rho_k - conj(rhobar_k)), drho/dt = -i[H, rho] (the same H acts on
  rhobar), for all 18 rows;
* vacuum: 14 of the 18 rows are -i[H0, rho] with H0 built from the
  reference's constants by their names (H0_emu = hemr + i hemi, H0_mutau =
  hmtr + i hmti, H0_taue = hter + i htei); the ee, bee, tt and btt rows are
  not (as written they do not conserve Tr rho; the btt row reads the nu
  fields) and are parity unpinned; the test checks that they differ;
* the vacuum constants: U diag(0, dm21, dm31) U^T - Tr/3 with the standard
  three-angle mixing matrix at the reference's angles (degrees/180, used as
  radians) and splittings (x 1267 x 2);
* with the tau fields zero and pmo = 0 the 8 two-flavour rows equal the
  2-flavour RHS bit for bit and the tau rows are zero;
* RK4 with mu = pmo = 0: a Fourier mode of every field evolves by the exact
  amplification G(z)^n of the advection + Kreiss-Oliger symbol.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle import oracle as O  # noqa: E402

(EE, MM, EMR, EMI, BEE, BMM, BEMR, BEMI, TT, MTR, MTI, TER, TEI, BTT, BMTR, BMTI, BTER, BTEI) = range(18)
NU = (EE, MM, TT, EMR, EMI, MTR, MTI, TER, TEI)
NUB = (BEE, BMM, BTT, BEMR, BEMI, BMTR, BMTI, BTER, BTEI)


def sim(nx=3, nz=4, nn=2, dx=0.1, dz=0.1, ko=0.0, mu=1.0, pmo=0.0, nflavor=3, gx=2, gz=2):
    return O.OracleSim(nn=nn, gx=gx, gz=gz, x0=-nx * dx / 2, x1=nx * dx / 2, z0=-nz * dz / 2, z1=nz * dz / 2,
                       dx=dx, dz=dz, ko=ko, mu=mu, pmo=pmo, nflavor=nflavor)


def to_rho(st, rows):
    ee, mm, tt, emr, emi, mtr, mti, ter, tei = (st[k] for k in rows)
    r = np.zeros(ee.shape + (3, 3), complex)
    r[..., 0, 0], r[..., 1, 1], r[..., 2, 2] = ee, mm, tt
    r[..., 0, 1] = emr + 1j * emi
    r[..., 1, 2] = mtr + 1j * mti
    r[..., 2, 0] = ter + 1j * tei
    r[..., 1, 0] = np.conj(r[..., 0, 1])
    r[..., 2, 1] = np.conj(r[..., 1, 2])
    r[..., 0, 2] = np.conj(r[..., 2, 0])
    return r


def from_rho(d, out, rows):
    vals = (d[..., 0, 0].real, d[..., 1, 1].real, d[..., 2, 2].real, d[..., 0, 1].real, d[..., 0, 1].imag,
            d[..., 1, 2].real, d[..., 1, 2].imag, d[..., 2, 0].real, d[..., 2, 0].imag)
    for k, v in zip(rows, vals):
        out[k] = v


def homogeneous_state(rng, shape, amp=0.3):
    """Random 3-flavour state, the same at every site: advection and KO vanish."""
    nx, nz, nv = shape
    st = amp * rng.standard_normal((18, 1, 1, nv))
    for k in (EE, BEE):
        st[k] += 1.0
    return np.broadcast_to(st, (18, nx, nz, nv)).copy()


def commutator_rhs(st, H):
    """-i[H, rho] and -i[H, rhobar] for a per-point H (..., 3, 3) as 18 rows."""
    out = np.zeros_like(st)
    for rows in (NU, NUB):
        r = to_rho(st, rows)
        from_rho(-1j * (H @ r - r @ H), out, rows)
    return out


@pytest.mark.parametrize("mu", [1.0, 0.37])
def test_interaction_is_the_commutator(mu):
    o = sim(mu=mu)
    rng = np.random.default_rng(11)
    st = homogeneous_state(rng, o.shape)
    vx, vy, vz, vw = o.vgrid()
    D = to_rho(st, NU) - np.conj(to_rho(st, NUB))
    m = [np.einsum("k,xzkab->xzab", w, D) for w in (vw, vw * vx, vw * vy, vw * vz)]
    H = mu * (m[0][:, :, None] - vx[None, None, :, None, None] * m[1][:, :, None]
              - vy[None, None, :, None, None] * m[2][:, :, None] - vz[None, None, :, None, None] * m[3][:, :, None])
    want = commutator_rhs(st, H)
    got = o.calrhs(st)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    o.close()


def test_vacuum_rows():
    o = sim(mu=0.0, pmo=0.8)
    rng = np.random.default_rng(4)
    st = homogeneous_state(rng, o.shape)
    hee, hmm, htt, hemr, hemi, hmtr, hmti, hter, htei = O.vacuum_h3()
    em, mt, te = hemr + 1j * hemi, hmtr + 1j * hmti, hter + 1j * htei
    H0 = np.array([[hee, em, np.conj(te)], [np.conj(em), hmm, mt], [te, np.conj(mt), htt]])
    want = commutator_rhs(st, 0.8 * H0)
    got = o.calrhs(st)
    err = np.max(np.abs(got - want), axis=(1, 2, 3))
    scale = np.max(np.abs(want))
    as_written = {EE, BEE, TT, BTT}
    for f in range(18):
        if f in as_written:
            assert err[f] > 1e-3 * scale, f      # the reference's rows, not the commutator
        else:
            assert err[f] <= 1e-13 * scale, (f, err[f])
    o.close()


def test_vacuum_constants_from_the_mixing_matrix():
    a12, a13, a23 = 33.82 / 180.0, 8.61 / 180.0, 48.3 / 180.0
    dm21, dm31 = 7.39e-5 * 1267.0 * 2.0, 2.5229e-3 * 1267.0 * 2.0

    def rot(i, j, a):
        r = np.eye(3)
        r[i, i] = r[j, j] = np.cos(a)
        r[i, j], r[j, i] = np.sin(a), -np.sin(a)
        return r

    U = rot(1, 2, a23) @ rot(0, 2, a13) @ rot(0, 1, a12)      # delta_cp = 0
    H = U @ np.diag([0.0, dm21, dm31]) @ U.T
    H -= np.trace(H) / 3.0 * np.eye(3)
    hv = O.vacuum_h3()
    # the reference's names: hmtr/hmti hold the (e, tau) element, hter/htei the (mu, tau) one
    want = [H[0, 0], H[1, 1], H[2, 2], H[0, 1], 0.0, H[0, 2], 0.0, H[1, 2], 0.0]
    assert np.allclose(hv, want, rtol=1e-14, atol=1e-15), (hv, want)


def test_reduces_to_two_flavours():
    o2 = sim(nx=5, nz=6, ko=0.1, mu=1.3, nflavor=2)
    o3 = sim(nx=5, nz=6, ko=0.1, mu=1.3, nflavor=3)
    o2.fill_init(ipt=5, alpha=0.9, eps0=0.3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=9)
    st = o2.get_state()
    st = st * (1 + 0.05 * np.random.default_rng(2).standard_normal(st.shape))
    st3 = np.zeros((18,) + st.shape[1:])
    st3[:8] = st
    r2, r3 = o2.calrhs(st), o3.calrhs(st3)
    assert np.array_equal(r3[:8], r2)
    assert not np.any(r3[8:])
    o2.close()
    o3.close()


def test_init_writes_two_flavour_fields_only():
    o = sim(nx=4, nz=5, nflavor=3)
    o.set_state(np.ones((18,) + o.shape))
    o.fill_init(ipt=5, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5, seed=1)
    st = o.get_state()
    o2 = sim(nx=4, nz=5, nflavor=2)
    o2.fill_init(ipt=5, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5, seed=1)
    assert np.array_equal(st[:8], o2.get_state())
    assert not np.any(st[8:])
    o.close()
    o2.close()


@pytest.mark.parametrize("axis", ["z", "x"])
def test_fourier_mode_rk4_all_fields(axis):
    """mu = pmo = 0: each of the 18 fields carries its own Fourier mode along
    one axis; after n RK4 steps it is G(lambda dt)^n times the initial mode,
    lambda = -i v (8 sin t - sin 2t)/(6h) - (ko/h) sin^4(t/2), t = k h."""
    n = 16
    ko, h = 0.2, 0.1
    o = sim(nx=n if axis == "x" else 1, nz=n if axis == "z" else 1, dx=h, dz=h, ko=ko, mu=0.0, pmo=0.0)
    nx, nz, nv = o.shape
    vx, vy, vz, vw = o.vgrid()
    vel = vz if axis == "z" else vx
    coord = np.arange(n)
    modes = [1 + f % 5 for f in range(18)]
    st = np.zeros((18, nx, nz, nv))
    for f, kk in enumerate(modes):
        prof = np.cos(2 * np.pi * kk * coord / n + 0.3 * f)
        st[f] = (prof[None, :, None] if axis == "z" else prof[:, None, None]) * np.ones((nx, nz, nv))
    o.set_state(st)
    steps = 7
    o.step(steps)
    got = o.get_state()
    dt = o.s.dt
    for f, kk in enumerate(modes):
        th = 2 * np.pi * kk / n
        lam = -1j * vel * (8 * np.sin(th) - np.sin(2 * th)) / (6 * h) - (ko / h) * np.sin(th / 2) ** 4
        z = lam * dt
        G = (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24) ** steps
        e = np.exp(1j * (2 * np.pi * kk * coord / n + 0.3 * f))
        want = (G[None, :] * e[:, None]).real                       # (n, nv)
        g = got[f][0] if axis == "z" else got[f][:, 0]
        assert np.max(np.abs(g - want)) <= 1e-13, f
    o.close()


def test_flavour_count_checked():
    with pytest.raises(ValueError):
        sim(nflavor=4)
