"""Pins the oracle's initial data and diagnostics.

Initial data (nuosc_init.cpp:17-137): sum_v w G0 = 1 and sum_v w G0b = alpha,
G0 proportional to the Gaussian of vz (the vx width is infinite), trace
ee + mm = G0, pure states (det rho = 0), the ipt-specific perturbation
profiles in closed form, and the seeded generator's uniform range.

Diagnostics (nuosc_ana.cpp:34-137): an independent numpy formulation through
the polarisation vectors, plus the identities a pure, unperturbed state
satisfies; analysis.dat formatting (nuosc_ana.cpp:126-129).
"""
import numpy as np
import pytest

BOX = dict(nn=2, x0=-0.4, x1=0.4, z0=-0.8, z1=0.8, dx=0.1, dz=0.1, cfl=0.4, ko=0.1, mu=1.0, pmo=0.0)


def sim(O, **over):
    return O.OracleSim(**dict(BOX, **over))


@pytest.mark.parametrize("ipt", [0, 2, 3, 5])
def test_init_identities(oracle_mod, ipt):
    o = sim(oracle_mod)
    alpha, eps0, sigma, lnue, lnueb = 0.85, 0.2, 0.3, 0.6, 0.5
    o.fill_init(ipt=ipt, alpha=alpha, eps0=eps0, sigma=sigma, lnue=lnue, lnueb=lnueb, seed=77)
    vx, vy, vz, vw = o.vgrid()
    G0, G0b = o.G0, o.G0b
    assert abs(np.sum(vw * G0) - 1) < 1e-15
    assert abs(np.sum(vw * G0b) - alpha) < 1e-15
    shape = np.exp(-(vz - 1) ** 2 / (2 * lnue ** 2))
    assert np.allclose(G0 / G0[0], shape / shape[0], rtol=1e-14, atol=0)
    s = o.get_state()
    assert np.max(np.abs(s[0] + s[1] - G0)) < 4e-16
    assert np.max(np.abs(s[4] + s[5] - G0b)) < 4e-16
    # pure states: ee*mm = |rho_ex|^2
    assert np.max(np.abs(s[0] * s[1] - s[2] ** 2 - s[3] ** 2)) < 1e-16
    assert np.max(np.abs(s[4] * s[5] - s[6] ** 2 - s[7] ** 2)) < 1e-16
    Z = (np.arange(o.shape[1]) + 0.5) * 0.1 - 0.8
    tmpr = 2 * s[2] / G0[None, None, :]
    if ipt == 0:
        want = eps0 * np.exp(-Z ** 2 / (2 * sigma ** 2))
        assert np.allclose(tmpr, want[None, :, None], rtol=1e-14, atol=0)
    elif ipt == 2:
        want = eps0 * (1 + np.cos(2 * np.pi * Z / (2 * sigma ** 2))) * 0.5
        assert np.allclose(tmpr, want[None, :, None], rtol=1e-13, atol=1e-17)
    elif ipt == 3:
        assert np.allclose(tmpr, eps0, rtol=1e-15, atol=0)
    else:
        tmpi = 2 * s[3] / G0[None, None, :]
        assert np.all(np.abs(tmpr) <= eps0) and np.all(np.abs(tmpi) <= eps0)
        # independent draws, centred, spread like U[-eps, eps]
        assert abs(np.mean(tmpr)) < 0.1 * eps0 and abs(np.std(tmpr) - eps0 / np.sqrt(3)) < 0.05 * eps0
        assert abs(np.corrcoef(tmpr.ravel(), tmpi.ravel())[0, 1]) < 0.1
        # the antineutrino coherence is the conjugate perturbation
        assert np.allclose(2 * s[7] / G0b, -tmpi, rtol=1e-14, atol=1e-20)
    if ipt != 5:
        assert not s[3].any() and not s[7].any()
    n0 = np.sum(vw[None, None, :] * s[0]) * o.s.ds_L
    # the oracle sums sequentially (as the reference), numpy pairwise
    assert abs(o.n_nue0[0] - n0) < 1e-12 * abs(n0)


def test_seeded_generator(oracle_mod):
    O = oracle_mod
    vals = np.array([O.seed_uniforms(123, k) for k in range(20000)])
    assert np.all(vals >= -1) and np.all(vals < 1)
    assert abs(vals.mean()) < 0.02
    assert O.seed_uniforms(123, 5) == O.seed_uniforms(123, 5)
    assert O.seed_uniforms(123, 5) != O.seed_uniforms(124, 5)


def numpy_analysis(o):
    s = o.get_state()
    vx, vy, vz, vw = o.vgrid()
    G0, G0b = o.G0, o.G0b
    P = np.stack([2 * s[2], -2 * s[3], s[0] - s[1]]) / G0
    Pb = np.stack([2 * s[6], 2 * s[7], s[4] - s[5]]) / G0b
    dP = np.abs(1 - np.linalg.norm(P, axis=0))
    dPb = np.abs(1 - np.linalg.norm(Pb, axis=0))
    W = np.broadcast_to(vw, s[0].shape)
    nor, norb = np.sum(W * G0), np.sum(W * G0b)
    surv, survb = np.sum(W * s[0]) / nor, np.sum(W * s[4]) / norb
    M = np.array([np.sum(W * (s[2] - s[6])), np.sum(W * (-s[3] - s[7])),
                  np.sum(W * 0.5 * (s[0] - s[1] - s[4] + s[5]))])
    n0, n0b = o.n_nue0
    return np.array([o.phy_time, max(dP.max(), dPb.max()), surv, survb,
                     np.sum(W * G0 * dP) / nor, np.sum(W * G0b * dPb) / norb,
                     np.linalg.norm(M) * o.s.ds_L, M[2] * o.s.ds_L,
                     abs(n0 * (1 - surv) - n0b * (1 - survb)) / (n0 + n0b)])


def test_analysis_vs_numpy(oracle_mod):
    o = sim(oracle_mod, pmo=1.0)
    o.fill_init(ipt=5, alpha=0.9, eps0=0.3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=5)
    o.step(30)
    got, status = o.analysis()
    want = numpy_analysis(o)
    assert status == 0
    assert got[0] == want[0]
    for k in range(1, 9):
        assert abs(got[k] - want[k]) <= 1e-13 * abs(want[k]) + 1e-15, k


def test_analysis_identities_unperturbed(oracle_mod):
    o = sim(oracle_mod)
    o.fill_init(ipt=3, alpha=0.9, eps0=0.0, sigma=1.0, lnue=0.6, lnueb=0.5)
    out, status = o.analysis()
    assert status == 0
    assert out[1] < 4e-16          # |P| = 1
    assert out[2] == 1.0 and out[3] == 1.0   # Pee = Peebar = 1
    assert out[8] == 0.0           # ELNe
    # |M0| = |sum w (G0 - G0b)/2| ds_L = (1 - alpha)/2 * nx*nz * ds_L
    nx, nz, _ = o.shape
    assert abs(out[6] - 0.05 * nx * nz * o.s.ds_L) < 1e-14
    assert abs(out[7] - out[6]) < 1e-16


def test_analysis_flags_blowup(oracle_mod):
    o = sim(oracle_mod)
    o.fill_init(ipt=3, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5)
    s = o.get_state()
    s[0] *= 12.0
    o.set_state(s)
    _, status = o.analysis()
    assert status == 1
    s[2, 0, 0, 0] = np.nan
    o.set_state(s)
    _, status = o.analysis()
    assert status == 1


def test_analysis_dat_formatting(oracle_mod):
    o = sim(oracle_mod)
    o.fill_init(ipt=0, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5)
    out, _ = o.analysis()
    out[0] = 0.123456789
    first = o.analysis_line(out)
    second = o.analysis_line(out)
    # the stream's precision(13) is set after the first time value is written
    assert first.startswith("0.123457 ") and second.startswith("0.123456789 ")
    assert first.endswith(" \n") and len(first.split()) == 9
    assert first.split()[1:] == second.split()[1:]
