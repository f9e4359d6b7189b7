"""GPU parity of the 3-flavour path (18 fields; the reference's N_FLAVOR_3,
nuosc_class.cpp:28-82, 189-277, 543-569) against the oracle:

  * one calRHS, FP64, over shapes that leave x quads, z tiles and the last
    64-direction chunk partly filled: relative L-inf <= 1e-12 per field;
  * 30 RK4 steps, FP64, from a state with all 18 fields nonzero:
    relative L-inf <= 1e-10 per field; the analysis line agrees to 1e-12;
  * 12 SSP-RK3 steps, FP64: relative L-inf <= 1e-10 per field;
  * open z (one x row and a 6-row box), 20 RK4 steps: relative L-inf <=
    1e-10 per field;
  * calRHS of the FP32 form against FP64 on a state exactly representable in
    FP32: within 64 eps32 max|field| / h (FP32 rounding of the stencil);
  * 50 RK4 steps of the FP32 form against FP64: Pee and fields within
    4 * 50 * 2^-24 * 2 = 2.4e-5 (of max|field|): one FP32 rounding per stage
    store, never cancelling, amplified at most twofold;
  * the FP32 form on open z (6 x rows, RK4) and with SSP-RK3, 20 steps each,
    against FP64 under the same kind of bound;
  * renormalising runs, 10 steps: relative L-inf <= 1e-10 per field;
  * hipGraph replay bitwise equal to eager launches; per-stage timing;
  * the CLI with --nflavor 3, and the arguments the path refuses.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cfg_for(nx, nz, nn=2, dx=0.1, ko=0.1, pmo=0.7, mu=1.0, **over):
    c = dict(nv=nn, x0=-nx * dx / 2, x1=nx * dx / 2, z0=-nz * dx / 2, z1=nz * dx / 2, dx=dx, dz=dx, CFL=0.4,
             ko=ko, mu=mu, pmo=pmo)
    c.update(over)
    return c


def oracle_sim(O, c, **over):
    return O.OracleSim(nn=c["nv"], x0=c["x0"], x1=c["x1"], z0=c["z0"], z1=c["z1"], dx=c["dx"], dz=c["dz"],
                       cfl=c["CFL"], ko=c["ko"], mu=c["mu"], pmo=c["pmo"], nflavor=3, **over)


def rel_linf(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def full_state(o, seed, amp=0.05):
    """2-flavour seeded init for fields 0..7, small random tau fields."""
    o.fill_init(ipt=5, alpha=0.9, eps0=0.3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=seed)
    st = o.get_state()
    rng = np.random.default_rng(seed)
    st[:8] *= 1 + 0.05 * rng.standard_normal(st[:8].shape)
    st[8:] = amp * rng.standard_normal(st[8:].shape) * np.max(np.abs(st[0]))
    return st


SHAPES = [(1, 5), (3, 9), (5, 6), (13, 11), (7, 33)]


# 642 directions (nn = 8) on the small shapes only
CALRHS3_CASES = [(nx, nz, nn) for nx, nz in SHAPES for nn in (2, 4, 8) if nn < 8 or nx * nz <= 60]


@pytest.mark.parametrize("nx,nz,nn", CALRHS3_CASES)
def test_calrhs3_fp64(oracle_mod, nuosc_mod, nx, nz, nn):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(nx, nz, nn)
    o = oracle_sim(O, c)
    g = N.NuOsc(**c, nflavor=3)
    assert g.nvar == 18
    st = full_state(o, 10 * nx + nz + nn)
    ro, rg = o.calrhs(st), g.calRHS(st)
    for f in range(18):
        assert rel_linf(rg[f], ro[f]) <= 1e-12, (f, rel_linf(rg[f], ro[f]))
    g.close()
    o.close()


def test_rk4_flavor3_30_steps(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(10, 12, 2, ko=0.05, pmo=0.5)
    o = oracle_sim(O, c)
    g = N.NuOsc(**c, nflavor=3)
    st = full_state(o, 77)
    o.set_state(st)
    g.fillInitValue(ipt=5, alpha=0.9, eps0=0.3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=77)  # G0, n_nue0
    g.set_state(st)
    o.step(30)
    g.step_rk4(30)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(18):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, (f, rel_linf(gs[f], os_[f]))
    ag, (ao, _) = g.analysis(), o.analysis()
    assert abs(ag.surv - ao[2]) <= 1e-12 and abs(ag.survb - ao[3]) <= 1e-12
    g.close()
    o.close()


def test_srk3_flavor3(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(6, 9, 2, ko=0.05, pmo=0.5)
    o = oracle_sim(O, c)
    g = N.NuOsc(**c, nflavor=3)
    st = full_state(o, 5)
    o.set_state(st)
    g.fillInitValue(ipt=5, alpha=0.9, eps0=0.3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=5)
    g.set_state(st)
    o.step_srk3(12)
    g.step_srk3(12)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(18):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, (f, rel_linf(gs[f], os_[f]))
    g.close()
    o.close()


@pytest.mark.parametrize("x1", [0.1, 0.6])   # one x row (nx = 1) and a 6-row box (nx = 6)
def test_open_column_flavor3(oracle_mod, nuosc_mod, x1):
    O, N = oracle_mod, nuosc_mod
    c = dict(nv=2, x0=0.0, x1=x1, z0=-3.2, z1=3.2, dx=0.1, dz=0.1, CFL=0.4, ko=0.05, mu=1.0, pmo=0.3)
    o = O.OracleSim(nn=2, gx=2, gz=2, x0=0.0, x1=x1, z0=-3.2, z1=3.2, dx=0.1, dz=0.1, cfl=0.4, ko=0.05, mu=1.0,
                    pmo=0.3, bc=1, zc=-1.0, nflavor=3)
    g = N.NuOsc(**c, bc=N.BC_OPEN_Z, zc=-1.0, nflavor=3)
    init = dict(ipt=0, alpha=0.9, eps0=1e-2, sigma=0.5, lnue=0.6, lnueb=0.5)
    o.fill_init(**init)
    g.fillInitValue(**init)
    o.step(20)
    g.step_rk4(20)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(18):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, (f, rel_linf(gs[f], os_[f]))
    g.close()
    o.close()


@pytest.mark.parametrize("nx,nz,nn", [(5, 6, 2), (13, 11, 4)])
def test_calrhs3_fp32(oracle_mod, nuosc_mod, nx, nz, nn):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(nx, nz, nn)
    o = oracle_sim(O, c)
    a = N.NuOsc(**c, nflavor=3)
    b = N.NuOsc(**c, nflavor=3, precision=N.PREC_FP32)
    st = full_state(o, 3 + nx).astype(np.float32).astype(np.float64)
    r64, r32 = a.calRHS(st), b.calRHS(st)
    eps32 = np.finfo(np.float32).eps
    for f in range(18):
        bound = 64 * eps32 * np.max(np.abs(st)) / c["dx"]
        assert np.max(np.abs(r32[f] - r64[f])) <= bound, f
    a.close()
    b.close()
    o.close()


def test_init_zeroes_tau_fields(nuosc_mod):
    N = nuosc_mod
    c = cfg_for(4, 5, 2)
    g = N.NuOsc(**c, nflavor=3)
    g.set_state(np.ones((18, 4, 5, g.nv)))
    g.fillInitValue(ipt=5, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5, seed=2)
    g2 = N.NuOsc(**c)
    g2.fillInitValue(ipt=5, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5, seed=2)
    st = g.get_state()
    assert np.array_equal(st[:8], g2.get_state())
    assert not np.any(st[8:])
    g.close()
    g2.close()


def test_flavor3_arguments_checked(nuosc_mod):
    N = nuosc_mod
    c = cfg_for(8, 8)
    with pytest.raises(N.NuOscError):
        N.NuOsc(**c, nflavor=4)
    with pytest.raises(N.NuOscError):
        N.NuOsc(**c, nflavor=3, ko_order=5, gx=3, gz=3)
    with pytest.raises(N.NuOscError):
        N.NuOsc(**c, nflavor=3, adv=N.ADV_UPWIND2)
    g = N.NuOsc(**c, nflavor=3)
    with pytest.raises(N.NuOscError):
        g.get_field(18)
    b = N.Batch(precision=N.PREC_FP64)
    with pytest.raises(N.NuOscError):
        b.add(g)
    assert g.autotune(1)[0] == 0
    b.close()
    g.close()


def test_cli_nflavor3(tmp_path):
    exe = os.path.join(ROOT, "nuosc_amd", "bin", "nuosc")
    p = subprocess.run([exe, "--zmax", "0.8", "--dz", "0.1", "--nv", "2", "--ko", "0.05", "--END_STEP", "10",
                        "--ANA_EVERY", "5", "--nflavor", "3", "--ipt", "5", "--eps0", "0.1"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Fieldvar: 18" in p.stdout
    rows = [l for l in open(tmp_path / "analysis.dat").read().splitlines()[1:] if l.strip()]
    assert len(rows) == 3


def test_flavor3_graph_replay_bitwise(nuosc_mod):
    """hipGraph replay of the 3-flavour step (stage kernels + moment totals)
    equals eager launches bit for bit, over several replays."""
    N = nuosc_mod
    c = cfg_for(9, 10, 2, ko=0.05, pmo=0.5)
    init = dict(ipt=5, alpha=0.9, eps0=0.2, sigma=1.0, lnue=0.6, lnueb=0.5, seed=21)
    a = N.NuOsc(**c, nflavor=3, use_graph=True)
    b = N.NuOsc(**c, nflavor=3, use_graph=False)
    for g in (a, b):
        g.fillInitValue(**init)
    for _ in range(3):
        a.step_rk4(4)
        b.step_rk4(4)
        assert np.array_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


def test_flavor3_renorm(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    c = cfg_for(6, 8, 2, ko=0.05, pmo=0.5)
    o = oracle_sim(O, c, renorm=1)
    g = N.NuOsc(**c, nflavor=3, renorm=True)
    init = dict(ipt=5, alpha=0.9, eps0=0.2, sigma=1.0, lnue=0.6, lnueb=0.5, seed=4)
    o.fill_init(**init)
    g.fillInitValue(**init)
    o.step(10)
    g.step_rk4(10)
    gs, os_ = g.get_state(), o.get_state()
    for f in range(18):
        assert rel_linf(gs[f], os_[f]) <= 1e-10, (f, rel_linf(gs[f], os_[f]))
    g.close()
    o.close()


def test_flavor3_profile_stages(nuosc_mod):
    N = nuosc_mod
    g = N.NuOsc(**cfg_for(8, 16, 2), nflavor=3)
    g.fillInitValue(ipt=5, alpha=0.9, eps0=0.1, sigma=1.0, lnue=0.6, lnueb=0.5, seed=1)
    ms = g.profile_stages(2)
    assert len(ms) == 4 and all(x > 0 for x in ms)
    g.close()


def test_server_refuses_flavor3(nuosc_mod):
    N = nuosc_mod
    srv = N.Server(precision=N.PREC_FP64, hbm_budget_bytes=1 << 30)
    with pytest.raises(N.NuOscError):
        srv.submit(10, nflavor=3)
    srv.close()


def test_flavor3_fp32_tracks_fp64(nuosc_mod):
    """50 RK4 steps of the FP32 form against FP64 from the same seeded
    state.  Bound: every stage store rounds each field to FP32 (relative
    u = 2^-24), 4 stores per step; if those roundings never cancel and the
    dynamics amplify an error at most twofold over the run (t = 2), the
    FP32 state is within 4 * 50 * u * 2 = 2.4e-5 of max|field|, and so is
    Pee = sum w ee / sum w G0, a weighted mean of ee/G0.  (Measured: fields
    2.8e-6, Pee 1.7e-6.)  The 3-flavour vacuum terms as written do not
    conserve Tr rho, so Pee drifts above 1 here in both precisions."""
    N = nuosc_mod
    c = cfg_for(12, 16, 4, ko=0.05, pmo=0.5)
    init = dict(ipt=5, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=8)
    a = N.NuOsc(**c, nflavor=3)
    b = N.NuOsc(**c, nflavor=3, precision=N.PREC_FP32)
    for g in (a, b):
        g.fillInitValue(**init)
        g.step_rk4(50)
    sa, sb = a.get_state(), b.get_state()
    bound = 4 * 50 * 2.0 ** -24 * 2
    assert np.max(np.abs(sa - sb)) <= bound * np.max(np.abs(sa))
    assert np.any(sa[8:])                        # the vacuum term populated the tau fields
    assert abs(a.analysis().surv - b.analysis().surv) <= bound
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["open_rk4", "periodic_srk3"])
def test_flavor3_fp32_tracks_fp64_open_srk3(nuosc_mod, case):
    """The FP32 z march on open z (6 x rows, 20 RK4 steps) and with SSP-RK3
    (12 x 16, 20 steps) against FP64 from the same state.  Bound as in
    test_flavor3_fp32_tracks_fp64: s FP32 stage stores per step (RK4 4,
    SSP-RK3 3), each rounding every field by at most 2^-24 relative, never
    cancelling, amplified at most twofold: s * 20 * 2^-24 * 2 of max|field|."""
    N = nuosc_mod
    if case == "open_rk4":
        c = dict(nv=2, x0=0.0, x1=0.6, z0=-3.2, z1=3.2, dx=0.1, dz=0.1, CFL=0.4, ko=0.05, mu=1.0, pmo=0.3,
                 bc=N.BC_OPEN_Z, zc=-1.0)
        init = dict(ipt=0, alpha=0.9, eps0=1e-2, sigma=0.5, lnue=0.6, lnueb=0.5)
        stores = 4
    else:
        c = cfg_for(12, 16, 4, ko=0.05, pmo=0.5)
        init = dict(ipt=5, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=11)
        stores = 3
    a = N.NuOsc(**c, nflavor=3)
    b = N.NuOsc(**c, nflavor=3, precision=N.PREC_FP32)
    assert a.shape[0] > 1                        # a box: the z-march kernel
    for g in (a, b):
        g.fillInitValue(**init)
        if case == "open_rk4":
            g.step_rk4(20)
        else:
            g.step_srk3(20)
    sa, sb = a.get_state(), b.get_state()
    bound = stores * 20 * 2.0 ** -24 * 2
    assert np.all(np.isfinite(sb))
    assert np.max(np.abs(sa - sb)) <= bound * np.max(np.abs(sa)), np.max(np.abs(sa - sb)) / np.max(np.abs(sa))
    a.close()
    b.close()
