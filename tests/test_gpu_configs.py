"""The benchmark workloads, each exercised on the GPU at its own shape (run
lengths reduced where stated).

512 x 512 x 642 periodic box, FP64 (--zmax 12.8 --dz 0.05 --nv 8, mu 1,
ko 0.1, cfl 0.4, pmo 0, seeded eps = 1e-6), 10 RK4 steps:
  * per direction, the box sums of ee + mm and of bee + bmm are conserved to
    1e-13 relative: the interaction is trace-free (Imm = -Iee) and the
    central advection and Kreiss-Oliger differences telescope over the
    periodic box, so only rounding changes them;
  * max |P| - 1 <= 1e-11: advection mixes neighbouring points whose
    polarisation vectors differ in direction by O(eps) (the seeded
    perturbation varies point to point), which shortens them at O(eps^2) =
    1e-12 (the oracle's 128 x 128 crop of the same physics reads 1e-12 after
    one step and 3e-12 after ten);
  * Pee and |M0| finite, the state free of non-finite values.
128 x 128 x 642 crop of the same physics (--zmax 3.2 --dz 0.05 --nv 8)
against the oracle, 10 steps: relative L-inf <= 1e-10 per field, Pee and
|M0| relative 1e-10, |P| - 1 absolute 1e-12 + relative 1e-10.

FP32 form at 512 x 512 x 642 against FP64 from the same seeded state, 50
steps.  Each of the 200 stage stores rounds a value to FP32 (relative
u = 2^-24 ~ 6e-8), and over 50 steps (t = 1) the linearised dynamics amplify
a perturbation by at most e^(gamma t) with gamma < 1, so the accumulated
error is at most 200 u e ~ 3.2e-5 of the field's scale: relative L-inf per
field <= 3.2e-5, Pee within 1e-6 and max |P| - 1 below 1e-5 (the bounds of
the FP32 growth-rate test).

1-D open column (1 x 4096 x 642, gz = 3, ipt-0 pulse entering through z0)
in FP32 against FP64, 1000 steps: growth rate of |rho_ex| within 1% of
FP64's, Pee within 1e-6, max |P| - 1 below 1e-5.

64-run batched scan (FP32; grids 64^2..256^2, 162/642 directions, uneven
physics, own seeds: the draw the batched benchmark uses), 5 steps: members
0, 1, 5 and the one with the largest grid bitwise equal to their runs alone; an
FP64 batch of the first 8 runs of the draw, 5 steps, its 64 x 64 x 162
member against the oracle at relative L-inf <= 1e-10 per field.

Served scan: requests drawn as the served benchmark draws them (grids up to
384^2, 162/642 directions; steps reduced to U[20, 60]), Poisson arrivals at
2/s into the run server (FP32): every request's series equals its run alone,
bit for bit.
"""
import os
import sys
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workload draws of the benchmark)
import growth as G  # noqa: E402

pytestmark = pytest.mark.gpu

BOX512 = dict(nv=8, x0=-12.8, x1=12.8, z0=-12.8, z1=12.8, dx=0.05, dz=0.05, CFL=0.4, ko=0.1, mu=1.0, pmo=0.0)
CROP128 = dict(BOX512, x0=-3.2, x1=3.2, z0=-3.2, z1=3.2)
SEEDED = dict(ipt=5, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5, seed=12345)


def rel_linf(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


def oracle_sim(O, c):
    return O.OracleSim(nn=c["nv"], x0=c["x0"], x1=c["x1"], z0=c["z0"], z1=c["z1"], dx=c["dx"], dz=c["dz"],
                       cfl=c["CFL"], ko=c["ko"], mu=c["mu"], pmo=c["pmo"])


def box_sum(f):
    """Per direction, the sum over the (x, z) box in extended precision, z
    first then x (512 terms each): a plain sum of 262144 doubles in a row
    carries ~1e-12 relative rounding, far above the drift measured here."""
    return f.sum(axis=1, dtype=np.longdouble).sum(axis=0)


def trace_sums(g):
    """Per direction: box sums of ee + mm and of bee + bmm."""
    out = []
    for a, b in ((0, 1), (4, 5)):
        s = box_sum(g.get_field(a))
        s += box_sum(g.get_field(b))
        out.append(s)
    return out


def test_box512_fp64_conservation(nuosc_mod):
    N = nuosc_mod
    g = N.NuOsc(**BOX512)
    assert g.shape == (512, 512, 642)
    g.fillInitValue(**SEEDED)
    t0 = trace_sums(g)
    g.step_rk4(10)
    a = g.analysis()
    t1 = trace_sums(g)
    drift = [float(np.max(np.abs(y - x) / np.abs(x))) for x, y in zip(t0, t1)]
    print(f"512x512x642 FP64, 10 steps: trace drift nu {drift[0]:.2e} nubar {drift[1]:.2e}, "
          f"max|P|-1 {a.maxdP:.2e}, Pee {a.surv:.15f}, |M0| {a.aM0:.6e}", flush=True)
    assert max(drift) <= 1e-13, drift
    assert a.maxdP <= 1e-11
    assert a.status == 0 and a.nonfinite == 0
    assert np.isfinite(a.surv) and np.isfinite(a.aM0) and np.isfinite(a.survb)
    assert abs(g.phy_time - 10 * g.dt) <= 1e-14
    g.close()


def test_crop128_fp64_against_oracle(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    o = oracle_sim(O, CROP128)
    g = N.NuOsc(**CROP128)
    assert g.shape == o.shape == (128, 128, 642)
    o.fill_init(**SEEDED)
    g.fillInitValue(**SEEDED)
    for step in range(10):
        o.step(1)
        g.step_rk4(1)
        ao, _ = o.analysis()
        ag = g.analysis()
        assert abs(ag.surv - ao[2]) <= 1e-10 * abs(ao[2])
        assert abs(ag.aM0 - ao[6]) <= 1e-10 * abs(ao[6])
        assert abs(ag.maxdP - ao[1]) <= 1e-12 + 1e-10 * abs(ao[1])
    gs, os_ = g.get_state(), o.get_state()
    errs = [rel_linf(gs[f], os_[f]) for f in range(8)]
    print(f"128x128x642 crop, 10 steps: max rel L-inf vs oracle {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-10, errs
    g.close()
    o.close()


def test_box512_fp32_against_fp64(nuosc_mod):
    N = nuosc_mod
    sims = {}
    for prec in (64, 32):
        g = N.NuOsc(precision=N.PREC_FP64 if prec == 64 else N.PREC_FP32, **BOX512)
        g.fillInitValue(**SEEDED)
        sims[prec] = g
    pee, dp = [], []
    for blk in range(5):
        for g in sims.values():
            g.step_rk4(10)
        a64, a32 = sims[64].analysis(), sims[32].analysis()
        pee.append(abs(a32.surv - a64.surv))
        dp.append(a32.maxdP)
        assert a32.status == 0 and a32.nonfinite == 0
    errs = []
    for f in range(8):
        errs.append(rel_linf(sims[32].get_field(f), sims[64].get_field(f)))
    print(f"512x512x642 FP32 vs FP64, 50 steps: rel L-inf per field {['%.1e' % e for e in errs]}, "
          f"Pee err {max(pee):.2e}, FP32 max|P|-1 {max(dp):.2e}", flush=True)
    assert max(errs) <= 3.2e-5, errs
    assert max(pee) <= 1e-6
    assert max(dp) <= 1e-5
    for g in sims.values():
        g.close()


def test_column4096_fp32_against_fp64(nuosc_mod):
    N = nuosc_mod
    col = dict(nv=8, x0=-0.025, x1=0.025, z0=-102.4, z1=102.4, dx=0.05, dz=0.05, CFL=0.4, ko=0.1, mu=1.0,
               pmo=0.0, gz=3, bc=N.BC_OPEN_Z, zc=-105.4)
    init = dict(ipt=0, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5)
    series = {}
    for prec in (64, 32):
        g = N.NuOsc(precision=N.PREC_FP64 if prec == 64 else N.PREC_FP32, **col)
        assert g.shape == (1, 4096, 642)
        g.fillInitValue(**init)
        rows = []
        for blk in range(101):
            if blk:
                g.step_rk4(10)
            a = g.analysis()
            rows.append((a.phy_time, a.rho_ex2, a.surv, a.maxdP))
        series[prec] = np.array(rows)
        g.close()
    s64, s32 = series[64], series[32]
    # the pulse enters through z0 and its |rho_ex| grows: fit over the run
    g64 = G.fitted_rate(s64[:, 0], s64[:, 1])
    g32 = G.fitted_rate(s32[:, 0], s32[:, 1])
    pee = np.max(np.abs(s32[:, 2] - s64[:, 2]))
    print(f"column 4096x642 FP32 vs FP64, 1000 steps: growth {g64:.6f} vs {g32:.6f} "
          f"(rel {abs(g32 - g64) / abs(g64):.1e}), Pee err {pee:.1e}, FP32 max|P|-1 {np.max(s32[:, 3]):.1e}",
          flush=True)
    assert g64 > 0
    assert abs(g32 - g64) <= 1e-2 * abs(g64)
    assert pee <= 1e-6
    assert np.max(s32[:, 3]) <= 1e-5


def test_batched_scan64(oracle_mod, nuosc_mod):
    O, N = oracle_mod, nuosc_mod
    rng = np.random.default_rng(7)
    specs = [bench.scan_run(rng, k) for k in range(64)]
    sims = []
    for geo, init in specs:
        g = N.NuOsc(precision=N.PREC_FP32, use_graph=True, **geo)
        g.fillInitValue(**init)
        sims.append(g)
    sizes = [g.nx * g.nz * g.nv for g in sims]
    picks = sorted({int(np.argmax(sizes)), 0, 1, 5})
    alone = []
    for k in picks:
        geo, init = specs[k]
        a = N.NuOsc(precision=N.PREC_FP32, use_graph=True, **geo)
        a.fillInitValue(**init)
        alone.append(a)
    b = N.Batch(precision=N.PREC_FP32)
    for g in sims:
        b.add(g)
    assert len(b) == 64
    b.step(5)
    for a in alone:
        a.step_rk4(5)
    for k, a in zip(picks, alone):
        assert np.array_equal(a.get_state(), sims[k].get_state()), k
        a.close()
    assert all(g.analysis().status == 0 for g in sims)
    print(f"64-run FP32 batch: {sum(sizes)} updates per step, members {picks} equal alone", flush=True)
    b.close()
    for g in sims:
        g.close()
    # FP64 batch of the first 8 runs of the draw; member 1 is 64 x 64 x 162
    geo1, init1 = specs[1]
    sims = []
    for geo, init in specs[:8]:
        g = N.NuOsc(precision=N.PREC_FP64, use_graph=True, **geo)
        g.fillInitValue(**init)
        sims.append(g)
    assert sims[1].shape == (64, 64, 162)
    b = N.Batch(precision=N.PREC_FP64)
    for g in sims:
        b.add(g)
    b.step(5)
    o = oracle_sim(O, geo1)
    o.fill_init(**init1)
    o.step(5)
    gs, os_ = sims[1].get_state(), o.get_state()
    errs = [rel_linf(gs[f], os_[f]) for f in range(8)]
    print(f"FP64 batch member 64x64x162 vs oracle, 5 steps: max rel L-inf {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-10, errs
    o.close()
    b.close()
    for g in sims:
        g.close()


def alone_series(N, geo, init, nsteps, ana, precision):
    g = N.NuOsc(precision=precision, use_graph=False, **geo)
    g.fillInitValue(**init)
    rows = []
    a = g.analysis()
    rows.append((a.phy_time, a.surv, a.survb, a.maxdP, a.aM0))
    done = 0
    while done < nsteps:
        k = min(ana, nsteps - done)
        g.step_rk4(k)
        done += k
        a = g.analysis()
        rows.append((a.phy_time, a.surv, a.survb, a.maxdP, a.aM0))
    g.close()
    return np.array(rows)


def test_served_scan_draw(nuosc_mod):
    N = nuosc_mod
    rng = np.random.default_rng(11)
    grids = (64, 128, 192, 256, 320, 384)
    srv = N.Server(device=0, precision=N.PREC_FP32, hbm_budget_bytes=200 << 30, max_tick_steps=10)
    jobs = {}
    t_next = 0.0
    t0 = time.perf_counter()
    for k in range(12):
        geo, init = bench.scan_run(rng, k, grids=grids)
        nsteps = int(rng.integers(20, 61))
        ana = max(1, nsteps // 5)
        t_next += rng.exponential(0.5)
        while time.perf_counter() - t0 < t_next:
            time.sleep(0.01)
        rid = srv.submit(nsteps=nsteps, ana_every=ana, **geo, **init)
        jobs[rid] = (geo, init, nsteps, ana)
    got = {}
    for _ in jobs:
        info, series = srv.take(timeout=300)
        assert info.status == 0
        got[info.id] = (info, series)
    srv.close()
    assert set(got) == set(jobs)
    for rid, (geo, init, nsteps, ana) in jobs.items():
        info, series = got[rid]
        assert info.nsteps == nsteps
        want = alone_series(N, geo, init, nsteps, ana, N.PREC_FP32)
        assert np.array_equal(series, want), rid
    n = max(int(round((g[0]["x1"] - g[0]["x0"]) / g[0]["dx"])) for g in jobs.values())
    print(f"served: {len(jobs)} requests (grids up to {n}^2) equal alone", flush=True)
