"""Run server: concurrent submissions are admitted within the HBM budget,
advanced together, and each run's series equals the series of the same run
stepped alone (bitwise: the batched kernels compute every run exactly as the
single-run path)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(nv=2, x0=-0.8, x1=0.8, z0=-1.2, z1=1.2, dx=0.1, dz=0.1, mu=1.0),
    dict(nv=3, x0=-0.4, x1=0.4, z0=-1.6, z1=1.6, dx=0.1, dz=0.1, mu=0.7),
    dict(nv=4, x0=-0.4, x1=0.4, z0=-0.4, z1=0.4, dx=0.05, dz=0.05, mu=1.3),
]


def alone_series(N, shape, nsteps, ana, seed, precision):
    g = N.NuOsc(precision=precision, CFL=0.4, ko=0.1, pmo=0.0, use_graph=False,
                **{k: v for k, v in shape.items()})
    g.fillInitValue(ipt=5, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=seed)
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
    return np.array(rows)


@pytest.mark.parametrize("precision", [64, 32])
def test_server_results_equal_alone(nuosc_mod, precision):
    N = nuosc_mod
    jobs = [(SHAPES[k % 3], 20 + 7 * k, 5 + (k % 2) * 3, 1000 + k) for k in range(6)]
    srv = N.Server(precision=precision, hbm_budget_bytes=1 << 34, max_tick_steps=4)
    ids = {}
    for shape, nsteps, ana, seed in jobs:
        ids[srv.submit(nsteps=nsteps, ana_every=ana, seed=seed, eps0=1e-3, CFL=0.4, ko=0.1, pmo=0.0, **shape)] = \
            (shape, nsteps, ana, seed)
        time.sleep(0.01)
    got = {}
    for _ in jobs:
        info, series = srv.take(timeout=120)
        assert info.status == 0
        got[info.id] = (info, series)
    assert set(got) == set(ids)
    for rid, (shape, nsteps, ana, seed) in ids.items():
        info, series = got[rid]
        assert info.nsteps == nsteps
        assert info.t_submit <= info.t_admit <= info.t_done
        want = alone_series(N, shape, nsteps, ana, seed, precision)
        assert series.shape == want.shape
        assert np.array_equal(series, want), rid
    st = srv.stats()
    assert st.finished == len(jobs) and st.active == 0 and st.queued == 0
    srv.close()


def test_server_respects_budget(nuosc_mod):
    N = nuosc_mod
    shape = SHAPES[0]
    one = N.estimate_device_bytes(precision=64, **{k: v for k, v in shape.items() if k != "mu"})
    srv = N.Server(precision=64, hbm_budget_bytes=int(2.5 * one), max_tick_steps=5)
    for k in range(5):
        srv.submit(nsteps=30, ana_every=10, seed=k, **shape)
    peak = 0
    for _ in range(400):
        st = srv.stats()
        peak = max(peak, st.resident_bytes)
        assert st.active <= 2
        if st.finished == 5:
            break
        time.sleep(0.01)
    assert peak <= 2.5 * one
    results = [srv.take(timeout=60) for _ in range(5)]
    assert all(r is not None and r[0].status == 0 for r in results)
    # later runs waited for a slot
    admits = sorted(r[0].t_admit for r in results)
    assert admits[-1] > admits[0]
    srv.close()


def test_server_refusals(nuosc_mod):
    """submit refuses runs a batch cannot take (renormalising, open z, a
    non-default discretisation, 3 flavours) and runs larger than the budget."""
    N = nuosc_mod
    shape = SHAPES[0]
    one = N.estimate_device_bytes(precision=64, **{k: v for k, v in shape.items() if k != "mu"})
    srv = N.Server(precision=64, hbm_budget_bytes=int(1.5 * one), max_tick_steps=5)
    req = N.Request(cfg=N.Config(nn=shape["nv"], gx=2, gz=2, x0=shape["x0"], x1=shape["x1"], z0=shape["z0"],
                                 z1=shape["z1"], dx=shape["dx"], dz=shape["dz"], cfl=0.4, ko=0.1, mu=1.0,
                                 renorm=1),
                    ipt=5, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5, lnuex=N.DBL_MAX,
                    lnuebx=N.DBL_MAX, seed=1, nsteps=5, ana_every=5)
    import ctypes as C
    rid = C.c_longlong()
    assert N.lib().nuosc_server_submit(srv._h, C.byref(req), C.byref(rid)) == N.NUOSC_EINVAL
    req.cfg.renorm = 0
    for field, val in (("bc", N.BC_OPEN_Z), ("ko_order", 5), ("adv", N.ADV_MUSCL), ("nflavor", 3)):
        setattr(req.cfg, field, val)
        assert N.lib().nuosc_server_submit(srv._h, C.byref(req), C.byref(rid)) == N.NUOSC_EINVAL, field
        setattr(req.cfg, field, 0)
    # twice the grid: larger than the budget on its own
    req.cfg.x1 = shape["x1"] + (shape["x1"] - shape["x0"])
    assert N.lib().nuosc_server_submit(srv._h, C.byref(req), C.byref(rid)) == N.NUOSC_EINVAL
    req.cfg.x1 = shape["x1"]
    assert N.lib().nuosc_server_submit(srv._h, C.byref(req), C.byref(rid)) == N.NUOSC_OK
    info, _ = srv.take(timeout=60)
    assert info.status == 0
    srv.close()


def test_server_admits_small_runs_past_a_large_one(nuosc_mod):
    """A queued run that does not fit does not hold back younger runs that
    do (within the age bound); with age bound 0 the order is strict."""
    N = nuosc_mod
    small = SHAPES[0]
    big = dict(SHAPES[0], x0=-1.6, x1=1.6)
    b_small = N.estimate_device_bytes(precision=64, **{k: v for k, v in small.items() if k != "mu"})
    b_big = N.estimate_device_bytes(precision=64, **{k: v for k, v in big.items() if k != "mu"})
    assert b_big > b_small
    for bound, small_first in ((60.0, True), (0.0, False)):
        srv = N.Server(precision=64, hbm_budget_bytes=b_big + b_small // 2, max_tick_steps=5)
        srv.set_age_bound(bound)
        # a long-running small run occupies memory, so the big one cannot fit;
        # a second small run arrives after the big one
        r0 = srv.submit(nsteps=20000, ana_every=20000, seed=1, **small)
        time.sleep(0.2)
        r_big = srv.submit(nsteps=5, ana_every=5, seed=2, **big)
        r_small = srv.submit(nsteps=5, ana_every=5, seed=3, **small)
        res = {}
        for _ in range(3):
            info, _ = srv.take(timeout=120)
            assert info.status == 0
            res[info.id] = info
        srv.close()
        assert (res[r_small].t_admit < res[r_big].t_admit) == small_first, bound
        assert res[r_big].t_admit >= res[r0].t_done - 1e-9
