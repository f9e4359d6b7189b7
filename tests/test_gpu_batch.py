"""Batched runs: members with different grids, direction counts and physics
advance together (one launch per RK4 stage for all, as an ordinary grid or a
persistent work-queue grid) and every member's state is bitwise the one it
reaches stepped alone; members can leave and join."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RUNS = [
    dict(nv=2, x0=-0.8, x1=0.8, z0=-1.2, z1=1.2, dx=0.1, dz=0.1, CFL=0.4, ko=0.1, mu=1.0, pmo=0.0),
    dict(nv=4, x0=-1.6, x1=1.6, z0=-0.8, z1=0.8, dx=0.1, dz=0.1, CFL=0.4, ko=0.05, mu=0.7, pmo=1.0),
    dict(nv=3, x0=-0.4, x1=0.4, z0=-0.4, z1=0.4, dx=0.05, dz=0.05, CFL=0.3, ko=0.1, mu=1.5, pmo=0.0),
    dict(nv=1, x0=0.0, x1=0.1, z0=-3.2, z1=3.2, dx=0.1, dz=0.1, CFL=0.4, ko=0.1, mu=1.0, pmo=0.0),
]


def make(N, cfg, seed, precision):
    g = N.NuOsc(precision=precision, **cfg)
    g.fillInitValue(ipt=5, alpha=0.85, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5, seed=seed)
    return g


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_equals_alone(nuosc_mod, precision):
    N = nuosc_mod
    alone = [make(N, c, 100 + k, precision) for k, c in enumerate(RUNS)]
    mixed = [make(N, c, 100 + k, precision) for k, c in enumerate(RUNS)]
    b = N.Batch(precision=precision)
    for g in mixed[:3]:
        b.add(g)
    assert len(b) == 3
    b.step(7)
    for g in alone[:3]:
        g.step_rk4(7)
    for a, m in zip(alone[:3], mixed[:3]):
        assert np.array_equal(a.get_state(), m.get_state())
        assert a.phy_time == m.phy_time
    # membership changes: one leaves, one joins
    b.remove(mixed[1])
    b.add(mixed[3])
    b.step(5)
    for k in (0, 2):
        alone[k].step_rk4(5)
    alone[3].step_rk4(5)
    for k in (0, 2, 3):
        assert np.array_equal(alone[k].get_state(), mixed[k].get_state()), k
        a1, a2 = alone[k].analysis(), mixed[k].analysis()
        assert a1.surv == a2.surv and a1.maxdP == a2.maxdP
    # the member that left keeps working alone
    mixed[1].step_rk4(5)
    alone[1].step_rk4(5)
    assert np.array_equal(alone[1].get_state(), mixed[1].get_state())
    b.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_persistent_launch(nuosc_mod, precision):
    """The persistent launch (resident blocks drawing logical blocks from
    per-XCD work queues) gives every member bitwise the state it reaches
    stepped alone, over replays of the captured step (the queue counters are
    zeroed by a kernel node in each); the autotuner's (zlen, launch form)
    choice keeps that, and an unknown form is refused."""
    N = nuosc_mod
    alone = [make(N, c, 200 + k, precision) for k, c in enumerate(RUNS)]
    mixed = [make(N, c, 200 + k, precision) for k, c in enumerate(RUNS)]
    b = N.Batch(precision=precision)
    for g in mixed:
        b.add(g)
    assert b.launch == N.BATCH_GRID
    b.set_launch(N.BATCH_PERSISTENT)
    assert b.launch == N.BATCH_PERSISTENT
    b.step(6)
    for g in alone:
        g.step_rk4(6)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a.get_state(), m.get_state()), k
    zl, ms = b.autotune(trials=1)
    assert zl in (16, 32, 64, 128, 256) and ms > 0
    assert b.launch in (N.BATCH_GRID, N.BATCH_PERSISTENT)
    b.step(3)
    for g in alone:
        g.step_rk4(3)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a.get_state(), m.get_state()), k
    with pytest.raises(N.NuOscError):
        b.set_launch(2)
    b.close()


def test_batch_rejects_mismatch(nuosc_mod):
    N = nuosc_mod
    b = N.Batch(precision=64)
    g32 = make(N, RUNS[0], 1, 32)
    with pytest.raises(N.NuOscError):
        b.add(g32)
    g = make(N, RUNS[0], 1, 64)
    b.add(g)
    with pytest.raises(N.NuOscError):
        b.add(g)


def test_estimate_matches_allocation(nuosc_mod):
    N = nuosc_mod
    for c in RUNS:
        for p in (64, 32):
            g = N.NuOsc(precision=p, **c)
            g.fillInitValue(ipt=0, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5)
            est = N.estimate_device_bytes(precision=p, **c)
            held = g.device_bytes()
            assert est >= held * 0.95 and est <= held * 1.5 + (1 << 20), (est, held)
            g.close()
            g3 = N.NuOsc(precision=p, nflavor=3, **c)
            est3 = N.estimate_device_bytes(precision=p, nflavor=3, **c)
            held3 = g3.device_bytes()
            assert est3 >= held3 * 0.95 and est3 <= held3 * 1.5 + (1 << 20), (est3, held3)
            g3.close()


def test_autotune_keeps_state_and_results(nuosc_mod):
    """The z-segment length changes only the decomposition: autotuning leaves
    the state untouched and every zlen steps to the same bits."""
    N = nuosc_mod
    cfg = RUNS[1]
    ref = make(N, cfg, 5, 64)
    ref.step_rk4(6)
    want = ref.get_state()
    for zl in (8, 16, 3, 1000):
        g = make(N, cfg, 5, 64)
        before = g.get_state()
        g.autotune(trials=1)
        assert np.array_equal(g.get_state(), before)
        g.set_zlen(zl)
        g.step_rk4(6)
        assert np.array_equal(g.get_state(), want), zl


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_zlen_and_autotune(nuosc_mod, precision):
    """The batch's shared z-segment length (and its autotuner) changes only
    the decomposition: states stay untouched by tuning and every length steps
    every member to the bits it reaches alone; an open-boundary run is
    refused."""
    N = nuosc_mod
    alone = [make(N, c, 200 + k, precision) for k, c in enumerate(RUNS)]
    for g in alone:
        g.step_rk4(5)
    for zl in (0, 4, 64, None):
        mixed = [make(N, c, 200 + k, precision) for k, c in enumerate(RUNS)]
        b = N.Batch(precision=precision)
        for g in mixed:
            b.add(g)
        if zl is None:
            before = [g.get_state() for g in mixed]
            best, ms = b.autotune(trials=1)
            assert best in (16, 32, 64, 128, 256) and ms > 0
            for g, s in zip(mixed, before):
                assert np.array_equal(g.get_state(), s)
        else:
            b.set_zlen(zl)
        b.step(5)
        for a, m in zip(alone, mixed):
            assert np.array_equal(a.get_state(), m.get_state()), zl
        b.close()
    op = N.NuOsc(precision=precision, bc=N.BC_OPEN_Z, zc=-8.0, **RUNS[3])
    op.fillInitValue(ipt=0, alpha=0.9, eps0=1e-3, sigma=1.0, lnue=0.6, lnueb=0.5)
    b = N.Batch(precision=precision)
    with pytest.raises(N.NuOscError):
        b.add(op)


@pytest.mark.parametrize("precision", [64, 32])
def test_member_physics_change_reaches_batch(nuosc_mod, precision):
    """set_mu / set_pmo on a member after it joined a batch take effect at the
    next batch step: the member stays bitwise equal to the same run alone."""
    N = nuosc_mod
    alone = [make(N, c, 300 + k, precision) for k, c in enumerate(RUNS[:2])]
    mixed = [make(N, c, 300 + k, precision) for k, c in enumerate(RUNS[:2])]
    b = N.Batch(precision=precision)
    for g in mixed:
        b.add(g)
    b.step(3)
    for g in alone:
        g.step_rk4(3)
    for g in (alone[0], mixed[0]):
        g.set_mu(1.7)
    for g in (alone[1], mixed[1]):
        g.set_pmo(-1.0)
    b.step(4)
    for g in alone:
        g.step_rk4(4)
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.get_state(), m.get_state())
    b.close()


def test_destroyed_member_leaves_its_batch(nuosc_mod):
    """Closing a member removes it from its batch (the batch must not keep
    pointers to freed buffers); the others keep stepping as they would alone.
    The batch is stepped only after its size shows the member gone."""
    N = nuosc_mod
    alone = [make(N, c, 400 + k, 64) for k, c in enumerate(RUNS[:3])]
    mixed = [make(N, c, 400 + k, 64) for k, c in enumerate(RUNS[:3])]
    b = N.Batch(precision=64)
    for g in mixed:
        b.add(g)
    b.step(2)
    for g in alone:
        g.step_rk4(2)
    mixed[1].close()
    assert len(b) == 2
    b.remove(mixed[1])          # already gone: the binding only forgets it
    assert len(b.members) == 2
    b.step(3)
    for k in (0, 2):
        alone[k].step_rk4(3)
        assert np.array_equal(alone[k].get_state(), mixed[k].get_state()), k
    b.close()


def test_member_of_one_batch_only(nuosc_mod):
    N = nuosc_mod
    g = make(N, RUNS[0], 1, 64)
    b1, b2 = N.Batch(precision=64), N.Batch(precision=64)
    b1.add(g)
    with pytest.raises(N.NuOscError):
        b2.add(g)
    b1.remove(g)
    b2.add(g)
    b1.close()
    b2.close()
    # a destroyed batch releases its members
    b3 = N.Batch(precision=64)
    b3.add(g)
    g.close()
